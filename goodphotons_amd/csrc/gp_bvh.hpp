// Host BVH builder: binned SAH (all three axes, 32 bins) down to leaves of at
// most 4 triangles, emitted as 64-byte child-pair nodes for the GPU traversal.
// The reference builds its own tree (src/bvh.h:89-176, 16 bins on the widest
// axis); closest-hit results do not depend on the tree, only speed does.
#pragma once
#include <cstdint>
#include <vector>

namespace gp {

struct BvhResult {
    std::vector<float> pairs;   // 16 floats per child-pair node
    std::vector<int> order;     // leaf order -> primitive index
    int depth = 0;              // pair nodes on the longest root-to-leaf path
    int64_t leaves = 0;
    double sahCost = 0.0;
};

// verts: n*9 floats. Throws std::runtime_error on inputs it cannot index.
void buildBvh(const float* verts, int64_t n, int threads, BvhResult& out);

}  // namespace gp
