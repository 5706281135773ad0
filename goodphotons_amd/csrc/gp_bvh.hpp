// Host BVH builder: binned SAH (all three axes, 32 bins) down to leaves of at
// most 4 triangles, emitted as 64-byte child-pair nodes and collapsed into
// four-wide nodes for the GPU traversal.
// The reference builds its own tree (src/bvh.h:89-176, 16 bins on the widest
// axis); closest-hit results do not depend on the tree, only speed does.
#pragma once
#include <cstdint>
#include <vector>

#include "goodphotons.h"

namespace gp {

struct BvhResult {
    std::vector<float> pairs;   // 16 floats per child-pair node
    std::vector<int> order;     // leaf order -> primitive index
    int depth = 0;              // pair nodes on the longest root-to-leaf path
    int64_t leaves = 0;
    double sahCost = 0.0;
};

// Four-wide nodes for the GPU traversal from the child-pair tree: the root pair node
// and every inner grandchild reached from it become nodes whose slots are their
// grandchildren (a child that is a leaf keeps a slot of its own). 32 floats (128 bytes,
// one cache line) per node: lo.x[4], hi.x[4], lo.y[4], hi.y[4], lo.z[4], hi.z[4],
// link[4] (int bits: >= 0 a node, < 0 a leaf code), 0[4]. An unused slot holds a point
// box far outside any scene, linked to slot 0's child, so even a ray that hit it would
// only revisit a valid subtree. depth4 = nodes on the longest root-to-leaf path.
void collapse4(const std::vector<float>& pairs, std::vector<float>& quads, int& depth4);

// verts: n*9 floats. Throws std::runtime_error on inputs it cannot index.
void buildBvh(const float* verts, int64_t n, int threads, BvhResult& out);

// The caller's binary tree (the reference's Bvh::nodes / primIdx) in the same
// child-pair form: boxes rounded outwards to float, leaves of at most 4 primitives.
// Checked as it is read: every node reached once from node 0, child boxes inside
// their parent's, every leaf's triangles inside its box, every triangle in some
// leaf. Throws std::runtime_error (with the reason) on a tree that fails a check.
void importBvh(const float* verts, int64_t nTris, const gp_bvh_node* nodes, int64_t nNodes,
               const int32_t* primIndex, int64_t nIndex, BvhResult& out);

}  // namespace gp
