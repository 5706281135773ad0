// Device BVH builder (linear BVH, collapsed into four-wide nodes) and the device-side
// eye-to-scene distance used when the tree never exists on the host.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/goodphotons.h"

namespace gp {

struct DeviceBvh {
    float4* nodes = nullptr;   // nPairs four-wide nodes x 8 float4 (node i = pair node i collapsed)
    float4* tris = nullptr;    // 3 float4 per triangle, leaf order
    float4* trin = nullptr;    // normal + material per triangle, leaf order
    int64_t nPairs = 0;
    int depth = 0;             // pair nodes on the longest root-to-leaf path
    int stackNeed = 0;         // traversal stack bound: 3 per four-wide level
    double buildMs = 0.0;
};

// dTrisIn / dTrinIn: input-order triangles (3 float4, .w = prim id / material) and
// normals; centerLo/Hi bound the triangles' box centres. Synchronous.
gp_status buildBvhDevice(const float4* dTrisIn, const float4* dTrinIn, int64_t n, const float centerLo[3],
                         const float centerHi[3], DeviceBvh& out);

// Smallest squared distance from p to any of n triangles (3 float4 each), in double.
gp_status deviceMinDistance2(const float4* dTris, int64_t n, const double p[3], double& d2);

}  // namespace gp
