// Device BVH builder: a linear BVH (Karras, "Maximizing parallelism in the
// construction of BVHs, octrees, and k-d trees", HPG 2012) emitted in the child-pair
// layout and collapsed into the four-wide nodes the traversal reads.
//
//   1. Morton code (21 bits per axis) of each triangle's box centre in the scene's
//      centre bounds; sort (code, index) pairs (rocPRIM radix sort).
//   2. One thread per internal node i finds the sorted range it covers and its split
//      from the longest common prefixes of neighbouring codes (equal codes are told
//      apart by their positions, so the tree is a valid binary tree for any input).
//   3. Boxes and heights bottom-up: one thread per primitive climbs from its leaf; at
//      each node the first arriving child stops, the second computes the union.
//   4. Every internal node whose range holds more than kLeafMax primitives becomes a
//      pair node; a child with at most kLeafMax primitives becomes a leaf over its
//      contiguous sorted range. Triangles are gathered into sorted (leaf) order.
//   5. Pair node i becomes four-wide node i, whose slots are its grandchildren (a
//      leaf child keeps its own slot), so links keep their indices; the root is
//      node 0, and nodes at odd pair levels are built but never reached.
//
// Boxes are padded and rounded outward exactly like the host builder's (1e-6 pad),
// so traversal stays conservative; closest-hit results do not depend on the tree.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>

#include "gp_bvh_device.hpp"
#include "gp_runtime.hpp"

namespace gp {
namespace {

constexpr double kPadD = 1e-6;

struct KNode {
    int left, right;     // child index; leaf children are encoded as -(k + 1)
    int first, last;     // sorted primitive range
};

__device__ __forceinline__ unsigned long long spread21(unsigned v) {
    unsigned long long x = v & 0x1fffffu;
    x = (x | (x << 32)) & 0x1f00000000ffffull;
    x = (x | (x << 16)) & 0x1f0000ff0000ffull;
    x = (x | (x << 8)) & 0x100f00f00f00f00full;
    x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

// Outward rounding of a double to float (as the host builder's down()/up()).
__device__ __forceinline__ float roundDown(double v) {
    float f = (float)v;
    return ((double)f > v) ? nextafterf(f, -FLT_MAX) : f;
}
__device__ __forceinline__ float roundUp(double v) {
    float f = (float)v;
    return ((double)f < v) ? nextafterf(f, FLT_MAX) : f;
}

// Padded box of input triangle t (float4 x3 per triangle, input order).
__device__ __forceinline__ void triBox(const float4* __restrict__ tris, int t, float lo[3], float hi[3]) {
    const float4 a = tris[3 * t], b = tris[3 * t + 1], c = tris[3 * t + 2];
    const float va[3] = {a.x, a.y, a.z}, vb[3] = {b.x, b.y, b.z}, vc[3] = {c.x, c.y, c.z};
    for (int k = 0; k < 3; ++k) {
        const double l = fmin(fmin((double)va[k], (double)vb[k]), (double)vc[k]) - kPadD;
        const double h = fmax(fmax((double)va[k], (double)vb[k]), (double)vc[k]) + kPadD;
        lo[k] = roundDown(l);
        hi[k] = roundUp(h);
    }
}

__global__ void k_morton(const float4* __restrict__ tris, int n, float3 lo, float3 inv,
                         unsigned long long* __restrict__ keys, unsigned* __restrict__ vals) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float bl[3], bh[3];
    triBox(tris, i, bl, bh);
    const float cx = 0.5f * (bl[0] + bh[0]), cy = 0.5f * (bl[1] + bh[1]), cz = 0.5f * (bl[2] + bh[2]);
    auto q = [](float c, float l, float s) {
        float u = (c - l) * s;
        u = fminf(fmaxf(u, 0.0f), 1.0f);
        return (unsigned)fminf(u * 2097152.0f, 2097151.0f);
    };
    const unsigned ux = q(cx, lo.x, inv.x), uy = q(cy, lo.y, inv.y), uz = q(cz, lo.z, inv.z);
    keys[i] = (spread21(ux) << 2) | (spread21(uy) << 1) | spread21(uz);
    vals[i] = (unsigned)i;
}

// longest common prefix of sorted keys i and j (positions break ties), -1 outside
__device__ __forceinline__ int lcp(const unsigned long long* __restrict__ k, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    const unsigned long long a = k[i], b = k[j];
    if (a == b) return 64 + __clz((unsigned)(i ^ j));
    return __clzll(a ^ b);
}

__global__ void k_karras(const unsigned long long* __restrict__ keys, int n, KNode* __restrict__ nodes,
                         int* __restrict__ parentInner, int* __restrict__ parentLeaf) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    const int d = (lcp(keys, n, i, i + 1) - lcp(keys, n, i, i - 1)) >= 0 ? 1 : -1;
    const int dmin = lcp(keys, n, i, i - d);
    long long lmax = 2;
    while (lcp(keys, n, i, (int)(i + lmax * d)) > dmin) lmax <<= 1;
    long long l = 0;
    for (long long t = lmax >> 1; t >= 1; t >>= 1)
        if (lcp(keys, n, i, (int)(i + (l + t) * d)) > dmin) l += t;
    const int j = (int)(i + l * d);
    const int dnode = lcp(keys, n, i, j);
    long long s = 0, t = l;
    do {
        t = (t + 1) >> 1;
        if (lcp(keys, n, i, (int)(i + (s + t) * d)) > dnode) s += t;
    } while (t > 1);
    const int gamma = (int)(i + s * d) + min(d, 0);
    KNode nd;
    nd.first = min(i, j);
    nd.last = max(i, j);
    if (nd.first == gamma) {
        nd.left = -(gamma + 1);
        parentLeaf[gamma] = i;
    } else {
        nd.left = gamma;
        parentInner[gamma] = i;
    }
    if (nd.last == gamma + 1) {
        nd.right = -(gamma + 2);
        parentLeaf[gamma + 1] = i;
    } else {
        nd.right = gamma + 1;
        parentInner[gamma + 1] = i;
    }
    nodes[i] = nd;
}

__device__ __forceinline__ float aload(const float* p) {
    return __int_as_float(__hip_atomic_load((const int*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}
__device__ __forceinline__ int aloadi(const int* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// box[6] per leaf (sorted position) and per internal node; height per internal node
// (pair-node levels below and including it; 0 for ranges of <= kLeafMax primitives)
__global__ void k_leaf_boxes(const float4* __restrict__ tris, const unsigned* __restrict__ order, int n,
                             float* __restrict__ leafBox) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    float lo[3], hi[3];
    triBox(tris, (int)order[k], lo, hi);
    for (int a = 0; a < 3; ++a) {
        leafBox[6 * k + a] = lo[a];
        leafBox[6 * k + 3 + a] = hi[a];
    }
}

__global__ void k_refit(const KNode* __restrict__ nodes, const int* __restrict__ parentInner,
                        const int* __restrict__ parentLeaf, int n, const float* __restrict__ leafBox,
                        float* __restrict__ box, int* __restrict__ height, unsigned* __restrict__ arrivals) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int p = parentLeaf[k];
    while (p >= 0) {
        __threadfence();
        if (atomicAdd(&arrivals[p], 1u) == 0u) return;   // the sibling subtree is still open
        __threadfence();
        const KNode nd = nodes[p];
        float lo[3], hi[3];
        int h = 0;
        for (int side = 0; side < 2; ++side) {
            const int c = side ? nd.right : nd.left;
            const float* b = c < 0 ? leafBox + 6 * (size_t)(-c - 1) : box + 6 * (size_t)c;
            for (int a = 0; a < 3; ++a) {
                const float l = aload(b + a), u = aload(b + 3 + a);
                lo[a] = side ? fminf(lo[a], l) : l;
                hi[a] = side ? fmaxf(hi[a], u) : u;
            }
            if (c >= 0) {
                const KNode cn = nodes[c];
                if (cn.last - cn.first + 1 > kLeafMax) h = max(h, aloadi(height + c));
            }
        }
        for (int a = 0; a < 3; ++a) {
            box[6 * (size_t)p + a] = lo[a];
            box[6 * (size_t)p + 3 + a] = hi[a];
        }
        height[p] = (nd.last - nd.first + 1 > kLeafMax) ? h + 1 : 0;
        p = parentInner[p];
    }
}

__device__ __forceinline__ int childLink(const KNode* __restrict__ nodes, int c) {
    if (c < 0) return leafCode(-c - 1, 1);
    const KNode cn = nodes[c];
    const int cnt = cn.last - cn.first + 1;
    return cnt <= kLeafMax ? leafCode(cn.first, cnt) : c;
}

__global__ void k_emit_pairs(const KNode* __restrict__ nodes, int n, const float* __restrict__ leafBox,
                             const float* __restrict__ box, float4* __restrict__ pairs) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    const KNode nd = nodes[i];
    if (nd.last - nd.first + 1 <= kLeafMax) return;   // collapsed into its parent's leaf
    const float* bl = nd.left < 0 ? leafBox + 6 * (size_t)(-nd.left - 1) : box + 6 * (size_t)nd.left;
    const float* br = nd.right < 0 ? leafBox + 6 * (size_t)(-nd.right - 1) : box + 6 * (size_t)nd.right;
    const int ll = childLink(nodes, nd.left), rl = childLink(nodes, nd.right);
    float4* q = pairs + 4 * (size_t)i;
    q[0] = make_float4(bl[0], bl[1], bl[2], __int_as_float(ll));
    q[1] = make_float4(bl[3], bl[4], bl[5], __int_as_float(rl));
    q[2] = make_float4(br[0], br[1], br[2], 0.0f);
    q[3] = make_float4(br[3], br[4], br[5], 0.0f);
}

// Four-wide node i from pair node i (see step 5; the layout is collapse4's in
// gp_bvh.hpp). `nodes` is null for the single pair of a lone leaf.
__global__ void k_collapse4(const KNode* __restrict__ nodes, int nInner, const float4* __restrict__ pairs,
                            float4* __restrict__ quads) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nInner) return;
    if (nodes) {
        const KNode nd = nodes[i];
        if (nd.last - nd.first + 1 <= kLeafMax) return;   // no pair node here
    }
    float lo[3][4], hi[3][4];
    int link[4];
    int n = 0;
    for (int side = 0; side < 2; ++side) {
        const float4* p = pairs + 4 * (size_t)i;
        const float4 a = side ? p[2] : p[0], b = side ? p[3] : p[1];
        const int c = __float_as_int(side ? p[1].w : p[0].w);
        if (c < 0) {
            lo[0][n] = a.x; lo[1][n] = a.y; lo[2][n] = a.z;
            hi[0][n] = b.x; hi[1][n] = b.y; hi[2][n] = b.z;
            link[n++] = c;
        } else {
            const float4* q = pairs + 4 * (size_t)c;
            for (int g = 0; g < 2; ++g) {
                const float4 ga = g ? q[2] : q[0], gb = g ? q[3] : q[1];
                lo[0][n] = ga.x; lo[1][n] = ga.y; lo[2][n] = ga.z;
                hi[0][n] = gb.x; hi[1][n] = gb.y; hi[2][n] = gb.z;
                link[n++] = __float_as_int(g ? q[1].w : q[0].w);
            }
        }
    }
    for (int k = n; k < 4; ++k) {   // unused slot: far point box, linked to slot 0
        for (int a = 0; a < 3; ++a) { lo[a][k] = 1e30f; hi[a][k] = 1e30f; }
        link[k] = link[0];
    }
    float4* o = quads + 8 * (size_t)i;
    for (int a = 0; a < 3; ++a) {
        o[2 * a] = make_float4(lo[a][0], lo[a][1], lo[a][2], lo[a][3]);
        o[2 * a + 1] = make_float4(hi[a][0], hi[a][1], hi[a][2], hi[a][3]);
    }
    o[6] = make_float4(__int_as_float(link[0]), __int_as_float(link[1]), __int_as_float(link[2]),
                       __int_as_float(link[3]));
    o[7] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// a scene of <= kLeafMax triangles: one leaf, tested through a pair with itself
__global__ void k_lone_leaf(const float* __restrict__ leafBox, int n, float4* __restrict__ pairs) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (int k = 0; k < n; ++k)
        for (int a = 0; a < 3; ++a) {
            lo[a] = fminf(lo[a], leafBox[6 * k + a]);
            hi[a] = fmaxf(hi[a], leafBox[6 * k + 3 + a]);
        }
    const int code = leafCode(0, n);
    pairs[0] = make_float4(lo[0], lo[1], lo[2], __int_as_float(code));
    pairs[1] = make_float4(hi[0], hi[1], hi[2], __int_as_float(code));
    pairs[2] = make_float4(lo[0], lo[1], lo[2], 0.0f);
    pairs[3] = make_float4(hi[0], hi[1], hi[2], 0.0f);
}

__global__ void k_gather(const float4* __restrict__ trisIn, const float4* __restrict__ trinIn,
                         const unsigned* __restrict__ order, int n, float4* __restrict__ tris,
                         float4* __restrict__ trin) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const unsigned t = order[k];
    tris[3 * (size_t)k + 0] = trisIn[3 * (size_t)t + 0];
    tris[3 * (size_t)k + 1] = trisIn[3 * (size_t)t + 1];
    tris[3 * (size_t)k + 2] = trisIn[3 * (size_t)t + 2];
    trin[k] = trinIn[t];
}

// ---- eye-to-scene distance (brute force, double; Ericson 5.1.5)
__device__ double dsub(double a, double b) { return a - b; }
__device__ double pointTriDist2D(const double p[3], const double a[3], const double b[3], const double c[3]) {
#pragma clang fp contract(off)
    double ab[3], ac[3], ap[3];
    for (int i = 0; i < 3; ++i) { ab[i] = b[i] - a[i]; ac[i] = c[i] - a[i]; ap[i] = p[i] - a[i]; }
    auto dot = [](const double* x, const double* y) { return x[0] * y[0] + x[1] * y[1] + x[2] * y[2]; };
    auto dist2 = [&](const double* x) {
        double d[3] = {p[0] - x[0], p[1] - x[1], p[2] - x[2]};
        return dot(d, d);
    };
    double q[3];
    const double d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0 && d2 <= 0) return dist2(a);
    double bp[3] = {p[0] - b[0], p[1] - b[1], p[2] - b[2]};
    const double d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0 && d4 <= d3) return dist2(b);
    const double vc = d1 * d4 - d3 * d2;
    if (vc <= 0 && d1 >= 0 && d3 <= 0) {
        const double v = d1 / (d1 - d3);
        for (int i = 0; i < 3; ++i) q[i] = a[i] + v * ab[i];
        return dist2(q);
    }
    double cp[3] = {p[0] - c[0], p[1] - c[1], p[2] - c[2]};
    const double d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0 && d5 <= d6) return dist2(c);
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0 && d6 <= 0) {
        const double w = d2 / (d2 - d6);
        for (int i = 0; i < 3; ++i) q[i] = a[i] + w * ac[i];
        return dist2(q);
    }
    const double va = d3 * d6 - d5 * d4;
    if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
        const double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        for (int i = 0; i < 3; ++i) q[i] = b[i] + w * (c[i] - b[i]);
        return dist2(q);
    }
    const double denom = 1.0 / (va + vb + vc);
    const double v = vb * denom, w = vc * denom;
    for (int i = 0; i < 3; ++i) q[i] = a[i] + ab[i] * v + ac[i] * w;
    return dist2(q);
}

__global__ void k_min_dist2(const float4* __restrict__ tris, size_t n, double px, double py, double pz,
                            unsigned long long* __restrict__ best) {
    __shared__ unsigned long long sBest;
    if (threadIdx.x == 0) sBest = ~0ull;
    __syncthreads();
    const double p[3] = {px, py, pz};
    unsigned long long mine = ~0ull;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride) {
        const float4 a = tris[3 * t], b = tris[3 * t + 1], c = tris[3 * t + 2];
        const double da[3] = {a.x, a.y, a.z}, db[3] = {b.x, b.y, b.z}, dc[3] = {c.x, c.y, c.z};
        const double d2 = pointTriDist2D(p, da, db, dc);
        const unsigned long long bits = (unsigned long long)__double_as_longlong(d2);
        mine = bits < mine ? bits : mine;
    }
    atomicMin(&sBest, mine);
    __syncthreads();
    if (threadIdx.x == 0) atomicMin(best, sBest);
}

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() { if (p) hipFree(p); }
};

int blocks(int64_t n) { return (int)std::max<int64_t>(1, (n + 255) / 256); }

}  // namespace

gp_status buildBvhDevice(const float4* dTrisIn, const float4* dTrinIn, int64_t n64, const float centerLo[3],
                         const float centerHi[3], DeviceBvh& out) {
    hipStream_t s = device().stream;
    const int n = (int)n64;
    auto t0 = std::chrono::steady_clock::now();
    auto alloc = [&](DevBuf& b, size_t bytes) -> gp_status {
        hipError_t e = hipMalloc(&b.p, std::max<size_t>(bytes, 16));
        if (e != hipSuccess) return fail(GP_ERR_OOM, std::string("BVH build: hipMalloc: ") + hipGetErrorString(e));
        return GP_OK;
    };
    gp_status st;
    DevBuf keys, keys2, vals, vals2, leafBox;
    if ((st = alloc(keys, sizeof(unsigned long long) * n)) != GP_OK) return st;
    if ((st = alloc(keys2, sizeof(unsigned long long) * n)) != GP_OK) return st;
    if ((st = alloc(vals, sizeof(unsigned) * n)) != GP_OK) return st;
    if ((st = alloc(vals2, sizeof(unsigned) * n)) != GP_OK) return st;
    if ((st = alloc(leafBox, sizeof(float) * 6 * (size_t)n)) != GP_OK) return st;
    float3 lo = make_float3(centerLo[0], centerLo[1], centerLo[2]);
    auto invExt = [](float l, float h) { return h > l ? 1.0f / (h - l) : 0.0f; };
    float3 inv = make_float3(invExt(centerLo[0], centerHi[0]), invExt(centerLo[1], centerHi[1]),
                             invExt(centerLo[2], centerHi[2]));
    hipLaunchKernelGGL(k_morton, dim3(blocks(n)), dim3(256), 0, s, dTrisIn, n, lo, inv,
                       (unsigned long long*)keys.p, (unsigned*)vals.p);
    GP_HIP(hipGetLastError());
    size_t tempBytes = 0;
    GP_HIP(rocprim::radix_sort_pairs(nullptr, tempBytes, (unsigned long long*)keys.p, (unsigned long long*)keys2.p,
                                     (unsigned*)vals.p, (unsigned*)vals2.p, (size_t)n, 0u, 63u, s));
    DevBuf temp;
    if ((st = alloc(temp, tempBytes)) != GP_OK) return st;
    GP_HIP(rocprim::radix_sort_pairs(temp.p, tempBytes, (unsigned long long*)keys.p, (unsigned long long*)keys2.p,
                                     (unsigned*)vals.p, (unsigned*)vals2.p, (size_t)n, 0u, 63u, s));
    const unsigned long long* skeys = (const unsigned long long*)keys2.p;
    const unsigned* order = (const unsigned*)vals2.p;
    hipLaunchKernelGGL(k_leaf_boxes, dim3(blocks(n)), dim3(256), 0, s, dTrisIn, order, n, (float*)leafBox.p);
    GP_HIP(hipGetLastError());

    const size_t nPairs = (size_t)std::max(1, n - 1);
    float4* pairs = nullptr;
    {
        hipError_t e = hipMalloc((void**)&pairs, nPairs * 4 * sizeof(float4));
        if (e != hipSuccess) return fail(GP_ERR_OOM, "BVH build: node allocation failed");
    }
    struct PairsGuard {
        float4*& p;
        bool keep = false;
        ~PairsGuard() { if (!keep && p) hipFree(p); }
    } pg{pairs};
    float4* quads = nullptr;
    {
        hipError_t e = hipMalloc((void**)&quads, nPairs * 8 * sizeof(float4));
        if (e != hipSuccess) return fail(GP_ERR_OOM, "BVH build: node allocation failed");
    }
    PairsGuard qg{quads};
    int height = 1;
    if (n <= kLeafMax) {
        hipLaunchKernelGGL(k_lone_leaf, dim3(1), dim3(64), 0, s, (const float*)leafBox.p, n, pairs);
        hipLaunchKernelGGL(k_collapse4, dim3(1), dim3(64), 0, s, (const KNode*)nullptr, 1, (const float4*)pairs,
                           quads);
        GP_HIP(hipGetLastError());
    } else {
        DevBuf nodes, parentInner, parentLeaf, box, heights, arrivals;
        if ((st = alloc(nodes, sizeof(KNode) * (size_t)(n - 1))) != GP_OK) return st;
        if ((st = alloc(parentInner, sizeof(int) * (size_t)(n - 1))) != GP_OK) return st;
        if ((st = alloc(parentLeaf, sizeof(int) * (size_t)n)) != GP_OK) return st;
        if ((st = alloc(box, sizeof(float) * 6 * (size_t)(n - 1))) != GP_OK) return st;
        if ((st = alloc(heights, sizeof(int) * (size_t)(n - 1))) != GP_OK) return st;
        if ((st = alloc(arrivals, sizeof(unsigned) * (size_t)(n - 1))) != GP_OK) return st;
        GP_HIP(hipMemsetAsync(arrivals.p, 0, sizeof(unsigned) * (size_t)(n - 1), s));
        GP_HIP(hipMemsetAsync(parentInner.p, 0xff, sizeof(int) * (size_t)(n - 1), s));   // root: -1
        hipLaunchKernelGGL(k_karras, dim3(blocks(n - 1)), dim3(256), 0, s, skeys, n, (KNode*)nodes.p,
                           (int*)parentInner.p, (int*)parentLeaf.p);
        hipLaunchKernelGGL(k_refit, dim3(blocks(n)), dim3(256), 0, s, (const KNode*)nodes.p,
                           (const int*)parentInner.p, (const int*)parentLeaf.p, n, (const float*)leafBox.p,
                           (float*)box.p, (int*)heights.p, (unsigned*)arrivals.p);
        hipLaunchKernelGGL(k_emit_pairs, dim3(blocks(n - 1)), dim3(256), 0, s, (const KNode*)nodes.p, n,
                           (const float*)leafBox.p, (const float*)box.p, pairs);
        hipLaunchKernelGGL(k_collapse4, dim3(blocks(n - 1)), dim3(256), 0, s, (const KNode*)nodes.p, n - 1,
                           (const float4*)pairs, quads);
        GP_HIP(hipGetLastError());
        GP_HIP(hipMemcpyAsync(&height, heights.p, sizeof(int), hipMemcpyDeviceToHost, s));
    }
    float4 *tris = nullptr, *trin = nullptr;
    {
        hipError_t e1 = hipMalloc((void**)&tris, sizeof(float4) * 3 * (size_t)n);
        hipError_t e2 = hipMalloc((void**)&trin, sizeof(float4) * (size_t)n);
        if (e1 != hipSuccess || e2 != hipSuccess) {
            if (tris) hipFree(tris);
            if (trin) hipFree(trin);
            return fail(GP_ERR_OOM, "BVH build: triangle allocation failed");
        }
    }
    hipLaunchKernelGGL(k_gather, dim3(blocks(n)), dim3(256), 0, s, dTrisIn, dTrinIn, order, n, tris, trin);
    hipError_t ge = hipGetLastError();
    hipError_t se = hipStreamSynchronize(s);
    if (ge != hipSuccess || se != hipSuccess) {
        hipFree(tris);
        hipFree(trin);
        return fail(GP_ERR_HIP, std::string("BVH build: ") + hipGetErrorString(ge != hipSuccess ? ge : se));
    }
    qg.keep = true;   // the pair nodes are freed on return
    out.nodes = quads;
    out.tris = tris;
    out.trin = trin;
    out.nPairs = (int64_t)nPairs;
    out.depth = std::max(1, height);
    out.stackNeed = 3 * ((out.depth + 1) / 2);   // four-wide levels: pair levels 1, 3, 5, ...
    out.buildMs = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return GP_OK;
}

gp_status deviceMinDistance2(const float4* dTris, int64_t n, const double p[3], double& d2) {
    hipStream_t s = device().stream;
    DevBuf best;
    GP_HIP(hipMalloc(&best.p, sizeof(unsigned long long)));
    GP_HIP(hipMemsetAsync(best.p, 0xff, sizeof(unsigned long long), s));
    const int g = (int)std::min<int64_t>(4096, std::max<int64_t>(1, (n + 255) / 256));
    hipLaunchKernelGGL(k_min_dist2, dim3(g), dim3(256), 0, s, dTris, (size_t)n, p[0], p[1], p[2],
                       (unsigned long long*)best.p);
    GP_HIP(hipGetLastError());
    unsigned long long bits = 0;
    GP_HIP(hipMemcpyAsync(&bits, best.p, sizeof bits, hipMemcpyDeviceToHost, s));
    GP_HIP(hipStreamSynchronize(s));
    std::memcpy(&d2, &bits, sizeof d2);
    return GP_OK;
}

}  // namespace gp
