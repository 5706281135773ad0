// Device functions for the forward light-tracing kernels (gfx950, wave64).
//
// Arithmetic follows the reference's CPU forward path in float32, except where a
// double keeps the same sample: the photon RNG (PCG32 seeded per photon by
// SplitMix64, src/rng.h:6-42) and the wavelength CDF inversion
// (EmissionSampler::sampleAt, src/spectrum.h:270-279) run in integer / double so the GPU draws exactly the
// same photons as the CPU reference for the same (seedBase, salt).
#pragma once
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

#include "gp_types.hpp"

namespace gp {

constexpr float kPiF = 3.14159265358979f;

// ------------------------------------------------------------------ vectors
struct F3 {
    float x, y, z;
};
__device__ __forceinline__ F3 f3(float x, float y, float z) { return F3{x, y, z}; }
__device__ __forceinline__ F3 operator+(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ F3 operator-(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ F3 operator*(F3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ F3 operator-(F3 a) { return {-a.x, -a.y, -a.z}; }
__device__ __forceinline__ float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ F3 normalize(F3 a) {
    float inv = 1.0f / sqrtf(dot(a, a));
    return a * inv;
}
__device__ __forceinline__ F3 reflectDir(F3 d, F3 n) { return d - n * (2.0f * dot(d, n)); }
__device__ __forceinline__ F3 load3(const float* p) { return F3{p[0], p[1], p[2]}; }
__device__ __forceinline__ F3 xyz(float4 v) { return F3{v.x, v.y, v.z}; }
// The same 16 bytes read as the other vector type.
__device__ __forceinline__ uint4 bitcast4(float4 v) {
    return make_uint4((unsigned)__float_as_int(v.x), (unsigned)__float_as_int(v.y), (unsigned)__float_as_int(v.z),
                      (unsigned)__float_as_int(v.w));
}
__device__ __forceinline__ float4 bitcast4f(uint4 v) {
    return make_float4(__int_as_float((int)v.x), __int_as_float((int)v.y), __int_as_float((int)v.z),
                       __int_as_float((int)v.w));
}

// Loads through a global-address-space pointer. Scene tables are reached through
// pointers stored in the job record, which the compiler cannot place in an address
// space; a plain dereference then becomes a FLAT load (address aperture checks, and
// it shares the LDS wait counter with the traversal stack). These force GLOBAL loads.
typedef float gp_f4v __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 ldg4(const float4* p) {
    const gp_f4v v = *(const __attribute__((address_space(1))) gp_f4v*)p;
    return make_float4(v.x, v.y, v.z, v.w);
}

// Orthonormal basis around unit n (Duff et al. 2017, the reference's onb).
__device__ __forceinline__ void frame(F3 n, F3& t, F3& b) {
    float sign = copysignf(1.0f, n.z);
    float a = -1.0f / (sign + n.z);
    float d = n.x * n.y * a;
    t = F3{1.0f + sign * n.x * n.x * a, sign * d, -sign * n.x};
    b = F3{d, sign + n.y * n.y * a, -n.y};
}

// ------------------------------------------------------------------ rng
// PCG-XSH-RR 32 (O'Neill). `inc` is not stored per path: it is a pure function of
// the photon index, recomputed where a kernel draws.
__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}
struct Rng {
    uint64_t state, inc;
    __device__ __forceinline__ uint32_t next() {
        uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((~rot + 1u) & 31u));
    }
    // (next >> 8) * 2^-24: exactly the reference's double uniform, representable in float.
    __device__ __forceinline__ float uniform() { return (float)(next() >> 8) * 5.9604644775390625e-8f; }
    __device__ __forceinline__ double uniformD() { return (double)(next() >> 8) * (1.0 / 16777216.0); }
};
__device__ __forceinline__ uint64_t photonInc(uint64_t k, uint64_t salt) {
    return (splitmix64(k ^ salt) << 1u) | 1u;
}
__device__ __forceinline__ Rng seedPhoton(uint64_t k, uint64_t salt) {
    Rng r;
    r.state = 0;
    r.inc = photonInc(k, salt);
    r.next();
    r.state += splitmix64(k + salt);
    r.next();
    return r;
}

// ------------------------------------------------------------------ spectra
__device__ __forceinline__ float specEval(const DevSpectrum& s, float w) {
    if (s.band_hi > s.band_lo && !(w >= s.band_lo && w <= s.band_hi)) return 0.0f;
    switch (s.kind) {
        case 0: return s.p[0];
        case 1: return (w < s.p[2]) ? s.p[0] : s.p[1];
        case 2: {
            float t = (w - s.p[2]) / s.p[3];
            return s.p[0] + s.p[1] * __expf(-0.5f * t * t);
        }
        case 4: {
            float l = w * 1e-3f;
            float l2 = l * l;
            float n2 = 1.0f + s.p[0] * l2 / (l2 - s.p[3]) + s.p[1] * l2 / (l2 - s.p[4]) +
                       s.p[2] * l2 / (l2 - s.p[5]);
            return sqrtf(n2 > 1.0f ? n2 : 1.0f);
        }
        case 5: {
            float um = w * 1e-3f;
            return s.p[0] + s.p[1] / (um * um);
        }
        case 6: {   // sampled: linear interpolation, clamped to the end samples
            const float f = (w - s.p[0]) * s.p[1];
            if (!(f > 0.0f)) return s.tab[0];
            if (f >= (float)(s.n - 1)) return s.tab[s.n - 1];
            const int i = (int)f;
            const float fr = f - (float)i;
            return s.tab[i] * (1.0f - fr) + s.tab[i + 1] * fr;
        }
        default: return 0.0f;
    }
}

// CIE 1931 observer, multi-lobe Gaussian fit (Wyman, Sloan, Shirley 2013).
__device__ __forceinline__ float lobe(float x, float mu, float s1, float s2) {
    float t = (x - mu) * ((x < mu) ? s1 : s2);
    return __expf(-0.5f * t * t);
}
__device__ __forceinline__ F3 cie(float w) {
    float X = 0.362f * lobe(w, 442.0f, 0.0624f, 0.0374f) + 1.056f * lobe(w, 599.8f, 0.0264f, 0.0323f) -
              0.065f * lobe(w, 501.1f, 0.0490f, 0.0382f);
    float Y = 0.821f * lobe(w, 568.8f, 0.0213f, 0.0247f) + 0.286f * lobe(w, 530.9f, 0.0613f, 0.0322f);
    float Z = 1.217f * lobe(w, 437.0f, 0.0845f, 0.0278f) + 0.681f * lobe(w, 459.0f, 0.0385f, 0.0725f);
    return F3{X, Y, Z};
}

// Inversion of a normalised 1-nm CDF over [360, 830), in double as on the host: the
// last bin lo with cdf[lo] <= u. The guide table brackets it (cdf[guide[k]] <= k/G <= u
// < (k+1)/G < cdf[guide[k+1] + 1] for k = floor(u G), exact since G is a power of two),
// so the binary search that follows takes ~3 steps instead of 9 and returns the same bin.
__device__ __forceinline__ double invertCdf(const double* __restrict__ cdf, const int32_t* __restrict__ guide,
                                            double u, double& pdf) {
    const int k = min(max((int)(u * (double)kCdfGuide), 0), kCdfGuide - 1);
    int lo = guide[k], hi = min(guide[k + 1] + 1, kLambdaBins);
    while (lo + 1 < hi) {
        int mid = (lo + hi) >> 1;
        if (cdf[mid] <= u) lo = mid;
        else hi = mid;
    }
    double c0 = cdf[lo], c1 = cdf[lo + 1];
    double frac = (c1 > c0) ? (u - c0) / (c1 - c0) : 0.5;
    pdf = c1 - c0;
    return 360.0 + ((double)lo + frac);
}

// ------------------------------------------------------------------ sampling
__device__ __forceinline__ F3 cosineDir(F3 n, Rng& rng) {
    float u1 = rng.uniform();
    float u2 = rng.uniform();
    float r = sqrtf(u1);
    float s, c;
    sincosf(6.28318530717959f * u2, &s, &c);
    float lz = sqrtf(1.0f - u1);
    F3 t, b;
    frame(n, t, b);
    return normalize(t * (r * c) + b * (r * s) + n * lz);
}
// Power-cosine lobe about mdir (the reference's glossyDirUV).
__device__ __forceinline__ F3 glossyDir(F3 mdir, float rough, Rng& rng) {
    float u1 = rng.uniform();
    float u2 = rng.uniform();
    float rr = rough < 1e-3f ? 1e-3f : rough;
    float e = 2.0f / (rr * rr) - 2.0f;
    if (e < 0.0f) e = 0.0f;
    float cosT = powf(u1, 1.0f / (e + 1.0f));
    float sinT = sqrtf(fmaxf(0.0f, 1.0f - cosT * cosT));
    float s, c;
    sincosf(6.28318530717959f * u2, &s, &c);
    F3 t, b;
    frame(mdir, t, b);
    return normalize(t * (sinT * c) + b * (sinT * s) + mdir * cosT);
}

// ------------------------------------------------------------------ intersection
struct RayPre {
    F3 o, d, inv;     // inv = 1 / d (infinite for a zero component)
    F3 invF, cn, cf;  // closest-hit slab test: the near plane of an axis is crossed at
                      // fma(x, invF, cn), the far plane at fma(x, invF, cf)
    int nx, ny, nz;   // float4 index of the near-plane bounds of axis x, y, z in a node
    int kx, ky, kz;
    float Sx, Sy, Sz;
};
__device__ __forceinline__ float pick(F3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

__device__ __forceinline__ RayPre prepareRay(F3 o, F3 d) {
    RayPre r;
    r.o = o;
    r.d = d;
    // Closest-hit slab distances as one FMA per plane, t = x * invF - o * invF,
    // conservative per axis: rounding o * invF and the addend costs at most 2u |o invF|
    // (u = 2^-24), so the lower (near) plane's addend is lowered and the upper (far)
    // plane's raised by e = 4u |o invF|; the FMA's own relative rounding is covered by
    // the 2^-21 widening of the exit distance (slab4). A zero (or sub-1e-30) direction
    // component gets a finite inverse of its sign, 1e30: with the margin e the slab then
    // spans every distance unless the origin lies outside it by more than 2^-22 of its
    // coordinate, where the ray cannot reach it. The any-hit test keeps (x - o) * inv
    // with the exact (possibly infinite) inverse.
    r.inv = F3{1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
    auto invOf = [](float c) { return fabsf(c) < 1e-30f ? copysignf(1e30f, c) : 1.0f / c; };
    r.invF = F3{invOf(d.x), invOf(d.y), invOf(d.z)};
    auto addends = [](float oc, float inv, float& cn, float& cf) {
        const float p = oc * inv;
        const float e = fabsf(p) * 2.38418579e-7f;   // 2^-22 |o inv|
        cn = -p - e;
        cf = -p + e;
    };
    addends(o.x, r.invF.x, r.cn.x, r.cf.x);
    addends(o.y, r.invF.y, r.cn.y, r.cf.y);
    addends(o.z, r.invF.z, r.cn.z, r.cf.z);
    // a node holds lo then hi of each axis: the near plane is lo for a positive
    // direction component, hi for a negative one (the inverse's sign, -0 included)
    r.nx = r.invF.x < 0.0f ? 1 : 0;
    r.ny = r.invF.y < 0.0f ? 3 : 2;
    r.nz = r.invF.z < 0.0f ? 5 : 4;
    float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    int kz = (ax >= ay && ax >= az) ? 0 : ((ay >= az) ? 1 : 2);
    int kx = kz + 1;
    if (kx == 3) kx = 0;
    int ky = kx + 1;
    if (ky == 3) ky = 0;
    float dz = pick(d, kz);
    if (dz < 0.0f) { int t = kx; kx = ky; ky = t; }
    r.kx = kx; r.ky = ky; r.kz = kz;
    r.Sx = pick(d, kx) / dz;
    r.Sy = pick(d, ky) / dz;
    r.Sz = 1.0f / dz;
    return r;
}

// Sign-exact fallback for an edge function that rounded to zero: the two float
// products are exact in double, so the double difference has the true sign.
__device__ __forceinline__ float edgeFallback(float a, float b, float c, float d) {
    double v = (double)a * (double)b - (double)c * (double)d;
    return v > 0.0 ? 1e-30f : (v < 0.0 ? -1e-30f : 0.0f);
}

// Watertight two-sided ray/triangle test (Woop, Benthin, Wald 2013) in float.
// The three edge-function differences are evaluated without FMA contraction so a
// shared edge is decided identically (with opposite sign) by both triangles.
// CLOSED: accept t == tmax too (the closest-hit tie rule decides such hits).
template <bool CLOSED = false>
__device__ __forceinline__ bool hitTri(const RayPre& r, F3 v0, F3 v1, F3 v2, float tmin, float tmax,
                                       float& tOut) {
#pragma clang fp contract(off)
    F3 A = v0 - r.o, B = v1 - r.o, C = v2 - r.o;
    float Akz = pick(A, r.kz), Bkz = pick(B, r.kz), Ckz = pick(C, r.kz);
    float Ax = pick(A, r.kx) - r.Sx * Akz, Ay = pick(A, r.ky) - r.Sy * Akz;
    float Bx = pick(B, r.kx) - r.Sx * Bkz, By = pick(B, r.ky) - r.Sy * Bkz;
    float Cx = pick(C, r.kx) - r.Sx * Ckz, Cy = pick(C, r.ky) - r.Sy * Ckz;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        if (U == 0.0f) U = edgeFallback(Cx, By, Cy, Bx);
        if (V == 0.0f) V = edgeFallback(Ax, Cy, Ay, Cx);
        if (W == 0.0f) W = edgeFallback(Bx, Ay, By, Ax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    float det = U + V + W;
    if (det == 0.0f) return false;
    float T = U * (r.Sz * Akz) + V * (r.Sz * Bkz) + W * (r.Sz * Ckz);
    // Cheap rejections ahead of the division, each implying the exact test below
    // fails: t = T / det <= 0 when the signs differ (tmin > 0), and t > tmax when
    // |T| exceeds tmax |det| by more than the ~2^-22 relative rounding of the two.
    if (tmin > 0.0f && !((T > 0.0f) == (det > 0.0f) && T != 0.0f)) return false;
    if (fabsf(T) > tmax * fabsf(det) * 1.0000010f) return false;
    float t = T * (1.0f / det);
    if (!(t >= tmin && (CLOSED ? t <= tmax : t < tmax))) return false;
    tOut = t;
    return true;
}

// Traversal kernels run workgroups of this many threads; the LDS stack is laid out
// with this stride so lane addresses need no multiply.
constexpr int kTravThreads = 256;

// Traversal stack: the top `SD` entries live in LDS at [slot * 256 + lane], deeper
// entries spill to a per-thread global area. The stack holds at most three entries
// per level of the four-wide tree (the hit siblings of the path's nodes), so a tree
// of depth D needs 3 D entries; the host guarantees SD + spill depth >= 3 D, so it
// never overflows. With WITH_T an entry is (node, entry distance) in one 64-bit LDS
// word; without (any-hit) just the node. SPILL = false is the variant for trees that
// fit in LDS (no spill area allocated, no spill branches).
template <int SD, bool WITH_T = true, bool SPILL = true, int ST = kTravThreads>
struct Stack {
    static constexpr int kStride = ST;   // lanes of the workgroup (LDS column stride)
    int2* e2;             // LDS, this lane's column of SD entries (WITH_T)
    int* e1;              // LDS, this lane's column of SD entries (!WITH_T)
    int2* spill;          // global, spillDepth entries for this thread (SPILL)
    int sp;
    __device__ __forceinline__ void push(int node, float t) {
        if (!SPILL || sp < SD) {
            if (WITH_T) e2[sp * ST] = make_int2(node, __float_as_int(t));
            else e1[sp * ST] = node;
        } else {
            spill[sp - SD] = make_int2(node, __float_as_int(t));
        }
        ++sp;
    }
    __device__ __forceinline__ void pop(int& node, float& t) {
        --sp;
        if (!SPILL || sp < SD) {
            if (WITH_T) {
                const int2 v = e2[sp * ST];
                node = v.x;
                t = __int_as_float(v.y);
            } else {
                node = e1[sp * ST];
                t = 0.0f;
            }
        } else {
            // a distinct load form keeps the compiler from merging the LDS and the
            // global load into one FLAT load through a selected pointer
            const unsigned long long v =
                __builtin_nontemporal_load((const unsigned long long*)(spill + (sp - SD)));
            node = (int)(unsigned)(v & 0xffffffffull);
            t = __int_as_float((int)(unsigned)(v >> 32));
        }
    }
};

// Slab test of the four slot boxes of a four-wide node: entry distances, hit flags and
// links. Exit distances are widened by 2^-21 relative so rounding can only over-report
// a box (Ize 2013, robust BVH traversal). FMA (closest hit): the ray loads each axis's
// near-plane bounds and far-plane bounds by its direction signs (RayPre::nx..nz), so
// entry = max of the near planes' fused multiply-adds and exit = min of the far planes'
// with no per-axis min/max (the same distances as min/max of the two planes: with
// lo <= hi and cn <= cf the near plane's FMA is never the larger); -24 instructions per
// node visit. Without FMA (any hit): (x - o) * inv with the exact inverse and per-axis
// min/max, which also orders the planes of a zero direction component (the addends'
// six registers would cost k_connect a wave per SIMD).
template <bool FMA>
__device__ __forceinline__ void slab4(const RayPre& r, const float4* __restrict__ nd, float tmin, float tmax,
                                      float tin[4], bool h[4], int l[4]) {
    const float widen = 1.00000048f;
    if constexpr (FMA) {
        const float4 nx = ldg4(nd + r.nx), fx = ldg4(nd + (r.nx ^ 1));
        const float4 ny = ldg4(nd + r.ny), fy = ldg4(nd + (r.ny ^ 1));
        const float4 nz = ldg4(nd + r.nz), fz = ldg4(nd + (r.nz ^ 1)), lk = ldg4(nd + 6);
        const float NX[4] = {nx.x, nx.y, nx.z, nx.w}, FX[4] = {fx.x, fx.y, fx.z, fx.w};
        const float NY[4] = {ny.x, ny.y, ny.z, ny.w}, FY[4] = {fy.x, fy.y, fy.z, fy.w};
        const float NZ[4] = {nz.x, nz.y, nz.z, nz.w}, FZ[4] = {fz.x, fz.y, fz.z, fz.w};
        l[0] = __float_as_int(lk.x); l[1] = __float_as_int(lk.y); l[2] = __float_as_int(lk.z); l[3] = __float_as_int(lk.w);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const float xn = fmaf(NX[k], r.invF.x, r.cn.x), xf = fmaf(FX[k], r.invF.x, r.cf.x);
            const float yn = fmaf(NY[k], r.invF.y, r.cn.y), yf = fmaf(FY[k], r.invF.y, r.cf.y);
            const float zn = fmaf(NZ[k], r.invF.z, r.cn.z), zf = fmaf(FZ[k], r.invF.z, r.cf.z);
            tin[k] = fmaxf(fmaxf(xn, yn), fmaxf(zn, tmin));
            const float tout = fminf(fminf(xf, yf), fminf(zf, tmax)) * widen;
            h[k] = tin[k] <= tout;
        }
    } else {
        const float4 lx = ldg4(nd), hx = ldg4(nd + 1), ly = ldg4(nd + 2), hy = ldg4(nd + 3);
        const float4 lz = ldg4(nd + 4), hz = ldg4(nd + 5), lk = ldg4(nd + 6);
        const float LX[4] = {lx.x, lx.y, lx.z, lx.w}, HX[4] = {hx.x, hx.y, hx.z, hx.w};
        const float LY[4] = {ly.x, ly.y, ly.z, ly.w}, HY[4] = {hy.x, hy.y, hy.z, hy.w};
        const float LZ[4] = {lz.x, lz.y, lz.z, lz.w}, HZ[4] = {hz.x, hz.y, hz.z, hz.w};
        l[0] = __float_as_int(lk.x); l[1] = __float_as_int(lk.y); l[2] = __float_as_int(lk.z); l[3] = __float_as_int(lk.w);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const float x0 = (LX[k] - r.o.x) * r.inv.x, x1 = (HX[k] - r.o.x) * r.inv.x;
            const float y0 = (LY[k] - r.o.y) * r.inv.y, y1 = (HY[k] - r.o.y) * r.inv.y;
            const float z0 = (LZ[k] - r.o.z) * r.inv.z, z1 = (HZ[k] - r.o.z) * r.inv.z;
            tin[k] = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), tmin));
            const float tout = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fminf(fmaxf(z0, z1), tmax)) * widen;
            h[k] = tin[k] <= tout;
        }
    }
}

// Closest-hit visit: the entry distances (+inf for a missed slot) sorted nearest first
// together with their links by a five-exchange network; returns the number of slots
// hit, which come first. A hit's distance is clamped to FLT_MAX so that hits sort
// ahead of misses even when a ray parallel to a slab makes it infinite.
__device__ __forceinline__ int visit4(const RayPre& r, const float4* __restrict__ nd, float tmin, float tmax,
                                      float t[4], int l[4]) {
    float tin[4];
    bool h[4];
    slab4<true>(r, nd, tmin, tmax, tin, h, l);
    int n = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        t[k] = h[k] ? fminf(tin[k], FLT_MAX) : __int_as_float(0x7f800000);
        n += h[k] ? 1 : 0;
    }
    auto cs = [&](int a, int b) {
        const bool sw = t[b] < t[a];
        const float ta = t[a], tb = t[b];
        const int la = l[a], lb = l[b];
        t[a] = sw ? tb : ta; t[b] = sw ? ta : tb;
        l[a] = sw ? lb : la; l[b] = sw ? la : lb;
    };
    cs(0, 1); cs(2, 3); cs(0, 2); cs(1, 3); cs(1, 2);
    return n;
}

// Any-hit visit: hit flags of the four slots, without the sort (any occluder ends the
// ray, so the visit order only matters for speed: -1 to -3% k_connect time against the
// sorted visit); returns the hit count.
__device__ __forceinline__ int visit4Any(const RayPre& r, const float4* __restrict__ nd, float tmin, float tmax,
                                         bool h[4], int l[4]) {
    float tin[4];
    slab4<false>(r, nd, tmin, tmax, tin, h, l);
    return (h[0] ? 1 : 0) + (h[1] ? 1 : 0) + (h[2] ? 1 : 0) + (h[3] ? 1 : 0);
}

// Traversal statistics of one query (the reference's TraversalStats, src/bvh.h:79,
// for the four-wide tree): node visits = four-wide nodes loaded and slab-tested (a
// leaf is a link in its parent, not a node), triangle tests = triangles intersected.
// Counted only by the statistics variants of the kernels; TravCount<false> compiles
// away.
template <bool ON>
struct TravCount {
    __device__ void node() {}
    __device__ void tris(int) {}
};
template <>
struct TravCount<true> {
    unsigned nodes = 0, tri = 0;
    __device__ void node() { ++nodes; }
    __device__ void tris(int c) { tri += (unsigned)c; }
};

template <int SD, bool SPILL, int ST>
__device__ __forceinline__ bool popNear(Stack<SD, true, SPILL, ST>& stk, float tmax, int& node, float& t) {
    while (stk.sp > 0) {
        stk.pop(node, t);
        if (t <= tmax) return true;
    }
    return false;
}

// Closest hit over the four-wide tree. Returns the leaf-order triangle index (or -1)
// and its distance. Front-to-back: the nearest hit slot is visited next and the
// others pushed farthest first with their entry distances; a popped entry farther
// than the current hit is skipped. Loop shape: while-while with postponed leaves
// (Aila & Laine 2009, speculative traversal): a lane that reaches a leaf keeps it and
// walks on through inner nodes until every lane of the wave still in the node loop
// holds one, so that the lanes visit nodes together and test leaves together.
// The visit order therefore depends on the wave's other rays; the result does not:
// among triangles at exactly the closest distance the lowest index wins (each of them
// is tested, since no box on its way is entered beyond that distance).
template <int SD, bool SPILL, class Cnt, int ST>
__device__ __forceinline__ int traceClosest(const float4* __restrict__ nodes,
                                            const float4* __restrict__ tris, const RayPre& r,
                                            float tmin, float& tmax, Stack<SD, true, SPILL, ST>& stk, Cnt& cnt) {
    constexpr int kDone = -2147483647 - 1;   // no node left (leaf links are ~code, code < 2^31 - 1)
    int best = -1;
    int node = 0, leaf = 0;          // leaf: a postponed leaf link, 0 = none
    float nt = 0.0f;                 // entry distance of `node`
    stk.sp = 0;
    while (true) {
        while (node >= 0) {
            float t[4];
            int l[4];
            const int n = visit4(r, nodes + 8 * node, tmin, tmax, t, l);
            cnt.node();
            if (!SPILL) {
                stk.e2[stk.sp * ST] = make_int2(l[3], __float_as_int(t[3]));
                stk.sp += n > 3 ? 1 : 0;
                stk.e2[stk.sp * ST] = make_int2(l[2], __float_as_int(t[2]));
                stk.sp += n > 2 ? 1 : 0;
                stk.e2[stk.sp * ST] = make_int2(l[1], __float_as_int(t[1]));
                stk.sp += n > 1 ? 1 : 0;
            } else {
                if (n > 3) stk.push(l[3], t[3]);
                if (n > 2) stk.push(l[2], t[2]);
                if (n > 1) stk.push(l[1], t[1]);
            }
            if (n > 0) {
                node = l[0];
                nt = t[0];
            } else if (!popNear(stk, tmax, node, nt)) {
                node = kDone;
            }
            if (node < 0 && node != kDone && leaf == 0) {
                leaf = node;
                if (!popNear(stk, tmax, node, nt)) node = kDone;
            }
            if (__ballot(leaf == 0) == 0ull) break;
        }
        while (leaf != 0) {
            const int code = ~leaf;
            const int first = code >> 3, count = (code & 7) + 1;
            cnt.tris(count);
            // one triangle at a time: a whole-leaf prefetch holds 48 more registers,
            // which costs the closest-hit kernel a resident wave per SIMD (c2: 4% slower)
            for (int k = 0; k < count; ++k) {
                const float4* tp = tris + 3 * (first + k);
                const float4 v0 = ldg4(tp), v1 = ldg4(tp + 1), v2 = ldg4(tp + 2);
                float t;
                if (hitTri<true>(r, xyz(v0), xyz(v1), xyz(v2), tmin, tmax, t) && (t < tmax || first + k < best)) {
                    tmax = t;
                    best = first + k;
                }
            }
            leaf = 0;
            // the node taken while the leaves waited may now lie beyond the hit (an
            // entry distance above tmax: nothing in it can be closer); -0.5% c2 time
            if (node != kDone && nt > tmax && !popNear(stk, tmax, node, nt)) node = kDone;
            if (node < 0 && node != kDone) {
                leaf = node;
                if (!popNear(stk, tmax, node, nt)) node = kDone;
            }
        }
        if (node == kDone) break;
    }
    return best;
}

// Any hit on [tmin, tmax) over the four-wide tree. Same loop shape; hit slots are
// visited lowest slot first.
template <int SD, bool SPILL, class Cnt>
__device__ __forceinline__ bool traceAny(const float4* __restrict__ nodes,
                                         const float4* __restrict__ tris, const RayPre& r,
                                         float tmin, float tmax, Stack<SD, false, SPILL>& stk, Cnt& cnt) {
    int node = 0;
    stk.sp = 0;
    while (true) {
        bool live = true;
        while (node >= 0) {
            bool h[4];
            int l[4];
            const int n = visit4Any(r, nodes + 8 * node, tmin, tmax, h, l);
            cnt.node();
            // The lowest hit slot is visited next; each higher hit is pushed when a lower
            // one turns up. Without spill the write is unconditional: the stack top
            // advances at most twice before the last write, so it stays within S.
            int next = 0;
            bool have = false;
#pragma unroll
            for (int k = 3; k >= 0; --k) {
                if (!SPILL) {
                    stk.e1[stk.sp * kTravThreads] = next;
                    stk.sp += (h[k] && have) ? 1 : 0;
                } else if (h[k] && have) {
                    stk.push(next, 0.0f);
                }
                next = h[k] ? l[k] : next;
                have = have || h[k];
            }
            if (n > 0) {
                node = next;
            } else if (stk.sp > 0) {
                float tt;
                stk.pop(node, tt);
            } else {
                live = false;
                break;
            }
        }
        if (!live) break;
        const int code = ~node;
        const int first = code >> 3, count = (code & 7) + 1;
        bool hit = false;
        for (int k = 0; k < count && !hit; ++k) {
            const float4* tp = tris + 3 * (first + k);
            const float4 v0 = ldg4(tp), v1 = ldg4(tp + 1), v2 = ldg4(tp + 2);
            float t;
            cnt.tris(1);
            hit = hitTri(r, xyz(v0), xyz(v1), xyz(v2), tmin, tmax, t);
        }
        if (hit) return true;
        if (stk.sp == 0) break;
        float tt;
        stk.pop(node, tt);
    }
    return false;
}

// ------------------------------------------------------------------ wave helpers
__device__ __forceinline__ unsigned laneId() { return __lane_id(); }
// Wave-aggregated append: every lane with `pred` gets a unique slot from *counter.
__device__ __forceinline__ unsigned waveAppend(unsigned* counter, bool pred) {
    unsigned long long mask = __ballot(pred);
    unsigned cnt = __popcll(mask);
    unsigned base = 0;
    int leader = __ffsll((long long)mask) - 1;
    if (cnt) {
        if ((int)laneId() == leader) base = atomicAdd(counter, cnt);
        base = __shfl(base, leader < 0 ? 0 : leader);
    }
    unsigned lo = (unsigned)(mask & 0xffffffffull), hi = (unsigned)(mask >> 32);
    unsigned rank = __builtin_amdgcn_mbcnt_hi(hi, __builtin_amdgcn_mbcnt_lo(lo, 0u));
    return base + rank;
}

}  // namespace gp
