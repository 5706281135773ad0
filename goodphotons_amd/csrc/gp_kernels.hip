// Wavefront forward light-tracing kernels for CDNA4 (gfx950).
//
// One iteration of the wavefront:
//   k_schedule  (1 block)  refill free queue slots from the jobs' photon budgets
//   k_emit                 emission + hero bundle + direct light->camera connection
//   k_bin_*                (gp_tuning.ray_sort only) counting sort of the paths by
//                          origin cell and direction octant into a third queue
//   k_extend               closest hit for every live path (four-wide nodes, LDS stack)
//   k_shade                material scatter, Russian roulette, camera connection,
//                          survivors compacted into the other queue by wave ballot
//   k_connect              any-hit for every connection; unoccluded ones splat into
//                          the int64 fixed-point film with 64-bit atomics
//   k_mirror               (scenes with flat mirrors only) camera connections of the
//                          queued vertices through each mirror plane
// All kernels are fixed-grid and grid-stride over counts that live on the device,
// so an iteration needs no host round trip and can be captured in a hipGraph.
//
// Per path the arithmetic follows the reference's Renderer::tracePhotonHero /
// tracePhotonHeroLoop (src/render.h:2252-2631), connectGeom / connectHero
// (:499-577) and refractOrReflect (:2657-2694), in float32.
#include <hip/hip_runtime.h>

#include "gp_device.hpp"
#include "gp_kernels.hpp"

namespace gp {

namespace {

constexpr float kInvPi = 0.318309886183791f;

__device__ __forceinline__ unsigned long long toFixed(double v, double scale) {
    double q = v * scale;
    return q > 0.0 ? (unsigned long long)__double2ll_rn(q) : 0ull;
}

__device__ __forceinline__ unsigned long long shflXor64(unsigned long long v, int m) {
    unsigned lo = (unsigned)v, hi = (unsigned)(v >> 32);
    lo = __shfl_xor(lo, m);
    hi = __shfl_xor(hi, m);
    return ((unsigned long long)hi << 32) | lo;
}

// Per-job sums (ledger terms, counters) without contended global atomics: each lane
// keeps a register sum for the job it is currently on; when its job changes the sum
// goes into a per-workgroup LDS table (one row per job slot), and the table is added
// to global memory once per workgroup. A queue holding one job never touches the
// table until the end; a mixed queue pays LDS atomics, not global ones.
template <int K>
struct JobAcc {
    int job = -1;
    unsigned long long v[K] = {};
};

template <int K>
__device__ __forceinline__ void initJobTable(unsigned long long* tab, int nJobs) {
    for (int r = threadIdx.x; r < nJobs * K; r += blockDim.x) tab[r] = 0ull;
    __syncthreads();
}

template <int K>
__device__ __forceinline__ void spill(JobAcc<K>& a, unsigned long long* tab) {
    if (a.job >= 0)
        for (int k = 0; k < K; ++k)
            if (a.v[k]) atomicAdd(&tab[a.job * K + k], a.v[k]);
    for (int k = 0; k < K; ++k) a.v[k] = 0;
}

template <int K>
__device__ __forceinline__ void accumulate(JobAcc<K>& a, int job, unsigned long long* tab) {
    if (a.job != job) {
        spill(a, tab);
        a.job = job;
    }
}

// Called by every thread of the workgroup (converged).
template <int K, class Addr>
__device__ __forceinline__ void flushJobTable(JobAcc<K>& a, unsigned long long* tab, int nJobs, Addr addr) {
    spill(a, tab);
    __syncthreads();
    for (int r = threadIdx.x; r < nJobs * K; r += blockDim.x) {
        const unsigned long long v = tab[r];
        if (v) atomicAdd(addr(r / K, r % K), v);
    }
}

// Ledger and counter sums go to the job's row (blockIdx.x % kSegs): eight rows on
// eight cache lines, so the per-workgroup flushes of a launch spread over eight
// addresses instead of serialising on one.
struct AccAddr {
    const DevJob* jobs;
    int base;   // first index of this kernel's quantities in the row
    __device__ unsigned long long* operator()(int j, int k) const {
        return jobs[j].acc + (size_t)(blockIdx.x % kSegs) * kAccRow + base + k;
    }
};

// Logical index -> slot of a segmented queue. `pre` is the prefix of the segment
// counts (wave-uniform); the segment is the last one whose prefix is <= i, which
// skips empty segments.
struct SegMap {
    unsigned pre[kSegs];
    unsigned cap;
    __device__ __forceinline__ unsigned slot(unsigned i) const {
        unsigned base = 0, off = i;
#pragma unroll
        for (int g = 1; g < kSegs; ++g)
            if (i >= pre[g]) { base = (unsigned)g * cap; off = i - pre[g]; }
        return base + off;
    }
    // The same slot through a per-workgroup LDS table tab[g] = g * cap - pre[g]
    // (initSegTab): the segment is counted with compares against the wave-uniform
    // prefixes and the one per-lane lookup is an LDS read, so no uniform value has to
    // be copied into vector registers (slot's selects need them there).
    __device__ __forceinline__ unsigned slotTab(unsigned i, const unsigned* tab) const {
        unsigned g = 0;
#pragma unroll
        for (int k = 1; k < kSegs; ++k) g += (i >= pre[k]) ? 1u : 0u;
        return i + tab[g];
    }
};

// Fills the LDS table of SegMap::slotTab; the caller synchronises the workgroup before
// the first lookup.
__device__ __forceinline__ void initSegTab(const SegMap& m, unsigned* tab) {
#pragma unroll
    for (int g = 0; g < kSegs; ++g)
        if (threadIdx.x == (unsigned)g) tab[g] = (unsigned)g * m.cap - m.pre[g];
}

__device__ __forceinline__ unsigned loadSegMap(const unsigned (*cnt)[kLine], unsigned cap, SegMap& m) {
    unsigned acc = 0;
#pragma unroll
    for (int g = 0; g < kSegs; ++g) {
        m.pre[g] = acc;
        // a segment that overflowed (the trace then fails with GP_ERR_CAPACITY) still
        // maps only its own slots
        acc += min(cnt[g][0], cap);
    }
    m.cap = cap;
    return acc;
}

// Append to this workgroup's segment (blockIdx.x % kSegs) of a segmented queue;
// returns the slot, or ~0u (and raises the overflow flag) if the segment is full.
__device__ __forceinline__ unsigned segAppend(unsigned (*cnt)[kLine], unsigned cap, bool pred,
                                              unsigned* overflow) {
    const unsigned g = blockIdx.x % kSegs;
    const unsigned off = waveAppend(&cnt[g][0], pred);
    if (pred && off >= cap) {
        *overflow = 1u;
        return ~0u;
    }
    return g * cap + off;
}

// Workgroups are dispatched to the eight XCDs round-robin (workgroup b to XCD b % 8),
// and each XCD has its own L2. Renumbering b as (b % 8) * (G / 8) + b / 8 gives every
// XCD a contiguous eighth of each grid-stride round (and adjacent numbers to the
// workgroups of one XCD). Identity when the grid G is not a multiple of 8.
constexpr unsigned kXcds = 8;
__device__ __forceinline__ unsigned xcdBlock() {
    const unsigned G = gridDim.x, b = blockIdx.x;
    return (G % kXcds) ? b : (b % kXcds) * (G / kXcds) + b / kXcds;
}

// Ray-binning key of a ray: the Morton code of its origin's cell in the job's
// kBinCellsPerAxis^3 grid over the triangle-centroid box, then its direction octant.
__device__ __forceinline__ unsigned spread3(unsigned v) {
    return (v & 1u) | ((v & 2u) << 2) | ((v & 4u) << 4);
}

__device__ __forceinline__ unsigned binCell(float x, float lo, float scale) {
    // fmaxf returns 0 for a NaN cell coordinate
    const float f = fminf(fmaxf((x - lo) * scale, 0.0f), (float)(kBinCellsPerAxis - 1));
    return (unsigned)f;
}

__device__ __forceinline__ unsigned short binKey(const DevJob& J, F3 o, F3 d) {
    const unsigned m = spread3(binCell(o.x, J.binLo[0], J.binScale[0])) |
                       (spread3(binCell(o.y, J.binLo[1], J.binScale[1])) << 1) |
                       (spread3(binCell(o.z, J.binLo[2], J.binScale[2])) << 2);
    const unsigned oct = (d.x < 0.0f ? 1u : 0u) | (d.y < 0.0f ? 2u : 0u) | (d.z < 0.0f ? 4u : 0u);
    return (unsigned short)((m << 3) | oct);
}

// Camera connection record for a diffuse (or emitter) vertex at p with shading
// normal n (flat triangles: n is also the oriented geometric normal, so the
// reference's terminator factor and adjoint correction are exactly 1).
__device__ __forceinline__ bool makeConnection(const DevCamera& K, unsigned tagJobCam, F3 p, F3 n,
                                               const float* lam, const float* beta, const float* rho,
                                               int nUp, float4& co, float4& cd, float4& cx) {
    const F3 eye = load3(K.eye);
    F3 toCam = eye - p;
    float dist = sqrtf(dot(toCam, toCam));
    F3 wdir = toCam * (1.0f / dist);
    float cosSurf = dot(n, wdir);
    if (!(cosSurf > 0.0f)) return false;
    F3 dd = p - eye;
    float cz = dot(dd, load3(K.w));
    if (!(cz > 1e-9f)) return false;
    float ix = (dot(dd, load3(K.u)) / cz) / K.tanHalfX;
    float iy = (dot(dd, load3(K.v)) / cz) / K.tanHalfY;
    if (!(ix >= -1.0f && ix < 1.0f && iy >= -1.0f && iy < 1.0f)) return false;
    int px = (int)((ix * 0.5f + 0.5f) * (float)K.resX);
    int py = (int)((iy * 0.5f + 0.5f) * (float)K.resY);
    px = min(max(px, 0), K.resX - 1);
    py = min(max(py, 0), K.resY - 1);
    float dist2 = dot(dd, dd);
    float cosCam = cz / sqrtf(dist2);
    float denom = dist2 * (K.pixelArea * cosCam * cosCam * cosCam);
    float X = 0.f, Y = 0.f, Z = 0.f;
    for (int i = 0; i < nUp; ++i) {
        float c = beta[i] * (rho[i] * kInvPi) * cosSurf / denom;
        F3 w = cie(lam[i]);
        X += w.x * c;
        Y += w.y * c;
        Z += w.z * c;
    }
    F3 org = p + n * kRayEps;
    co = make_float4(org.x, org.y, org.z, (dist - 2.0f * kRayEps) - kRayEps);
    cd = make_float4(wdir.x, wdir.y, wdir.z, __int_as_float(py * K.resX + px));
    cx = make_float4(X, Y, Z, __int_as_float(tagJobCam | ((unsigned)nUp << 16)));
    return true;
}

// Connections of one vertex to every camera of its job, appended to the workgroup's
// segment of the connection queue. Called by all lanes of the wave (nCams = 0 on
// lanes without a vertex); the loop runs while any lane has a camera left. Returns
// the lane's number of connections.
__device__ __forceinline__ unsigned connectCameras(const Queues& q, Ctrl* c, const DevCamera* cams, int nCams,
                                                   unsigned job, F3 p, F3 n, const float* lam,
                                                   const float* beta, const float* rho, int nUp) {
    unsigned made = 0;
    for (int k = 0; __any(k < nCams); ++k) {
        float4 co, cd, cx;
        bool conn = false;
        if (k < nCams) conn = makeConnection(cams[k], connTag(job, 0u, (unsigned)k), p, n, lam, beta, rho, nUp, co, cd, cx);
        const unsigned ci = segAppend(c->connCount, q.connSegCap, conn, &c->overflow);
        if (conn && ci != ~0u) {
            q.co[ci] = co;
            q.cd[ci] = cd;
            q.cx[ci] = cx;
            ++made;
        }
    }
    return made;
}

// A vertex to connect to its job's cameras through the scene's flat mirrors, queued
// for k_mirror (weights w: the albedo at a surface vertex, 1 at the emitter). Called
// by all active lanes of the wave.
__device__ __forceinline__ void queueMirrorVertex(const Queues& q, Ctrl* c, bool pred, unsigned job, F3 p,
                                                  F3 n, const float* lam, const float* beta,
                                                  const float* w, int nUp) {
    if (!__any(pred)) return;
    const unsigned mi = segAppend(c->mirrorCount, q.mirrorSegCap, pred, &c->overflow);
    if (pred && mi != ~0u) {
        q.mv[0][mi] = make_float4(p.x, p.y, p.z, __int_as_float((int)(job | ((unsigned)nUp << 16))));
        q.mv[1][mi] = make_float4(n.x, n.y, n.z, 0.f);
        q.mv[2][mi] = make_float4(lam[0], lam[1], lam[2], lam[3]);
        q.mv[3][mi] = make_float4(beta[0] * w[0], beta[1] * w[1], beta[2] * w[2], beta[3] * w[3]);
    }
}

}  // namespace

// ------------------------------------------------------------------ schedule
// One workgroup of kMaxJobs threads: thread k owns position k of the priority order.
// Photons still to emit per job are scanned (inclusive prefix in LDS) and each job
// takes what fits in the free queue slots after the jobs ahead of it. The emitted
// photons are then placed after the survivors of each segment, filling segments in
// order; the segments' free room adds up to at least the free capacity, so the
// number emitted depends only on the total number of survivors.
__global__ void __launch_bounds__(kMaxJobs) k_schedule(Queues q, int cur) {
    __shared__ unsigned sScan[kMaxJobs];
    Ctrl* c = q.ctrl;
    const int k = threadIdx.x;
    const int n = c->nOrder;
    const unsigned it = c->iter;
    unsigned alive = 0;
#pragma unroll
    for (int g = 0; g < kSegs; ++g) alive += c->pathCount[cur][g][0];
    const unsigned freeSlots = c->capacity > alive ? c->capacity - alive : 0u;
    int j = -1;
    unsigned rem = 0;
    if (k < n) {
        j = c->order[k];
        const unsigned long long nx = c->nextPhoton[j], en = c->endPhoton[j];
        const unsigned long long r = (c->active[j] && en > nx) ? en - nx : 0ull;
        rem = r > (unsigned long long)freeSlots ? freeSlots : (unsigned)r;
    }
    sScan[k] = rem;
    __syncthreads();
    for (int off = 1; off < kMaxJobs; off <<= 1) {
        unsigned v = (k >= off) ? sScan[k - off] : 0u;
        __syncthreads();
        if (k >= off) sScan[k] = min(sScan[k] + v, 0x7fffffffu);
        __syncthreads();
    }
    const unsigned before = (k > 0) ? sScan[k - 1] : 0u;
    if (k < n) {
        const unsigned room = freeSlots > before ? freeSlots - before : 0u;
        const unsigned take = rem < room ? rem : room;
        const unsigned start = before < freeSlots ? before : freeSlots;
        c->emitPrefix[k] = start;
        c->emitJob[k] = j;
        c->emitStart[k] = c->nextPhoton[j];
        if (take) {
            c->nextPhoton[j] += take;
            if (c->nextPhoton[j] == c->endPhoton[j]) c->lastEmitIter[j] = (int)it;
        }
    }
    __syncthreads();
    if (k == 0) {
        const unsigned total = (n > 0) ? min(sScan[n - 1], freeSlots) : 0u;
        c->emitPrefix[n] = total;
        c->nEmit = total;
        unsigned left = total, placed = 0;
        for (int g = 0; g < kSegs; ++g) {
            const unsigned cnt = c->pathCount[cur][g][0];
            const unsigned room = c->segCap > cnt ? c->segCap - cnt : 0u;
            const unsigned e = left < room ? left : room;
            c->segBase[g] = cnt;
            c->segEmit[g] = placed;
            c->pathCount[cur][g][0] = cnt + e;
            c->pathCount[cur ^ 1][g][0] = 0;
            c->connCount[g][0] = 0;
            c->mirrorCount[g][0] = 0;
            placed += e;
            left -= e;
        }
        c->segEmit[kSegs] = placed;
        if (left) c->overflow = 1u;
        c->iter = it + 1;
    }
}

// ------------------------------------------------------------------ emit
// Launch bound: five waves per SIMD (96 registers, no spills; -2% k_emit time on c2).
__global__ void __launch_bounds__(256, 5) k_emit(Queues q, int cur) {
    Ctrl* c = q.ctrl;
    const unsigned nEmit = c->nEmit;
    const int nOrder = c->nOrder;
    const int nJobs = c->nJobs;
    // emission slot s goes to segment g (segEmit[g] <= s < segEmit[g + 1]), after
    // that segment's segBase[g] survivors
    SegMap em;
    unsigned emBase[kSegs];
#pragma unroll
    for (int g = 0; g < kSegs; ++g) {
        em.pre[g] = c->segEmit[g];
        emBase[g] = c->segBase[g];
    }
    em.cap = q.segCap;
    __shared__ unsigned long long eTab[kMaxJobs], pTab[kMaxJobs * 2];
    initJobTable<1>(eTab, nJobs);
    initJobTable<2>(pTab, nJobs);
    JobAcc<1> eacc;   // emitted
    JobAcc<2> pacc;   // photons, direct camera connections
    AccAddr ea{q.jobs, kAccEnergy};
    AccAddr pa{q.jobs, kAccPhotons};
    const unsigned stride = gridDim.x * blockDim.x;
    for (unsigned s = blockIdx.x * blockDim.x + threadIdx.x; s < nEmit; s += stride) {
        int k = 0;
        if (nOrder > 1) {
            int lo = 0, hi = nOrder;   // largest k with emitPrefix[k] <= s
            while (lo + 1 < hi) {
                int mid = (lo + hi) >> 1;
                if (c->emitPrefix[mid] <= s) lo = mid;
                else hi = mid;
            }
            k = lo;
        }
        const int j = c->emitJob[k];
        const DevJob& J = q.jobs[j];
        const unsigned long long idx = c->emitStart[k] + (s - c->emitPrefix[k]);
        Rng rng = seedPhoton(J.seedBase + idx, J.salt);
        int ei = 0;
        if (J.nEm > 1) {
            double u = rng.uniformD();
            int lo = 0, hi = J.nEm - 1;
            while (lo < hi) {
                int mid = (lo + hi) >> 1;
                if (J.emCdf[mid] < u) lo = mid + 1;
                else hi = mid;
            }
            ei = lo;
        }
        const DevEmitter& E = J.ems[ei];
        float u1 = rng.uniform();
        float u2 = rng.uniform();
        F3 org = load3(E.origin) + load3(E.u) * u1 + load3(E.v) * u2;
        F3 en = load3(E.normal);
        F3 dir = E.collimated ? load3(E.beam) : cosineDir(en, rng);
        const int C = J.heroC;
        double ub = rng.uniformD();
        float lam[kMaxHero] = {0.f, 0.f, 0.f, 0.f};
        float beta[kMaxHero] = {0.f, 0.f, 0.f, 0.f};
        double pdf0;
        lam[0] = (float)invertCdf(E.cdf, E.guide, ub, pdf0);
        const bool dead = !(pdf0 > 0.0);
        for (int i = 1; i < C; ++i) {
            double uu = ub + (double)i / (double)C;
            if (uu >= 1.0) uu -= 1.0;
            double pd;
            lam[i] = (float)invertCdf(E.cdf, E.guide, uu, pd);
        }
        const float b = J.basePower / (float)C;
        for (int i = 0; i < C; ++i) beta[i] = b;
        accumulate(eacc, j, eTab);
        accumulate(pacc, j, pTab);
        pacc.v[0] += 1;
        if (!dead) eacc.v[0] += toFixed((double)b * (double)C, J.energyScale);
        {
            const float one[kMaxHero] = {1.f, 1.f, 1.f, 1.f};
            pacc.v[1] += connectCameras(q, c, J.cams, dead ? 0 : J.nCams, (unsigned)j, org, en, lam, beta, one, C);
            queueMirrorVertex(q, c, !dead && J.nMirrors > 0, (unsigned)j, org, en, lam, beta, one, C);
        }
        F3 o = org + dir * kRayEps;
        unsigned slot = 0;
        {
            unsigned off = s, g0 = 0;
#pragma unroll
            for (int g = 1; g < kSegs; ++g)
                if (s >= em.pre[g]) { g0 = (unsigned)g; off = s - em.pre[g]; }
            unsigned b = emBase[0];
#pragma unroll
            for (int g = 1; g < kSegs; ++g)
                if (g0 == (unsigned)g) b = emBase[g];
            slot = g0 * q.segCap + b + off;
        }
        q.ro[cur][slot] = make_float4(o.x, o.y, o.z, 0.f);
        q.rd[cur][slot] = make_float4(dir.x, dir.y, dir.z, 0.f);
        q.lam[cur][slot] = make_float4(lam[0], lam[1], lam[2], lam[3]);
        q.beta[cur][slot] = make_float4(beta[0], beta[1], beta[2], beta[3]);
        unsigned meta = packMeta((unsigned)j, 0u, (unsigned)C) | (dead ? 0x80000000u : 0u);
        q.st[cur][slot] = make_uint4((unsigned)rng.state, (unsigned)(rng.state >> 32), (unsigned)idx, meta);
        if (q.binned) q.key[cur][slot] = dead ? (unsigned short)0 : binKey(J, o, dir);
    }
    flushJobTable(eacc, eTab, nJobs, ea);
    flushJobTable(pacc, pTab, nJobs, pa);
}

// ------------------------------------------------------------------ extend
// One wave per workgroup: a workgroup's LDS (its lanes' stacks) is released as soon as
// its own rays are done, instead of waiting for the slowest of four waves.
constexpr int kExtendThreads = 64;
// STATS: also count the traversal statistics (node visits, triangle tests) per job.
// BIN: read the binned queue (ray binning) instead of queue `cur`.
// The traversal waits on dependent node loads, so resident waves are what hides the
// latency. Segments are counted by k_shade (same paths), so outside the statistics
// variant the kernel's LDS is the traversal stack alone (SD * 512 bytes per one-wave
// workgroup), and the launch bound asks for six waves per SIMD (at most 80 registers;
// the compiler applies it where the LDS allows six: SD = 8, two spilled registers in
// the spilling binned variant). Measured on c2, closest-hit time per 2^28 photons:
// 148.3 ms at SD 16 and five waves, 144.2 at SD 8 and five, 140.6 at SD 8 and six;
// seven or eight waves spill inside the loop (187 / 202 ms).
template <int SD, bool SPILL, bool STATS, bool BIN>
__global__ void __launch_bounds__(kExtendThreads, 6) k_extend(Queues q, int cur) {
    constexpr int NA = STATS ? 2 : 1;   // node visits, triangle tests
    __shared__ int2 sStk[SD * kExtendThreads];
    __shared__ unsigned segTab[kSegs];
    SegMap sm;
    const unsigned n = loadSegMap(q.ctrl->pathCount[cur], q.segCap, sm);
    if constexpr (!BIN) {
        initSegTab(sm, segTab);
        __syncthreads();
    }
    const unsigned stride = gridDim.x * blockDim.x;
    const unsigned gtid = blockIdx.x * blockDim.x + threadIdx.x;
    Stack<SD, true, SPILL, kExtendThreads> stk{sStk + threadIdx.x, nullptr,
                                               SPILL ? q.spill + (size_t)gtid * q.spillDepth : nullptr, 0};
    const int nJobs = q.ctrl->nJobs;
    __shared__ unsigned long long sTab[STATS ? kMaxJobs * NA : 1];
    if constexpr (STATS) initJobTable<NA>(sTab, nJobs);
    JobAcc<NA> sacc;
    AccAddr sa{q.jobs, kAccNodeVisits};
    for (unsigned li = gtid; li < n; li += stride) {
        const unsigned i = BIN ? li : sm.slotTab(li, segTab);
        const uint4 s4 = BIN ? bitcast4(q.pb[5 * (size_t)i + 4]) : q.st[cur][i];
        int best = -1;
        float t = INFINITY;
        if (!(s4.w & 0x80000000u)) {
            const int j = (int)(s4.w & 0xffffu);
            const DevJob& J = q.jobs[j];
            const float4 o = BIN ? q.pb[5 * (size_t)i] : q.ro[cur][i];
            const float4 d = BIN ? q.pb[5 * (size_t)i + 1] : q.rd[cur][i];
            RayPre r = prepareRay(xyz(o), xyz(d));
            float tmax = INFINITY;
            TravCount<STATS> tc;
            best = traceClosest<SD, SPILL>(J.nodes, J.tris, r, kRayEps, tmax, stk, tc);
            t = tmax;
            if constexpr (STATS) {
                accumulate(sacc, j, sTab);
                sacc.v[0] += tc.nodes;
                sacc.v[1] += tc.tri;
            }
        }
        q.hit[i] = make_float2(t, __int_as_float(best));
    }
    if constexpr (STATS) flushJobTable(sacc, sTab, nJobs, sa);
}

// ------------------------------------------------------------------ shade
// BIN: read the binned queue (ray binning) instead of queue `cur` and store the
// survivors' bin keys. Launch bound: five waves per SIMD (96 registers, no spills;
// -3% k_shade time on c2 binned).
template <bool BIN>
__global__ void __launch_bounds__(256, 5) k_shade(Queues q, int cur) {
    const int nxt = cur ^ 1;
    Ctrl* ctl = q.ctrl;
    SegMap sm;
    const unsigned n = loadSegMap(ctl->pathCount[cur], q.segCap, sm);
    __shared__ unsigned segTab[kSegs];
    if constexpr (!BIN) initSegTab(sm, segTab);   // initJobTable's barrier covers it
    const unsigned stride = gridDim.x * blockDim.x;
    const int nJobs = ctl->nJobs;
    __shared__ unsigned long long eTab[kMaxJobs * 4];
    initJobTable<4>(eTab, nJobs);
    // absorbed, escaped, residual -> ledger entries 1, 3, 4; the closest-hit queries of
    // the paths (every live path of the queue was traced by k_extend) -> segments
    JobAcc<4> eacc;
    AccAddr ea{q.jobs, kAccEnergy};
    auto eaddr = [&](int j, int k) {
        return ea(j, k == 0 ? 1 : (k == 1 ? 3 : (k == 2 ? 4 : kAccSegments - kAccEnergy)));
    };
    for (unsigned li = blockIdx.x * blockDim.x + threadIdx.x; li < n; li += stride) {
        const unsigned i = BIN ? li : sm.slotTab(li, segTab);
        const float4* rec = q.pb + 5 * (size_t)i;   // binned record (BIN only)
        const uint4 s4 = BIN ? bitcast4(rec[4]) : q.st[cur][i];
        bool alive = false;
        unsigned short bkey = 0;
        // diffuse vertex to connect to the cameras after shading
        int camN = 0;
        bool mirrorsToo = false;
        const DevCamera* cams = nullptr;
        unsigned cjob = 0;
        int cUp = 0;
        F3 cp{0.f, 0.f, 0.f}, cn{0.f, 0.f, 0.f};
        float cbeta[kMaxHero] = {0.f, 0.f, 0.f, 0.f}, crho[kMaxHero] = {0.f, 0.f, 0.f, 0.f};
        F3 no, nd;
        float lam[kMaxHero], beta[kMaxHero];
        unsigned meta = s4.w;
        Rng rng;
        rng.state = (uint64_t)s4.x | ((uint64_t)s4.y << 32);
        if (!(meta & 0x80000000u)) {
            const int j = (int)(meta & 0xffffu);
            const int bounce = (int)((meta >> 16) & 0xffu);
            int nUp = (int)((meta >> 24) & 7u);
            const DevJob& J = q.jobs[j];
            accumulate(eacc, j, eTab);
            eacc.v[3] += 1;
            const float4 l4 = BIN ? rec[2] : q.lam[cur][i], b4 = BIN ? rec[3] : q.beta[cur][i];
            lam[0] = l4.x; lam[1] = l4.y; lam[2] = l4.z; lam[3] = l4.w;
            beta[0] = b4.x; beta[1] = b4.y; beta[2] = b4.z; beta[3] = b4.w;
            auto activeSum = [&]() {
                double s = 0.0;
                for (int k = 0; k < nUp; ++k) s += (double)beta[k];
                return s;
            };
            const float2 h = q.hit[i];
            const int tri = __float_as_int(h.y);
            if (tri < 0) {
                eacc.v[1] += toFixed(activeSum(), J.energyScale);
            } else {
                const float4 o4 = BIN ? rec[0] : q.ro[cur][i], d4 = BIN ? rec[1] : q.rd[cur][i];
                const F3 d = xyz(d4);
                const F3 p = xyz(o4) + d * h.x;
                const float4 tn = J.trin[tri];
                const F3 gn = xyz(tn);
                const DevMaterial& M = J.mats[__float_as_int(tn.w)];
                const F3 n = (dot(d, gn) < 0.0f) ? gn : -gn;
                rng.inc = photonInc(J.seedBase + s4.z, J.salt);
                const bool roulette = bounce >= J.rrDepth;
                alive = true;
                if (M.type == 1) {
                    // Dielectric: de-hero, then Fresnel reflect or Snell refract (exterior
                    // index 1). The side is read off the geometric normal (entering when the
                    // ray runs against it), as the reference does.
                    if (nUp > 1) {
                        beta[0] *= (float)J.heroC;
                        nUp = 1;
                    }
                    const float eta_g = specEval(M.ior, lam[0]);
                    const bool entering = dot(d, gn) < 0.0f;
                    const F3 nl = entering ? gn : -gn;
                    const float n1 = entering ? 1.0f : eta_g;
                    const float n2 = entering ? eta_g : 1.0f;
                    const float eta = n1 / n2;
                    const float cosI = -dot(d, nl);
                    const float sin2t = eta * eta * (1.0f - cosI * cosI);
                    F3 out;
                    bool refracted = false;
                    if (sin2t > 1.0f) {
                        out = reflectDir(d, nl);
                    } else {
                        const float cosT = sqrtf(1.0f - sin2t);
                        const float rs = (n1 * cosI - n2 * cosT) / (n1 * cosI + n2 * cosT);
                        const float rp = (n1 * cosT - n2 * cosI) / (n1 * cosT + n2 * cosI);
                        const float R = 0.5f * (rs * rs + rp * rp);
                        if (rng.uniform() < R) {
                            out = reflectDir(d, nl);
                        } else {
                            out = d * eta + nl * (eta * cosI - cosT);
                            refracted = true;
                        }
                    }
                    out = normalize(out);
                    if (M.roughness > 1e-3f) {
                        F3 pert = glossyDir(out, M.roughness, rng);
                        bool ok = refracted ? (dot(pert, nl) < 0.0f) : (dot(pert, nl) > 0.0f);
                        if (ok) out = pert;
                    }
                    no = p + out * kRayEps;
                    nd = out;
                } else {
                    float c[kMaxHero] = {0.f, 0.f, 0.f, 0.f};
                    float q0 = 0.0f;
                    for (int k = 0; k < nUp; ++k) {
                        float v = specEval(M.reflect, lam[k]);
                        c[k] = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
                        q0 = (k == 0 || c[k] > q0) ? c[k] : q0;
                    }
                    const bool diffuse = (M.type == 0);
                    if (diffuse) {
                        camN = J.nCams;
                        mirrorsToo = J.nMirrors > 0;
                        cams = J.cams;
                        cjob = (unsigned)j;
                        cUp = nUp;
                        cp = p;
                        cn = n;
                        for (int k = 0; k < kMaxHero; ++k) { cbeta[k] = beta[k]; crho[k] = c[k]; }
                    }
                    if (roulette) {
                        if (rng.uniform() >= q0) {
                            eacc.v[0] += toFixed(activeSum(), J.energyScale);
                            alive = false;
                        } else {
                            double ab = 0.0;
                            for (int k = 0; k < nUp; ++k) {
                                float w = c[k] / q0;
                                ab += (double)beta[k] * (double)(1.0f - w);
                                beta[k] *= w;
                            }
                            eacc.v[0] += toFixed(ab, J.energyScale);
                        }
                    } else {
                        if (!(q0 > 0.0f)) {
                            eacc.v[0] += toFixed(activeSum(), J.energyScale);
                            alive = false;
                        } else {
                            double ab = 0.0;
                            for (int k = 0; k < nUp; ++k) {
                                ab += (double)beta[k] * (double)(1.0f - c[k]);
                                beta[k] *= c[k];
                            }
                            eacc.v[0] += toFixed(ab, J.energyScale);
                        }
                    }
                    if (alive) {
                        F3 wo;
                        if (diffuse) {
                            wo = cosineDir(n, rng);
                        } else if (M.type == 2) {
                            wo = reflectDir(d, n);
                        } else {
                            wo = glossyDir(reflectDir(d, n), M.roughness, rng);
                            if (dot(wo, n) <= 0.0f) {
                                eacc.v[0] += toFixed(activeSum(), J.energyScale);
                                alive = false;
                            }
                        }
                        no = p + n * kRayEps;
                        nd = wo;
                    }
                }
                if (alive && bounce + 1 >= J.maxDepth) {
                    eacc.v[2] += toFixed(activeSum(), J.energyScale);
                    alive = false;
                }
                if constexpr (BIN)
                    if (alive) bkey = binKey(J, no, nd);
                meta = packMeta((unsigned)j, (unsigned)(bounce + 1), (unsigned)nUp);
            }
        }
        connectCameras(q, ctl, cams, camN, cjob, cp, cn, lam, cbeta, crho, cUp);
        queueMirrorVertex(q, ctl, mirrorsToo, cjob, cp, cn, lam, cbeta, crho, cUp);
        const unsigned pi = segAppend(ctl->pathCount[nxt], q.segCap, alive, &ctl->overflow);
        if (alive && pi != ~0u) {
            q.ro[nxt][pi] = make_float4(no.x, no.y, no.z, 0.f);
            q.rd[nxt][pi] = make_float4(nd.x, nd.y, nd.z, 0.f);
            q.lam[nxt][pi] = make_float4(lam[0], lam[1], lam[2], lam[3]);
            q.beta[nxt][pi] = make_float4(beta[0], beta[1], beta[2], beta[3]);
            q.st[nxt][pi] = make_uint4((unsigned)rng.state, (unsigned)(rng.state >> 32), s4.z, meta);
            if constexpr (BIN) q.key[nxt][pi] = bkey;
        }
    }
    flushJobTable(eacc, eTab, nJobs, eaddr);
}

// ------------------------------------------------------------------ connect
// Splat table: a per-workgroup open-addressing table in LDS keyed by (job, camera,
// pixel).
// Unoccluded connections add their integer X, Y, Z and hit count into it with LDS
// 64-bit atomics. When a connection finds no slot within kSplatProbe probes, the
// workgroup flushes the whole table to the film and starts a fresh one, so every
// film update goes through the table. Two properties of the memory system shape
// this:
//   - film atomics are limited by requests, not operations: four lanes adding the
//     four words of one pixel in one instruction cost one request, four lanes each
//     adding to their own pixel cost four (measured ~91 vs ~23 G atomics/s on
//     MI355X), so the flush hands (slot, word) pairs to consecutive lanes;
//   - atomics on one address serialise (~12 ns each), and pixels that many
//     connections hit (the emitter's direct view, caustic foci) would queue on a few
//     addresses; the table sums them first.
// Integer sums are order-independent, so the film is bit-identical to splatting
// each connection directly.
constexpr int kSplatSlots = 256;
constexpr int kSplatProbe = 8;
constexpr unsigned long long kSplatEmpty = ~0ull;

// Slot of `key` (claimed if absent) within maxProbe linear probes, or -1.
__device__ __forceinline__ int splatSlot(unsigned long long* sKey, unsigned long long key, int maxProbe) {
    const unsigned h = (unsigned)((key * 0x9E3779B97F4A7C15ull) >> 56);
    for (int p = 0; p < maxProbe; ++p) {
        const int s = (int)((h + (unsigned)p) & (kSplatSlots - 1));
        unsigned long long cur = __hip_atomic_load(&sKey[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (cur == kSplatEmpty) cur = atomicCAS(&sKey[s], kSplatEmpty, key);
        if (cur == kSplatEmpty || cur == key) return s;
    }
    return -1;
}

__device__ __forceinline__ void splatAdd(unsigned long long* sVal, int s, unsigned long long X,
                                         unsigned long long Y, unsigned long long Z, unsigned long long H) {
    if (X) atomicAdd(&sVal[4 * s + 0], X);
    if (Y) atomicAdd(&sVal[4 * s + 1], Y);
    if (Z) atomicAdd(&sVal[4 * s + 2], Z);
    atomicAdd(&sVal[4 * s + 3], H);
}

// Called by every thread of the workgroup after a barrier: adds every occupied
// slot into its film, lane t taking word t % 4 of slot t / 4 (+ 64 per pass).
__device__ __forceinline__ void flushSplats(const DevJob* jobs, const unsigned long long* sKey,
                                            const unsigned long long* sVal) {
    const int w = threadIdx.x & 3;
    for (int s = threadIdx.x >> 2; s < kSplatSlots; s += blockDim.x >> 2) {
        const unsigned long long key = sKey[s];
        if (key == kSplatEmpty) continue;
        const unsigned long long v = sVal[4 * s + w];
        if (v) {
            const unsigned jc = (unsigned)(key >> 32);
            unsigned long long* f = (unsigned long long*)jobs[jc / kMaxCams].cams[jc % kMaxCams].film;
            atomicAdd(f + 4 * (size_t)(key & 0xffffffffull) + w, v);
        }
    }
}

__device__ __forceinline__ void clearSplats(unsigned long long* sKey, unsigned long long* sVal) {
    for (int k = threadIdx.x; k < kSplatSlots; k += blockDim.x) {
        sKey[k] = kSplatEmpty;
        sVal[4 * k + 0] = 0; sVal[4 * k + 1] = 0; sVal[4 * k + 2] = 0; sVal[4 * k + 3] = 0;
    }
}

// Launch bound: six waves per SIMD where the LDS allows it (the default 8-entry stack:
// 22.8 KB per workgroup; 80 registers, five spilled): k_connect 26.7 -> 25.6 ms on c2
// against five waves with a 16-entry stack.
template <int SD, bool SPILL, bool STATS>
__global__ void __launch_bounds__(kTravThreads, 6) k_connect(Queues q) {
    constexpr int NA = STATS ? 4 : 2;   // connections, splats[, node visits, triangle tests]
    __shared__ int sStk[SD * kTravThreads];
    __shared__ unsigned long long sKey[kSplatSlots];
    __shared__ unsigned long long sVal[kSplatSlots * 4];
    clearSplats(sKey, sVal);
    SegMap sm;
    const unsigned n = loadSegMap(q.ctrl->connCount, q.connSegCap, sm);
    __shared__ unsigned segTab[kSegs];
    initSegTab(sm, segTab);   // initJobTable's barrier covers it
    const unsigned stride = gridDim.x * blockDim.x;
    const unsigned gtid = blockIdx.x * blockDim.x + threadIdx.x;
    Stack<SD, false, SPILL> stk{nullptr, sStk + threadIdx.x,
                                SPILL ? q.spill + (size_t)gtid * q.spillDepth : nullptr, 0};
    const int nJobs = q.ctrl->nJobs;
    __shared__ unsigned long long cTab[kMaxJobs * NA];
    initJobTable<NA>(cTab, nJobs);   // (its barrier also orders clearSplats)
    JobAcc<NA> cacc;
    AccAddr ca{q.jobs, kAccConnections};
    // Rounds are workgroup-uniform so the table can be flushed between them.
    for (unsigned base = blockIdx.x * blockDim.x; base < n; base += stride) {
        const unsigned li = base + threadIdx.x;
        bool pending = false;
        unsigned long long key = 0, X = 0, Y = 0, Z = 0, H = 0;
        if (li < n) {
            const unsigned i = sm.slotTab(li, segTab);
            const float4 co = q.co[i], cd = q.cd[i], cx = q.cx[i];
            const unsigned tag = (unsigned)__float_as_int(cx.w);
            const int j = (int)(tag & 0xffffu);
            const int cam = (int)((tag >> 19) & 15u);
            const DevJob& J = q.jobs[j];
            accumulate(cacc, j, cTab);
            cacc.v[0] += 1;
            RayPre r = prepareRay(xyz(co), xyz(cd));
            TravCount<STATS> tc;
            const bool occ = traceAny<SD, SPILL>(J.nodes, J.tris, r, kRayEps, co.w, stk, tc);
            if constexpr (STATS) {
                cacc.v[2] += tc.nodes;
                cacc.v[3] += tc.tri;
            }
            if (!occ) {
                cacc.v[1] += 1;
                const DevCamera& K = J.cams[cam];
                const unsigned pix = (unsigned)__float_as_int(cd.w);
                X = min(toFixed((double)cx.x, K.filmScale), K.filmQmax);
                Y = min(toFixed((double)cx.y, K.filmScale), K.filmQmax);
                Z = min(toFixed((double)cx.z, K.filmScale), K.filmQmax);
                H = (unsigned long long)((tag >> 16) & 7u);
                key = ((unsigned long long)(j * kMaxCams + cam) << 32) | pix;
                const int s = splatSlot(sKey, key, kSplatProbe);
                if (s >= 0) splatAdd(sVal, s, X, Y, Z, H);
                else pending = true;
            }
        }
        if (__syncthreads_or(pending)) {
            flushSplats(q.jobs, sKey, sVal);
            __syncthreads();
            clearSplats(sKey, sVal);
            __syncthreads();
            // Only the pending keys, at most one per thread, go into the empty table
            // before the barrier below lets any thread start the next round, so a full
            // scan always finds its key or a free slot. (Without the barrier, waves
            // racing ahead could fill the table first.) A failed scan still falls
            // back to the film itself rather than index the table with -1.
            if (pending) {
                const int s = splatSlot(sKey, key, kSplatSlots);
                if (s >= 0) {
                    splatAdd(sVal, s, X, Y, Z, H);
                } else {
                    const unsigned jc = (unsigned)(key >> 32);
                    unsigned long long* f =
                        (unsigned long long*)q.jobs[jc / kMaxCams].cams[jc % kMaxCams].film +
                        4 * (size_t)(key & 0xffffffffull);
                    if (X) atomicAdd(f + 0, X);
                    if (Y) atomicAdd(f + 1, Y);
                    if (Z) atomicAdd(f + 2, Z);
                    atomicAdd(f + 3, H);
                }
            }
            __syncthreads();
        }
    }
    flushJobTable(cacc, cTab, nJobs, ca);   // ends with a barrier-ordered read of cTab
    __syncthreads();
    flushSplats(q.jobs, sKey, sVal);
}

// ------------------------------------------------------------------ mirrors
// Camera connections through flat mirrors (the reference's connectSpecularPlane,
// src/render.h:1334-1376, GPU twin src/render_cuda.cu:5004-5052). Mirroring the eye
// in the plane turns vertex -> mirror -> eye into the straight segment from the
// vertex to the eye's image, so the mirror point R is a plane intersection and the
// specular geometry factor is 1 / D^2 over the unfolded length D. The geometry is
// solved in double; the two legs are traced in float: the light side must be clear
// (any hit), and the eye's closest hit must be a mirror triangle at R (within
// 1e-4 (1 + |R - eye|)), whose reflectance tints the splat. One thread per queued
// vertex walks the scene's planes (outer) and the job's cameras (inner); splats go
// straight to the film with 64-bit atomics (they are few next to the direct
// connections, and integer sums keep the film independent of their order).
struct MirrorSolve {
    double wP[3];       // vertex -> mirror point, unit
    double dP;          // vertex -> mirror point distance
    double wE[3];       // eye -> mirror point, unit
    double dE;          // eye -> mirror point distance
    double weight;      // (cos at the vertex / pi) / (D^2 * pixel solid angle)
    unsigned pix;       // film pixel of the mirror point
};

__device__ __forceinline__ bool solveMirror(const DevMirrorPlane& P, const DevCamera& K, const double p[3],
                                            const double np[3], MirrorSolve& o) {
    const double e[3] = {(double)K.eye[0], (double)K.eye[1], (double)K.eye[2]};
    const double pd = P.n[0] * p[0] + P.n[1] * p[1] + P.n[2] * p[2];
    const double se = (P.n[0] * e[0] + P.n[1] * e[1] + P.n[2] * e[2]) - P.d;
    const double sp = pd - P.d;
    if (!(se * sp > 0.0)) return false;   // eye and vertex on the same face
    double w[3];
    for (int a = 0; a < 3; ++a) w[a] = (e[a] - P.n[a] * (2.0 * se)) - p[a];
    const double D = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    if (D < 1e-9) return false;
    const double inv = 1.0 / D;
    for (int a = 0; a < 3; ++a) w[a] *= inv;
    const double cosSurf = np[0] * w[0] + np[1] * w[1] + np[2] * w[2];
    if (!(cosSurf > 0.0)) return false;
    const double denom = P.n[0] * w[0] + P.n[1] * w[1] + P.n[2] * w[2];
    if (fabs(denom) < 1e-12) return false;
    const double dP = (P.d - pd) / denom;
    if (dP <= 1e-7 || dP >= D - 1e-7) return false;
    double R[3];
    for (int a = 0; a < 3; ++a) {
        R[a] = p[a] + w[a] * dP;
        if (R[a] < P.lo[a] || R[a] > P.hi[a]) return false;
    }
    // pinhole projection of R (Camera::project / pixelSolidAngle)
    const double d[3] = {R[0] - e[0], R[1] - e[1], R[2] - e[2]};
    const double cz = d[0] * K.w[0] + d[1] * K.w[1] + d[2] * K.w[2];
    if (!(cz > 1e-9)) return false;
    const double ix = ((d[0] * K.u[0] + d[1] * K.u[1] + d[2] * K.u[2]) / cz) / (double)K.tanHalfX;
    const double iy = ((d[0] * K.v[0] + d[1] * K.v[1] + d[2] * K.v[2]) / cz) / (double)K.tanHalfY;
    if (!(ix >= -1.0 && ix < 1.0 && iy >= -1.0 && iy < 1.0)) return false;
    const int px = min(max((int)((ix * 0.5 + 0.5) * (double)K.resX), 0), K.resX - 1);
    const int py = min(max((int)((iy * 0.5 + 0.5) * (double)K.resY), 0), K.resY - 1);
    const double dist2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    const double cosCam = cz / sqrt(dist2);
    const double omega = (double)K.pixelArea * cosCam * cosCam * cosCam;
    if (!(omega > 0.0)) return false;
    const double dE = D - dP;
    for (int a = 0; a < 3; ++a) {
        o.wP[a] = w[a];
        o.wE[a] = (R[a] - e[a]) * (1.0 / dE);
    }
    o.dP = dP;
    o.dE = dE;
    o.weight = (cosSurf / 3.141592653589793) / (D * D * omega);
    o.pix = (unsigned)(py * K.resX + px);
    return true;
}

template <int SD, bool SPILL>
__global__ void __launch_bounds__(kTravThreads) k_mirror(Queues q) {
    // Separate LDS arrays for the two traversals: lanes of a workgroup run the light leg
    // (any hit, int entries) and the eye leg (closest hit, int2 entries) at the same
    // time, and two views of one array would overlap across lanes (lane t's int entry k
    // and lane u's int2 entry r share a word whenever k*256 + t = 2*(r*256 + u) + 0/1).
    __shared__ int2 sStk[SD * kTravThreads];
    __shared__ int sStkAny[SD * kTravThreads];
    __shared__ unsigned long long cTab[kMaxJobs * 2];
    SegMap sm;
    const unsigned n = loadSegMap(q.ctrl->mirrorCount, q.mirrorSegCap, sm);
    const int nJobs = q.ctrl->nJobs;
    initJobTable<2>(cTab, nJobs);
    JobAcc<2> cacc;   // ray queries of the connections, splats (legs not in the statistics)
    AccAddr ca{q.jobs, kAccConnections};
    TravCount<false> tc;
    const unsigned stride = gridDim.x * blockDim.x;
    const unsigned gtid = blockIdx.x * blockDim.x + threadIdx.x;
    int2* spill = SPILL ? q.spill + (size_t)gtid * q.spillDepth : nullptr;
    for (unsigned li = gtid; li < n; li += stride) {
        const unsigned i = sm.slot(li);
        const float4 a = q.mv[0][i], b = q.mv[1][i], l4 = q.mv[2][i], w4 = q.mv[3][i];
        const unsigned tag = (unsigned)__float_as_int(a.w);
        const int j = (int)(tag & 0xffffu);
        const int nUp = (int)((tag >> 16) & 7u);
        const DevJob& J = q.jobs[j];
        accumulate(cacc, j, cTab);
        const float lam[kMaxHero] = {l4.x, l4.y, l4.z, l4.w};
        const float bw[kMaxHero] = {w4.x, w4.y, w4.z, w4.w};
        const double p[3] = {(double)a.x, (double)a.y, (double)a.z};
        const double np[3] = {(double)b.x, (double)b.y, (double)b.z};
        for (int m = 0; m < J.nMirrors; ++m) {
            const DevMirrorPlane P = J.mirrors[m];
            for (int k = 0; k < J.nCams; ++k) {
                const DevCamera& K = J.cams[k];
                MirrorSolve g;
                if (!solveMirror(P, K, p, np, g)) continue;
                cacc.v[0] += 1;
                const F3 wl{(float)g.wP[0], (float)g.wP[1], (float)g.wP[2]};
                const F3 ol{(float)(p[0] + g.wP[0] * 1e-6), (float)(p[1] + g.wP[1] * 1e-6),
                            (float)(p[2] + g.wP[2] * 1e-6)};
                Stack<SD, false, SPILL> as{nullptr, sStkAny + threadIdx.x, spill, 0};
                if (traceAny<SD, SPILL>(J.nodes, J.tris, prepareRay(ol, wl), kRayEps,
                                        (float)((g.dP - 2e-6) - 1e-6), as, tc))
                    continue;
                cacc.v[0] += 1;
                Stack<SD, true, SPILL> cs{sStk + threadIdx.x, nullptr, spill, 0};
                float th = INFINITY;
                const int tri = traceClosest<SD, SPILL>(
                    J.nodes, J.tris, prepareRay(load3(K.eye), F3{(float)g.wE[0], (float)g.wE[1], (float)g.wE[2]}),
                    kRayEps, th, cs, tc);
                if (tri < 0 || fabs((double)th - g.dE) > 1e-4 * (1.0 + g.dE)) continue;
                const DevMaterial& M = J.mats[__float_as_int(J.trin[tri].w)];
                if (M.type != 2) continue;   // only Mirror is a planar mirror here
                cacc.v[1] += 1;
                double X = 0.0, Y = 0.0, Z = 0.0;
                unsigned long long H = 0;
                for (int c = 0; c < nUp; ++c) {
                    float r = specEval(M.reflect, lam[c]);
                    r = r < 0.0f ? 0.0f : (r > 1.0f ? 1.0f : r);
                    const double contrib = (double)bw[c] * g.weight * (double)r;
                    if (!(contrib > 0.0)) continue;
                    const F3 cw = cie(lam[c]);
                    X += (double)cw.x * contrib;
                    Y += (double)cw.y * contrib;
                    Z += (double)cw.z * contrib;
                    ++H;
                }
                if (H) {
                    unsigned long long* f = (unsigned long long*)K.film + 4 * (size_t)g.pix;
                    const unsigned long long qx = min(toFixed(X, K.filmScale), K.filmQmax);
                    const unsigned long long qy = min(toFixed(Y, K.filmScale), K.filmQmax);
                    const unsigned long long qz = min(toFixed(Z, K.filmScale), K.filmQmax);
                    if (qx) atomicAdd(f + 0, qx);
                    if (qy) atomicAdd(f + 1, qy);
                    if (qz) atomicAdd(f + 2, qz);
                    atomicAdd(f + 3, H);
                }
            }
        }
    }
    flushJobTable(cacc, cTab, nJobs, ca);
}

// ------------------------------------------------------------------ ray binning
// Counting sort of queue `cur` into the binned queue by the 12-bit keys the emission
// and shading kernels stored (binKey). Column c (workgroup xcdBlock() of C = q.binCols) owns the
// contiguous logical range [n c / C, n (c + 1) / C). The order inside a (column,
// bin) run is that of LDS atomics; nothing downstream depends on it (films, ledgers
// and counters are integer sums).
__device__ __forceinline__ void binRange(unsigned n, unsigned col, unsigned cols, unsigned& lo, unsigned& hi) {
    lo = (unsigned)(((unsigned long long)n * col) / cols);
    hi = (unsigned)(((unsigned long long)n * (col + 1)) / cols);
}

// Histogram of column xcdBlock() into binHist[col][*].
__global__ void __launch_bounds__(kBinThreads) k_bin_count(Queues q, int cur) {
    __shared__ unsigned h[kBins];
    for (int b = threadIdx.x; b < kBins; b += blockDim.x) h[b] = 0u;
    __syncthreads();
    SegMap sm;
    const unsigned n = loadSegMap(q.ctrl->pathCount[cur], q.segCap, sm);
    const unsigned col = xcdBlock();
    unsigned lo, hi;
    binRange(n, col, (unsigned)q.binCols, lo, hi);
    const unsigned short* __restrict__ keys = q.key[cur];
    // keys are masked: paths queued before binning was switched on carry no key (both
    // passes read the same words, so any value still gives a permutation)
    for (unsigned li = lo + threadIdx.x; li < hi; li += blockDim.x) atomicAdd(&h[keys[sm.slot(li)] & (kBins - 1)], 1u);
    __syncthreads();
    unsigned* out = q.binHist + (size_t)col * kBins;
    for (int b = threadIdx.x; b < kBins; b += blockDim.x) out[b] = h[b];
}

// Per bin: exclusive prefix of the columns inside each of the kBinChunks column
// chunks (in place), the chunks' exclusive prefix (binChunk[chunk][bin]) and the
// bin's total (binTotal[bin]). A workgroup takes 32 bins x kBinChunks chunks at a
// time, grid-stride over the kBins / 32 groups.
__global__ void __launch_bounds__(256) k_bin_chunks(Queues q) {
    const unsigned per = (unsigned)q.binCols / kBinChunks;
    __shared__ unsigned part[kBinChunks][32];
    const unsigned bl = threadIdx.x & 31u, ch = threadIdx.x >> 5;
    for (unsigned grp = blockIdx.x; grp < kBins / 32; grp += gridDim.x) {
        const unsigned bin = grp * 32u + bl;
        unsigned run = 0;
        // eight columns loaded before any is rewritten, then the rest one at a time
        unsigned c = ch * per;
        const unsigned cEnd = (ch + 1) * per;
        for (; c + 8 <= cEnd; c += 8) {
            unsigned v[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) v[k] = q.binHist[(size_t)(c + k) * kBins + bin];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                q.binHist[(size_t)(c + k) * kBins + bin] = run;
                run += v[k];
            }
        }
        for (; c < cEnd; ++c) {
            unsigned* e = q.binHist + (size_t)c * kBins + bin;
            const unsigned v = *e;
            *e = run;
            run += v;
        }
        part[ch][bl] = run;
        __syncthreads();
        if (ch == 0) {
            unsigned acc = 0;
            for (unsigned k = 0; k < kBinChunks; ++k) {
                q.binChunk[(size_t)k * kBins + bin] = acc;
                acc += part[k][bl];
            }
            q.binTotal[bin] = acc;
        }
        __syncthreads();
    }
}

// Exclusive scan of the kBins totals in place (one workgroup of 1024 threads).
constexpr int kBinScanThreads = kBins < 1024 ? kBins : 1024;
__global__ void __launch_bounds__(kBinScanThreads) k_bin_scan(Queues q) {
    constexpr int kPer = kBins / kBinScanThreads;
    __shared__ unsigned sum[kBinScanThreads];
    const unsigned t = threadIdx.x;
    unsigned v[kPer], acc = 0;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        v[k] = q.binTotal[t * kPer + k];
        acc += v[k];
    }
    sum[t] = acc;
    __syncthreads();
    for (unsigned off = 1; off < kBinScanThreads; off <<= 1) {
        const unsigned add = t >= off ? sum[t - off] : 0u;
        __syncthreads();
        sum[t] += add;
        __syncthreads();
    }
    unsigned base = t ? sum[t - 1] : 0u;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        q.binTotal[t * kPer + k] = base;
        base += v[k];
    }
}

__device__ __forceinline__ void scatterWave(float4* __restrict__ out, float4* st, unsigned lane, unsigned dst,
                                            const float4& o, const float4& d, const float4& l, const float4& b,
                                            const uint4& s) {
    // Lane L stages its record at st[5 L ..]; then round k stores float4 64 k + L,
    // element e of record r = (64 k + L) / 5, at the destination of lane r: a store
    // instruction writes about 13 whole records (80 contiguous bytes each) instead of
    // one 16-byte piece of 64 records.
    st[5 * lane + 0] = o;
    st[5 * lane + 1] = d;
    st[5 * lane + 2] = l;
    st[5 * lane + 3] = b;
    st[5 * lane + 4] = bitcast4f(s);
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (unsigned k = 0; k < 5; ++k) {
        const unsigned f = 64u * k + lane, r = f / 5u, e = f - 5u * r;
        const unsigned t = (unsigned)__shfl((int)dst, (int)r);
        const float4 v = st[f];
        if (t != ~0u) out[5 * (size_t)t + e] = v;
    }
    __builtin_amdgcn_wave_barrier();
}

// Moves column xcdBlock()'s paths to their bins' slots of the binned queue.
__global__ void __launch_bounds__(kBinThreads) k_bin_scatter(Queues q, int cur) {
    __shared__ unsigned slot[kBins];
    __shared__ float4 stage[kBinThreads / 64][5 * 64];
    const unsigned col = xcdBlock();
    const unsigned ch = col / ((unsigned)q.binCols / kBinChunks);
    const unsigned* colHist = q.binHist + (size_t)col * kBins;
    const unsigned* chunk = q.binChunk + (size_t)ch * kBins;
    for (int b = threadIdx.x; b < kBins; b += blockDim.x) slot[b] = q.binTotal[b] + chunk[b] + colHist[b];
    __syncthreads();
    SegMap sm;
    const unsigned n = loadSegMap(q.ctrl->pathCount[cur], q.segCap, sm);
    const unsigned room = kSegs * q.segCap;
    unsigned lo, hi;
    binRange(n, col, (unsigned)q.binCols, lo, hi);
    const unsigned short* __restrict__ keys = q.key[cur];
    const uint4* __restrict__ sst = q.st[cur];
    const float4* __restrict__ sro = q.ro[cur];
    const float4* __restrict__ srd = q.rd[cur];
    const float4* __restrict__ slam = q.lam[cur];
    const float4* __restrict__ sbeta = q.beta[cur];
    float4* __restrict__ out = q.pb;
    const unsigned lane = threadIdx.x & 63u;
    float4* st = stage[threadIdx.x >> 6];
    // two paths per lane and step, both loaded before either is stored; the loop is
    // wave-uniform (a lane past the end loads its wave's first path and stores nothing)
    for (unsigned wb = lo + (threadIdx.x & ~63u); wb < hi; wb += 2 * blockDim.x) {
        const unsigned li = wb + lane, lj = li + blockDim.x;
        const bool v0 = li < hi, v1 = lj < hi;
        const unsigned i0 = sm.slot(v0 ? li : wb), i1 = sm.slot(v1 ? lj : wb);
        const unsigned k0 = keys[i0] & (kBins - 1), k1 = keys[i1] & (kBins - 1);
        const uint4 s0 = sst[i0], s1 = sst[i1];
        const float4 o0 = sro[i0], o1 = sro[i1], d0 = srd[i0], d1 = srd[i1];
        const float4 l0 = slam[i0], l1 = slam[i1], b0 = sbeta[i0], b1 = sbeta[i1];
        unsigned t0 = v0 ? atomicAdd(&slot[k0], 1u) : ~0u;
        unsigned t1 = v1 ? atomicAdd(&slot[k1], 1u) : ~0u;
        if ((v0 && t0 >= room) || (v1 && t1 >= room)) {   // impossible: the two passes read the same keys
            q.ctrl->overflow = 1u;
            t0 = t0 >= room ? ~0u : t0;
            t1 = t1 >= room ? ~0u : t1;
        }
        scatterWave(out, st, lane, t0, o0, d0, l0, b0, s0);
        scatterWave(out, st, lane, t1, o1, d1, l1, b1, s1);
    }
}

// ------------------------------------------------------------------ probes
template <int SD, bool SPILL>
__global__ void __launch_bounds__(kTravThreads) k_probe_closest(const float4* nodes, const float4* tris,
                                                               const float* rays, long long n, float tmin,
                                                               int* prim, float* tout, int2* spill,
                                                               int spillDepth) {
    __shared__ int2 sStk[SD * kTravThreads];
    const long long stride = (long long)gridDim.x * blockDim.x;
    const long long gtid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    Stack<SD, true, SPILL> stk{sStk + threadIdx.x, nullptr,
                               SPILL ? spill + (size_t)gtid * spillDepth : nullptr, 0};
    for (long long i = gtid; i < n; i += stride) {
        const float* rp = rays + 6 * i;
        RayPre r = prepareRay(F3{rp[0], rp[1], rp[2]}, F3{rp[3], rp[4], rp[5]});
        float tmax = INFINITY;
        TravCount<false> tc;
        int best = traceClosest<SD, SPILL>(nodes, tris, r, tmin, tmax, stk, tc);
        prim[i] = best >= 0 ? __float_as_int(tris[3 * best].w) : -1;
        tout[i] = best >= 0 ? tmax : INFINITY;
    }
}

template <int SD, bool SPILL>
__global__ void __launch_bounds__(kTravThreads) k_probe_any(const float4* nodes, const float4* tris,
                                                           const float* rays, const float* tmax, long long n,
                                                           float tmin, unsigned char* occ, int2* spill,
                                                           int spillDepth) {
    __shared__ int sStk[SD * kTravThreads];
    const long long stride = (long long)gridDim.x * blockDim.x;
    const long long gtid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    Stack<SD, false, SPILL> stk{nullptr, sStk + threadIdx.x,
                                SPILL ? spill + (size_t)gtid * spillDepth : nullptr, 0};
    for (long long i = gtid; i < n; i += stride) {
        const float* rp = rays + 6 * i;
        RayPre r = prepareRay(F3{rp[0], rp[1], rp[2]}, F3{rp[3], rp[4], rp[5]});
        TravCount<false> tc;
        occ[i] = traceAny<SD, SPILL>(nodes, tris, r, tmin, tmax[i], stk, tc) ? 1 : 0;
    }
}

// ------------------------------------------------------------------ job table
// Installs a job into slot a.slot: its record, photon range, emission state and zeroed
// ledger rows. A job without photons is complete once maxDepth iterations have run
// from now, like a job whose last photon was emitted in the current iteration.
__global__ void __launch_bounds__(256) k_install_job(Queues q, JobInstall a) {
    Ctrl* c = q.ctrl;
    const int slot = a.slot;
    if (threadIdx.x == 0) {
        const_cast<DevJob*>(q.jobs)[slot] = a.job;
        const_cast<DevCamera*>(a.job.cams)[0] = a.cam;
        c->nextPhoton[slot] = a.next;
        c->endPhoton[slot] = a.end;
        c->lastEmitIter[slot] = (a.end > a.next) ? -1 : (int)c->iter;
        c->active[slot] = 1;
    }
    for (int k = threadIdx.x; k < kSegs * kAccRow; k += blockDim.x) a.job.acc[k] = 0ull;
}

__global__ void __launch_bounds__(kMaxJobs) k_set_order(Queues q, OrderUpdate a) {
    Ctrl* c = q.ctrl;
    c->order[threadIdx.x] = a.order[threadIdx.x];
    if (threadIdx.x == 0) {
        c->nOrder = a.nOrder;
        c->nJobs = a.nJobs;
        if (a.clearSlot >= 0) c->active[a.clearSlot] = 0;
    }
}

__global__ void k_fill64(unsigned long long* p, size_t n, unsigned long long v) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

// ------------------------------------------------------------------ launchers
void launchSchedule(const Queues& q, int cur, hipStream_t s) {
    hipLaunchKernelGGL(k_schedule, dim3(1), dim3(kMaxJobs), 0, s, q, cur);
}
void launchEmit(const Queues& q, int cur, const LaunchShape& ls, hipStream_t s) {
    hipLaunchKernelGGL(k_emit, dim3(ls.blocksShade), dim3(256), 0, s, q, cur);
}
void launchShade(const Queues& q, int cur, const LaunchShape& ls, hipStream_t s) {
    if (q.binned) hipLaunchKernelGGL(k_shade<true>, dim3(ls.blocksShade), dim3(256), 0, s, q, cur);
    else hipLaunchKernelGGL(k_shade<false>, dim3(ls.blocksShade), dim3(256), 0, s, q, cur);
}
// Kernel instances by LDS stack depth and spill use.
#define GP_BY_STACK(LDS, SPILL, CALL) \
    do {                                   \
        if (SPILL) {                       \
            switch (LDS) {                 \
                case 8: CALL(8, true); break;   \
                case 24: CALL(24, true); break; \
                case 32: CALL(32, true); break; \
                default: CALL(16, true); break; \
            }                              \
        } else {                           \
            switch (LDS) {                 \
                case 8: CALL(8, false); break;   \
                case 24: CALL(24, false); break; \
                case 32: CALL(32, false); break; \
                default: CALL(16, false); break; \
            }                              \
        }                                  \
    } while (0)

// The closest-hit kernel also has a 12-entry stack (6 KB of LDS per one-wave
// workgroup: six waves per SIMD still fit): c2 closest-hit time 139.0 ms against 140.9
// at 8 entries, 176.7 against 191.6 in queue order.
#define GP_BY_STACK_EXT(LDS, SPILL, CALL)  \
    do {                                   \
        if ((LDS) == 12) {                 \
            if (SPILL) CALL(12, true);     \
            else CALL(12, false);          \
        } else {                           \
            GP_BY_STACK(LDS, SPILL, CALL); \
        }                                  \
    } while (0)

void launchExtend(const Queues& q, int cur, const LaunchShape& ls, hipStream_t s) {
    const dim3 grid(ls.blocksExtend * kExtendGridScale * (kTravThreads / kExtendThreads));
#define GP_CALL(SD, SP) hipLaunchKernelGGL((k_extend<SD, SP, false, false>), grid, dim3(kExtendThreads), 0, s, q, cur)
#define GP_CALLS(SD, SP) hipLaunchKernelGGL((k_extend<SD, SP, true, false>), grid, dim3(kExtendThreads), 0, s, q, cur)
#define GP_CALLB(SD, SP) hipLaunchKernelGGL((k_extend<SD, SP, false, true>), grid, dim3(kExtendThreads), 0, s, q, cur)
#define GP_CALLSB(SD, SP) hipLaunchKernelGGL((k_extend<SD, SP, true, true>), grid, dim3(kExtendThreads), 0, s, q, cur)
    if (q.binned) {
        if (q.stats) GP_BY_STACK_EXT(q.lds, q.useSpill, GP_CALLSB);
        else GP_BY_STACK_EXT(q.lds, q.useSpill, GP_CALLB);
    } else {
        if (q.stats) GP_BY_STACK_EXT(q.lds, q.useSpill, GP_CALLS);
        else GP_BY_STACK_EXT(q.lds, q.useSpill, GP_CALL);
    }
#undef GP_CALLSB
#undef GP_CALLB
#undef GP_CALLS
#undef GP_CALL
}
void launchBin(const Queues& q, int cur, hipStream_t s) {
    hipLaunchKernelGGL(k_bin_count, dim3(q.binCols), dim3(kBinThreads), 0, s, q, cur);
    hipLaunchKernelGGL(k_bin_chunks, dim3(std::min(kBins / 32, q.binCols)), dim3(kBinChunks * 32), 0, s, q);
    hipLaunchKernelGGL(k_bin_scan, dim3(1), dim3(kBinScanThreads), 0, s, q);
    hipLaunchKernelGGL(k_bin_scatter, dim3(q.binCols), dim3(kBinThreads), 0, s, q, cur);
}
void launchConnect(const Queues& q, const LaunchShape& ls, hipStream_t s) {
#define GP_CALL(SD, SP) hipLaunchKernelGGL((k_connect<SD, SP, false>), dim3(ls.blocksExtend), dim3(kTravThreads), 0, s, q)
#define GP_CALLS(SD, SP) hipLaunchKernelGGL((k_connect<SD, SP, true>), dim3(ls.blocksExtend), dim3(kTravThreads), 0, s, q)
    if (q.stats) GP_BY_STACK(q.connLds, q.connUseSpill, GP_CALLS);
    else GP_BY_STACK(q.connLds, q.connUseSpill, GP_CALL);
#undef GP_CALLS
#undef GP_CALL
}
void launchMirror(const Queues& q, const LaunchShape& ls, hipStream_t s) {
#define GP_CALL(SD, SP) hipLaunchKernelGGL((k_mirror<SD, SP>), dim3(ls.blocksExtend), dim3(kTravThreads), 0, s, q)
    GP_BY_STACK(q.lds, q.useSpill, GP_CALL);
#undef GP_CALL
}
void launchProbeClosest(const float4* nodes, const float4* tris, const float* rays, long long n,
                        float tmin, int* prim, float* t, int2* spill, int spillDepth, int lds,
                        int blocks, int threads, hipStream_t s) {
    (void)threads;
#define GP_CALL(SD, SP) hipLaunchKernelGGL((k_probe_closest<SD, SP>), dim3(blocks), dim3(kTravThreads), 0, s, nodes, tris, rays, n, tmin, prim, t, spill, spillDepth)
    GP_BY_STACK(lds, spillDepth > 0, GP_CALL);
#undef GP_CALL
}
void launchProbeAny(const float4* nodes, const float4* tris, const float* rays, const float* tmax,
                    long long n, float tmin, unsigned char* occ, int2* spill, int spillDepth,
                    int lds, int blocks, int threads, hipStream_t s) {
    (void)threads;
#define GP_CALL(SD, SP) hipLaunchKernelGGL((k_probe_any<SD, SP>), dim3(blocks), dim3(kTravThreads), 0, s, nodes, tris, rays, tmax, n, tmin, occ, spill, spillDepth)
    GP_BY_STACK(lds, spillDepth > 0, GP_CALL);
#undef GP_CALL
}
#undef GP_BY_STACK
void launchInstallJob(const Queues& q, const JobInstall& a, hipStream_t s) {
    hipLaunchKernelGGL(k_install_job, dim3(1), dim3(256), 0, s, q, a);
}
void launchSetOrder(const Queues& q, const OrderUpdate& a, hipStream_t s) {
    hipLaunchKernelGGL(k_set_order, dim3(1), dim3(kMaxJobs), 0, s, q, a);
}
void launchFill64(unsigned long long* p, size_t n, unsigned long long v, hipStream_t s) {
    if (n == 0) return;
    size_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_fill64, dim3((unsigned)blocks), dim3(256), 0, s, p, n, v);
}

}  // namespace gp
