// Host runtime: device discovery, camera basis, film scaling, wavefront engine.
#include "gp_runtime.hpp"

#include "gp_bvh_device.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace gp {

namespace {
thread_local std::string g_err;
}

gp_status fail(gp_status s, const std::string& msg) {
    g_err = msg;
    return s;
}
const char* lastError() { return g_err.c_str(); }

// ------------------------------------------------------------------ device
DeviceState& device() {
    static DeviceState st = [] {
        DeviceState d;
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
            d.why = "no HIP device";
            return d;
        }
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, 0) != hipSuccess) {
            d.why = "hipGetDeviceProperties failed";
            return d;
        }
        d.name = prop.name;
        d.arch = prop.gcnArchName;
        d.cus = prop.multiProcessorCount;
        d.ldsPerBlock = prop.sharedMemPerBlock;
        if (d.arch.rfind("gfx950", 0) != 0) {
            d.why = "device arch " + d.arch + " is not gfx950";
            return d;
        }
        if (hipSetDevice(0) != hipSuccess || hipStreamCreate(&d.stream) != hipSuccess ||
            hipStreamCreate(&d.copyStream) != hipSuccess) {
            d.why = "stream creation failed";
            return d;
        }
        d.ok = true;
        d.status = GP_OK;
        return d;
    }();
    return st;
}

gp_status requireDevice() {
    DeviceState& d = device();
    if (!d.ok) return fail(GP_ERR_NO_DEVICE, d.why);
    return GP_OK;
}

// ------------------------------------------------------------------ camera
static void sub3(const double* a, const double* b, double* o) { for (int i = 0; i < 3; ++i) o[i] = a[i] - b[i]; }
static double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static void cross3(const double* a, const double* b, double* o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
static void unit3(double* a) {
    double l = std::sqrt(dot3(a, a));
    for (int i = 0; i < 3; ++i) a[i] = a[i] / l;
}

bool cameraBasis(const gp_camera_desc& c, CamBasis& o) {
    if (c.res_x <= 0 || c.res_y <= 0 || !(c.fov_y_deg > 0.0 && c.fov_y_deg < 180.0)) return false;
    for (int i = 0; i < 3; ++i) o.eye[i] = c.eye[i];
    sub3(c.target, c.eye, o.w);
    unit3(o.w);
    cross3(o.w, c.up, o.u);
    unit3(o.u);
    cross3(o.u, o.w, o.v);
    o.tanHalfY = std::tan(0.5 * c.fov_y_deg * 3.141592653589793 / 180.0);
    o.tanHalfX = o.tanHalfY * (double)c.res_x / (double)c.res_y;
    o.pixelArea = (4.0 * o.tanHalfX * o.tanHalfY) / ((double)c.res_x * (double)c.res_y);
    o.resX = c.res_x;
    o.resY = c.res_y;
    for (int i = 0; i < 3; ++i)
        if (!std::isfinite(o.u[i]) || !std::isfinite(o.v[i]) || !std::isfinite(o.w[i])) return false;
    return true;
}

// ------------------------------------------------------------------ distance bound
// Closest point on a triangle (Ericson, Real-Time Collision Detection 5.1.5).
static double pointTriDist2(const double* p, const double* a, const double* b, const double* c) {
    double ab[3], ac[3], ap[3];
    sub3(b, a, ab); sub3(c, a, ac); sub3(p, a, ap);
    double d1 = dot3(ab, ap), d2 = dot3(ac, ap);
    double q[3];
    auto dist2 = [&](const double* x) { double d[3]; sub3(p, x, d); return dot3(d, d); };
    if (d1 <= 0 && d2 <= 0) return dist2(a);
    double bp[3]; sub3(p, b, bp);
    double d3 = dot3(ab, bp), d4 = dot3(ac, bp);
    if (d3 >= 0 && d4 <= d3) return dist2(b);
    double vc = d1 * d4 - d3 * d2;
    if (vc <= 0 && d1 >= 0 && d3 <= 0) {
        double v = d1 / (d1 - d3);
        for (int i = 0; i < 3; ++i) q[i] = a[i] + v * ab[i];
        return dist2(q);
    }
    double cp[3]; sub3(p, c, cp);
    double d5 = dot3(ab, cp), d6 = dot3(ac, cp);
    if (d6 >= 0 && d5 <= d6) return dist2(c);
    double vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0 && d6 <= 0) {
        double w = d2 / (d2 - d6);
        for (int i = 0; i < 3; ++i) q[i] = a[i] + w * ac[i];
        return dist2(q);
    }
    double va = d3 * d6 - d5 * d4;
    if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
        double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        for (int i = 0; i < 3; ++i) q[i] = b[i] + w * (c[i] - b[i]);
        return dist2(q);
    }
    double denom = 1.0 / (va + vb + vc);
    double v = vb * denom, w = vc * denom;
    for (int i = 0; i < 3; ++i) q[i] = a[i] + ab[i] * v + ac[i] * w;
    return dist2(q);
}
static double pointBoxDist2(const double* p, const float* lo, const float* hi) {
    double s = 0;
    for (int a = 0; a < 3; ++a) {
        double d = 0;
        if (p[a] < lo[a]) d = lo[a] - p[a];
        else if (p[a] > hi[a]) d = p[a] - hi[a];
        s += d * d;
    }
    return s;
}

double sceneMinDistance(const gp_scene& s, const double p[3]) {
    double best = INFINITY;
    if (s.pairs.empty()) {
        // tree built on the device: brute force there, once per eye position
        const std::array<double, 3> key{p[0], p[1], p[2]};
        auto it = s.eyeDist2.find(key);
        if (it != s.eyeDist2.end()) {
            best = it->second;
        } else {
            double d2 = 0.0;   // on failure the bound degrades to the 1e-6 floor (smaller scale)
            if (deviceMinDistance2(s.dTris, s.nTris, p, d2) == GP_OK) s.eyeDist2[key] = d2;
            best = d2;
        }
    }
    // Depth-first over the pair nodes, nearer child first, pruned by box distance.
    std::vector<int> stack;
    if (!s.pairs.empty()) stack.push_back(0);
    auto visitLeaf = [&](int code) {
        int c = ~code;
        int first = c >> 3, count = (c & 7) + 1;
        for (int k = 0; k < count; ++k) {
            const float* t = &s.leafTris[9 * (size_t)(first + k)];
            double a[3] = {t[0], t[1], t[2]}, b[3] = {t[3], t[4], t[5]}, cc[3] = {t[6], t[7], t[8]};
            best = std::min(best, pointTriDist2(p, a, b, cc));
        }
    };
    while (!stack.empty()) {
        int node = stack.back();
        stack.pop_back();
        const float* q = &s.pairs[16 * (size_t)node];
        int links[2];
        std::memcpy(&links[0], &q[3], 4);
        std::memcpy(&links[1], &q[7], 4);
        double dl = pointBoxDist2(p, q + 0, q + 4), dr = pointBoxDist2(p, q + 8, q + 12);
        int order[2] = {0, 1};
        double dd[2] = {dl, dr};
        if (dr < dl) { order[0] = 1; order[1] = 0; }
        for (int k = 1; k >= 0; --k) {
            int side = order[k];
            if (dd[side] >= best) continue;
            if (links[side] >= 0) stack.push_back(links[side]);
            else visitLeaf(links[side]);
        }
    }
    for (const auto& e : s.emitQuads) {
        const double* o = &e[0];
        const double* u = &e[3];
        const double* v = &e[6];
        double b[3], c[3], d[3];
        for (int i = 0; i < 3; ++i) { b[i] = o[i] + u[i]; c[i] = o[i] + u[i] + v[i]; d[i] = o[i] + v[i]; }
        best = std::min(best, pointTriDist2(p, o, b, c));
        best = std::min(best, pointTriDist2(p, o, c, d));
    }
    return std::sqrt(best);
}

// ------------------------------------------------------------------ film scale
// The largest value one camera connection can add to one film channel is
//   B = P * (1/pi) * max_lambda(cie) / (dmin^2 * pixelArea * cosMin^3)
// with P the photon power (sum of the live per-wavelength weights never exceeds it),
// dmin the eye's distance to the scene and cosMin the cosine at the film corner.
// Splats are rounded to the nearest integer unit and saturated at qmax = 2^K, and a
// photon splats at most (maxDepth + 1) times, so a film sized for N photons holds at
// most N * (maxDepth + 1) * 2^K <= 2^62 in any channel: no int64 overflow.
gp_status prepareFilm(gp_film* film, const gp_scene* s, const CamBasis& cb,
                      const gp_trace_params& p) {
    // every vertex splats at most once directly and once through each mirror plane
    const double perVertex = 1.0 + (double)s->nMirrors;
    const double splats = (double)p.n_photons * (double)(p.max_depth + 1) * perVertex;
    if (film->scale > 0.0) {
        if (film->splatUsed + splats > film->splatBudget * (1.0 + 1e-12))
            return fail(GP_ERR_CAPACITY, "film photon budget exceeded (create it with a larger max_photons)");
        film->splatUsed += splats;
        return GP_OK;
    }
    const double dmin = std::max(sceneMinDistance(*s, cb.eye), 1e-6);
    const double cosMin = 1.0 / std::sqrt(1.0 + cb.tanHalfX * cb.tanHalfX + cb.tanHalfY * cb.tanHalfY);
    const double cieMax = 1.8;   // max over [360, 830] nm of the three fitted CIE curves is < 1.8
    const double B = s->basePower * cieMax / 3.141592653589793 /
                     (dmin * dmin * cb.pixelArea * cosMin * cosMin * cosMin);
    const double budget =
        std::max((double)film->maxPhotons, (double)p.n_photons) * (double)(p.max_depth + 1) * perVertex;
    double K = std::floor(62.0 - std::log2(budget) - 1e-9);
    K = std::min(K, 52.0);
    if (K < 8.0) return fail(GP_ERR_CAPACITY, "photon budget too large for a 64-bit film");
    film->qmax = 1ull << (int)K;
    film->scale = std::ldexp(1.0, (int)K) / B;
    film->splatBudget = budget;
    film->splatUsed = splats;
    return GP_OK;
}

gp_status fillJob(const gp_scene* s, const gp_camera_desc* cams, gp_film* const* films, int nCams,
                  const gp_trace_params& p, int slot, DevJob& J, DevCamera* camsOut) {
    if (nCams < 1 || nCams > kMaxCams) return fail(GP_ERR_INVALID, "1..16 cameras per job");
    CamBasis cbs[kMaxCams];
    for (int c = 0; c < nCams; ++c) {
        if (!cameraBasis(cams[c], cbs[c])) return fail(GP_ERR_INVALID, "degenerate camera");
        const gp_film* film = films[c];
        if (!film) return fail(GP_ERR_INVALID, "null film");
        if (cams[c].res_x != film->rx || cams[c].res_y != film->ry)
            return fail(GP_ERR_INVALID, "camera resolution differs from the film's");
        for (int d = 0; d < c; ++d)
            if (films[d] == film) return fail(GP_ERR_INVALID, "two cameras share one film");
    }
    // Films take the job's photons into their budgets; if one refuses, the films
    // before it give them back, so a refused job leaves every film as it was.
    struct Budget { double scale, budget, used; unsigned long long qmax; };
    Budget saved[kMaxCams];
    for (int c = 0; c < nCams; ++c) {
        gp_film* film = films[c];
        saved[c] = {film->scale, film->splatBudget, film->splatUsed, film->qmax};
        gp_status st = prepareFilm(film, s, cbs[c], p);
        if (st != GP_OK) {
            for (int d = 0; d <= c; ++d) {
                films[d]->scale = saved[d].scale;
                films[d]->splatBudget = saved[d].budget;
                films[d]->splatUsed = saved[d].used;
                films[d]->qmax = saved[d].qmax;
            }
            return st;
        }
    }
    for (int c = 0; c < nCams; ++c) {
        const CamBasis& cb = cbs[c];
        gp_film* film = films[c];
        DevCamera& K = camsOut[c];
        std::memset(&K, 0, sizeof K);
        for (int i = 0; i < 3; ++i) {
            K.eye[i] = (float)cb.eye[i];
            K.u[i] = (float)cb.u[i];
            K.v[i] = (float)cb.v[i];
            K.w[i] = (float)cb.w[i];
        }
        K.tanHalfX = (float)cb.tanHalfX;
        K.tanHalfY = (float)cb.tanHalfY;
        K.pixelArea = (float)cb.pixelArea;
        K.resX = cb.resX;
        K.resY = cb.resY;
        K.film = (long long*)film->d;
        K.filmScale = film->scale;
        K.filmQmax = film->qmax;
    }
    Engine& E = engine();
    std::memset(&J, 0, sizeof J);
    J.nodes = s->dNodes;
    J.tris = s->dTris;
    J.trin = s->dTrin;
    J.mats = s->dMats;
    J.ems = s->dEms;
    J.emCdf = s->dEmCdf;
    J.nEm = s->nEm;
    J.basePower = (float)s->basePower;
    J.cams = E.dJobCams + (size_t)slot * kMaxCams;
    J.nCams = nCams;
    J.nMirrors = s->nMirrors;
    J.mirrors = s->dMirrors;
    J.acc = E.dJobAcc + (size_t)slot * kSegs * kAccRow;
    const double ebits = std::floor(60.0 - std::log2(std::max(1.0, (double)p.n_photons)) - 2.0);
    J.energyScale = std::ldexp(1.0, (int)std::max(8.0, ebits)) / std::max(s->basePower, 1e-300);
    J.seedBase = p.seed_base;
    J.salt = p.salt;
    J.heroC = p.hero_c;
    J.maxDepth = p.max_depth;
    J.rrDepth = p.rr_depth;
    for (int a = 0; a < 3; ++a) {
        J.binLo[a] = s->binLo[a];
        J.binScale[a] = s->binScale[a];
    }
    return GP_OK;
}

// ------------------------------------------------------------------ engine
gp_tuning defaultTuning() {
    gp_tuning t;
    // Measured on MI355X (c1, 2^26 photons): 2^23 slots and a 64-waves-per-CU grid
    // trace 2.57 G segments/s against 2.35 G with 2^22 and 16 (c2: 1.84 vs 1.41).
    t.queue_capacity = 1 << 23;
    t.block_extend = 256;
    t.block_shade = 256;
    t.grid_waves_per_cu = 64;
    t.iters_per_graph = 8;
    t.use_graph = 1;
    // 12 LDS stack entries (deeper entries spill to global memory): the closest-hit
    // kernel then holds 6 KB of LDS per one-wave workgroup and six waves per SIMD
    // (c2: 139.0 ms of closest-hit time per 2^28 photons, 140.9 with 8 entries, 148.3
    // with 16 entries and five waves); k_connect keeps 8.
    t.lds_stack = 12;
    t.profile = 0;
    t.ray_sort = 2;   // automatic: binned for scenes of kBinAutoTris to kBinAutoMaxTris triangles
    return t;
}

Engine& engine() {
    static Engine e;
    return e;
}

void Engine::releaseQueues() {
    if (graph) { hipGraphExecDestroy(graph); graph = nullptr; }
    for (int i = 0; i < nbufs; ++i) hipFree(bufs[i]);
    nbufs = 0;
    if (dCtrl) hipFree(dCtrl);
    if (dJobs) hipFree(dJobs);
    if (hCtrl) hipHostFree(hCtrl);
    if (hJobs) hipHostFree(hJobs);
    if (dJobAcc) hipFree(dJobAcc);
    if (dJobCams) hipFree(dJobCams);
    if (hJobCams) hipHostFree(hJobCams);
    dCtrl = nullptr; dJobs = nullptr; hCtrl = nullptr; hJobs = nullptr;
    dJobAcc = nullptr; dJobCams = nullptr; hJobCams = nullptr;
    int2* sp = q.spill;
    q = Queues{};
    q.spill = sp;
    q.spillDepth = spillDepth;
    capacity = 0;
    segCap = 0;
}

void Engine::releaseSpill() {
    if (graph) { hipGraphExecDestroy(graph); graph = nullptr; }
    if (q.spill) hipFree(q.spill);
    q.spill = nullptr;
    q.spillDepth = 0;
    spillDepth = 0;
    spillThreads = 0;
}

void Engine::release() {
    releaseQueues();
    releaseSpill();
}

// Largest number of items that the workgroups b with b % kSegs == g visit, over g,
// when `blocks` workgroups of 256 threads grid-stride over cap items (the loop
// shape of k_emit and k_shade). A segment receives at most one path and one
// connection per item its workgroups visit, so this bounds every segment's fill.
// The items sum to cap over the segments, so kSegs * bound >= cap.
unsigned segmentBound(unsigned cap, unsigned blocks) {
    uint64_t per[kSegs] = {};
    const uint64_t stride = (uint64_t)blocks * 256;
    for (unsigned b = 0; b < blocks; ++b) {
        uint64_t items = 0;
        for (uint64_t start = (uint64_t)b * 256; start < cap; start += stride)
            items += std::min<uint64_t>(256, cap - start);
        per[b % kSegs] += items;
    }
    uint64_t m = 0;
    for (int g = 0; g < kSegs; ++g) m = std::max(m, per[g]);
    return (unsigned)((m + 63) & ~uint64_t(63));
}

void Engine::initCtrl() {
    std::memset(hCtrl, 0, sizeof(Ctrl));
    hCtrl->capacity = capacity;
    hCtrl->segCap = segCap;
    hCtrl->connSegCap = 2 * segCap * (unsigned)camCap;
    hCtrl->mirrorSegCap = 2 * segCap;
    for (int k = 0; k < kMaxJobs; ++k) hCtrl->lastEmitIter[k] = -1;
}

gp_status Engine::resetJobAcc(int slot) {
    launchFill64(dJobAcc + (size_t)slot * kSegs * kAccRow, (size_t)kSegs * kAccRow, 0ull, device().stream);
    GP_HIP(hipGetLastError());
    return GP_OK;
}

void sumAccRows(const unsigned long long* rows, unsigned long long ledger[5], unsigned long long* counters) {
    for (int k = 0; k < 5; ++k) ledger[k] = 0;
    for (int k = 0; k < kAccCounters; ++k) counters[k] = 0;
    for (int g = 0; g < kSegs; ++g) {
        for (int k = 0; k < 5; ++k) ledger[k] += rows[g * kAccRow + kAccEnergy + k];
        for (int k = 0; k < kAccCounters; ++k) counters[k] += rows[g * kAccRow + kAccPhotons + k];
    }
}

void addCounters(const unsigned long long* cv, gp_trace_stats& st) {
    st.photons += cv[kAccPhotons - kAccPhotons];
    st.direct_connections += cv[kAccDirect - kAccPhotons];
    st.segments += cv[kAccSegments - kAccPhotons];
    st.node_visits += cv[kAccNodeVisits - kAccPhotons];
    st.tri_tests += cv[kAccTriTests - kAccPhotons];
    st.connections += cv[kAccConnections - kAccPhotons];
    st.splats += cv[kAccSplats - kAccPhotons];
    st.shadow_node_visits += cv[kAccAnyNodeVisits - kAccPhotons];
    st.shadow_tri_tests += cv[kAccAnyTriTests - kAccPhotons];
}

gp_status Engine::readJobAcc(int slot, unsigned long long ledger[5], unsigned long long* counters) {
    unsigned long long rows[kSegs * kAccRow];
    GP_HIP(hipMemcpyAsync(rows, dJobAcc + (size_t)slot * kSegs * kAccRow, sizeof rows,
                          hipMemcpyDeviceToHost, device().stream));
    GP_HIP(hipStreamSynchronize(device().stream));
    sumAccRows(rows, ledger, counters);
    return GP_OK;
}

bool binningFor(const gp_tuning& t, int64_t maxTris) {
    return t.ray_sort == 1 || (t.ray_sort == 2 && maxTris >= kBinAutoTris && maxTris < kBinAutoMaxTris);
}

gp_status Engine::ensure(unsigned cap, int lds, int stackNeed, int cams, bool mirrors, int64_t maxTris) {
    if (!tunSet) { tun = defaultTuning(); tunSet = true; }
    DeviceState& d = device();
    const int cus = d.cus > 0 ? d.cus : 256;
    ls.threadsExtend = 256;
    ls.threadsShade = 256;
    ls.blocksExtend = std::max(1, cus * tun.grid_waves_per_cu / 4);
    ls.blocksShade = std::max(1, cus * tun.grid_waves_per_cu / 4);
    const size_t threads = (size_t)ls.blocksExtend * 256 * kExtendGridScale;
    const unsigned sc = segmentBound(cap, (unsigned)ls.blocksShade);
    const int camsNeed = std::max(camCap, std::max(1, cams));
    const bool needQueues = (cap != capacity) || sc != segCap || camsNeed != camCap || !dCtrl;
    const bool wantMirror = mirrors || mirrorQueue;
    const bool needMirrorAlloc = wantMirror && (needQueues || !q.mv[0]);
    const int binned = binningFor(tun, maxTris) ? 1 : 0;
    const bool needBinAlloc = binned && (needQueues || !q.pb);
    // Entries beyond the LDS stack live in the spill area (stack index sp >= SD maps to
    // spill[sp - SD], and sp < stackNeed). k_connect keeps as many entries in LDS as the
    // closest-hit kernel, 8 for its 12 (4-byte entries: at 8, six waves per SIMD fit its
    // LDS; 12 or 16 leave room for five, 24 for four); k_mirror runs its 16-entry
    // variant for 12. The spill area covers the shallowest of them.
    const int connLds = lds == 12 ? 8 : lds;
    const int needSpill = std::max(0, stackNeed - std::min(lds, connLds));
    const int connUseSpill = stackNeed > connLds ? 1 : 0;
    const bool needSpillAlloc = needSpill > spillDepth || (needSpill > 0 && threads > spillThreads);
    // The kernels without spill code keep the whole stack in LDS: they may only run
    // when every traced scene fits. A change of variant re-captures the graph.
    const int useSpill = needSpill > 0 ? 1 : 0;
    if (!needQueues && !needSpillAlloc && !needMirrorAlloc && !needBinAlloc && q.lds == lds &&
        q.useSpill == useSpill && q.connLds == connLds && q.connUseSpill == connUseSpill &&
        graphShape == ls.blocksExtend && q.binned == binned)
        return GP_OK;
    if (graph) { hipGraphExecDestroy(graph); graph = nullptr; }
    if (needQueues) {
        releaseQueues();
        auto alloc = [&](size_t bytes, void** out) -> gp_status {
            hipError_t e = hipMalloc(out, bytes);
            if (e != hipSuccess) return fail(GP_ERR_OOM, std::string("hipMalloc: ") + hipGetErrorString(e));
            bufs[nbufs++] = *out;
            return GP_OK;
        };
        gp_status st;
        const size_t pcap = (size_t)kSegs * sc;
        for (int k = 0; k < 2; ++k) {
            if ((st = alloc(pcap * sizeof(float4), (void**)&q.ro[k])) != GP_OK) return st;
            if ((st = alloc(pcap * sizeof(float4), (void**)&q.rd[k])) != GP_OK) return st;
            if ((st = alloc(pcap * sizeof(float4), (void**)&q.lam[k])) != GP_OK) return st;
            if ((st = alloc(pcap * sizeof(float4), (void**)&q.beta[k])) != GP_OK) return st;
            if ((st = alloc(pcap * sizeof(uint4), (void**)&q.st[k])) != GP_OK) return st;
        }
        if ((st = alloc(pcap * sizeof(float2), (void**)&q.hit)) != GP_OK) return st;
        // Connections per iteration and camera: one per emitted photon (direct view of
        // the light) plus at most one per live path; k_emit and k_shade run the same
        // grid, so a segment receives at most segCap of each per camera.
        const size_t ccap = 2 * pcap * (size_t)camsNeed;
        if ((st = alloc(ccap * sizeof(float4), (void**)&q.co)) != GP_OK) return st;
        if ((st = alloc(ccap * sizeof(float4), (void**)&q.cd)) != GP_OK) return st;
        if ((st = alloc(ccap * sizeof(float4), (void**)&q.cx)) != GP_OK) return st;
        GP_HIP(hipMalloc(&dCtrl, sizeof(Ctrl)));
        GP_HIP(hipMalloc(&dJobs, sizeof(DevJob) * kMaxJobs));
        GP_HIP(hipHostMalloc(&hCtrl, sizeof(Ctrl)));
        GP_HIP(hipHostMalloc(&hJobs, sizeof(DevJob) * kMaxJobs));
        GP_HIP(hipMalloc(&dJobAcc, sizeof(unsigned long long) * kMaxJobs * kSegs * kAccRow));
        GP_HIP(hipMalloc(&dJobCams, sizeof(DevCamera) * kMaxJobs * kMaxCams));
        GP_HIP(hipHostMalloc(&hJobCams, sizeof(DevCamera) * kMaxJobs * kMaxCams));
        std::memset(hCtrl, 0, sizeof(Ctrl));
        std::memset(hJobs, 0, sizeof(DevJob) * kMaxJobs);
        capacity = cap;
        segCap = sc;
        q.ctrl = dCtrl;
        q.jobs = dJobs;
        q.segCap = sc;
        q.connSegCap = 2 * sc * (unsigned)camsNeed;
        q.mirrorSegCap = 2 * sc;
        camCap = camsNeed;
    }
    if (needMirrorAlloc) {
        // one vertex per emitted photon and per shaded path, like the connections of
        // one camera
        const size_t mcap = 2 * (size_t)kSegs * segCap;
        float4* mv[4] = {};
        for (int k = 0; k < 4; ++k) {
            hipError_t e = hipMalloc(&mv[k], mcap * sizeof(float4));
            if (e != hipSuccess) {
                for (int i = 0; i < k; ++i) hipFree(mv[i]);
                return fail(GP_ERR_OOM, std::string("hipMalloc: ") + hipGetErrorString(e));
            }
        }
        for (int k = 0; k < 4; ++k) {
            q.mv[k] = mv[k];
            bufs[nbufs++] = mv[k];
        }
        mirrorQueue = true;
    }
    if (needBinAlloc) {
        // the binned path queue (same slots as the others, 80-byte records), the keys of
        // both path queues and the counting-sort tables
        const size_t pcap = (size_t)kSegs * segCap;
        void* p[6] = {};
        const size_t bytes[6] = {pcap * 5 * sizeof(float4), pcap * sizeof(unsigned short),
                                 pcap * sizeof(unsigned short), (size_t)kBinCols * kBins * sizeof(unsigned),
                                 (size_t)kBinChunks * kBins * sizeof(unsigned), (size_t)kBins * sizeof(unsigned)};
        for (int k = 0; k < 6; ++k) {
            hipError_t e = hipMalloc(&p[k], bytes[k]);
            if (e != hipSuccess) {
                for (int i = 0; i < k; ++i) hipFree(p[i]);
                return fail(GP_ERR_OOM, std::string("hipMalloc: ") + hipGetErrorString(e));
            }
        }
        for (int k = 0; k < 6; ++k) bufs[nbufs++] = p[k];
        q.pb = (float4*)p[0];
        q.key[0] = (unsigned short*)p[1];
        q.key[1] = (unsigned short*)p[2];
        q.binHist = (unsigned*)p[3];
        q.binChunk = (unsigned*)p[4];
        q.binTotal = (unsigned*)p[5];
    }
    if (needSpillAlloc) {
        const int depthAlloc = std::max(needSpill, spillDepth);
        if (q.spill) hipFree(q.spill);
        q.spill = nullptr;
        hipError_t e = hipMalloc(&q.spill, threads * (size_t)depthAlloc * sizeof(int2));
        if (e != hipSuccess) { spillDepth = 0; spillThreads = 0; return fail(GP_ERR_OOM, "spill stack allocation failed"); }
        spillDepth = depthAlloc;
        spillThreads = threads;
    }
    q.spillDepth = spillDepth;
    q.lds = lds;
    q.useSpill = useSpill;
    q.connLds = connLds;
    q.connUseSpill = connUseSpill;
    q.binned = binned;
    q.binCols = std::min(kBinCols, std::max(kBinColStep, (cus * 4) / kBinColStep * kBinColStep));
    static_assert(kBinColStep % kBinChunks == 0 && kBinCols % kBinColStep == 0, "binning column shape");
    graphShape = ls.blocksExtend;
    return GP_OK;
}

gp_status Engine::enqueueIterations(int n) {
    hipStream_t s = device().stream;
    const bool prof = tun.profile != 0;
    auto timed = [&](int k, auto&& launch) -> gp_status {
        if (!prof) { launch(); return GP_OK; }
        std::array<hipEvent_t, 2> ev;
        GP_HIP(hipEventCreate(&ev[0]));
        GP_HIP(hipEventCreate(&ev[1]));
        GP_HIP(hipEventRecord(ev[0], s));
        launch();
        GP_HIP(hipEventRecord(ev[1], s));
        pending[k].push_back(ev);
        return GP_OK;
    };
    gp_status st;
    for (int i = 0; i < n; ++i) {
        const int cur = parity;
        if ((st = timed(0, [&] { launchSchedule(q, cur, s); })) != GP_OK) return st;
        if ((st = timed(1, [&] { launchEmit(q, cur, ls, s); })) != GP_OK) return st;
        if (q.binned)
            if ((st = timed(5, [&] { launchBin(q, cur, s); })) != GP_OK) return st;
        if ((st = timed(2, [&] { launchExtend(q, cur, ls, s); })) != GP_OK) return st;
        if ((st = timed(3, [&] { launchShade(q, cur, ls, s); })) != GP_OK) return st;
        if ((st = timed(4, [&] { launchConnect(q, ls, s); })) != GP_OK) return st;
        if (mirrorActive && q.mv[0])
            if ((st = timed(4, [&] { launchMirror(q, ls, s); })) != GP_OK) return st;
        parity ^= 1;
    }
    GP_HIP(hipGetLastError());
    return GP_OK;
}

void Engine::resetProfile() {
    for (int k = 0; k < kKernelSlots; ++k) {
        for (auto& ev : pending[k]) { hipEventDestroy(ev[0]); hipEventDestroy(ev[1]); }
        pending[k].clear();
        kernelMs[k] = 0.0;
        kernelLaunches[k] = 0;
    }
}

gp_status Engine::collectProfile() {
    for (int k = 0; k < kKernelSlots; ++k) {
        for (auto& ev : pending[k]) {
            float ms = 0.f;
            GP_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
            kernelMs[k] += ms;
            kernelLaunches[k] += 1;
            hipEventDestroy(ev[0]);
            hipEventDestroy(ev[1]);
        }
        pending[k].clear();
    }
    return GP_OK;
}

gp_status Engine::runIterations(int n) {
    hipStream_t s = device().stream;
    const int K = chunkIterations(tun);
    if (!tun.use_graph || tun.profile || q.stats || parity != 0 || n < K) return enqueueIterations(n);
    if (!graph || graphIters != K || graphLds != q.lds || graphMirror != mirrorActive || graphBinned != q.binned) {
        if (graph) { hipGraphExecDestroy(graph); graph = nullptr; }
        hipGraph_t g;
        GP_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        int saved = parity;
        gp_status st = enqueueIterations(K);
        hipError_t ce = hipStreamEndCapture(s, &g);
        parity = saved;
        if (st != GP_OK) return st;
        if (ce != hipSuccess) return fail(GP_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ce));
        hipError_t ie = hipGraphInstantiate(&graph, g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        if (ie != hipSuccess) { graph = nullptr; return fail(GP_ERR_HIP, "hipGraphInstantiate failed"); }
        graphIters = K;
        graphLds = q.lds;
        graphMirror = mirrorActive;
        graphBinned = q.binned;
    }
    int done = 0;
    while (n - done >= K) {
        GP_HIP(hipGraphLaunch(graph, s));
        done += K;
    }
    if (done < n) return enqueueIterations(n - done);
    return GP_OK;
}

gp_status Engine::readCtrl() {
    hipStream_t s = device().stream;
    GP_HIP(hipMemcpyAsync(hCtrl, dCtrl, sizeof(Ctrl), hipMemcpyDeviceToHost, s));
    GP_HIP(hipStreamSynchronize(s));
    if (hCtrl->overflow) return fail(GP_ERR_CAPACITY, "wavefront queue segment overflow");
    if (tun.profile) return collectProfile();
    return GP_OK;
}

}  // namespace gp
