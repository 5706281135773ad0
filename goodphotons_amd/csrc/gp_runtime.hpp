// Host runtime internals shared by the C ABI implementation files.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/goodphotons.h"
#include "gp_kernels.hpp"
#include "gp_types.hpp"

struct gp_scene {
    int64_t nTris = 0;
    int64_t nNodes = 0;                // four-wide nodes on the device
    int depth = 0;                     // child-pair levels of the built tree
    int stackNeed = 0;                 // traversal stack entries a ray can need (3 per four-wide level)
    // ray-binning grid over the triangle-centroid box: cell = (x - binLo) * binScale
    float binLo[3] = {0.f, 0.f, 0.f};
    float binScale[3] = {0.f, 0.f, 0.f};
    float4* dNodes = nullptr;
    float4* dTris = nullptr;
    float4* dTrin = nullptr;
    gp::DevMaterial* dMats = nullptr;
    gp::DevEmitter* dEms = nullptr;
    double* dCdfs = nullptr;
    double* dEmCdf = nullptr;
    int32_t* dGuides = nullptr;        // per emitter kCdfGuide + 1 CDF guide entries
    gp::DevMirrorPlane* dMirrors = nullptr;   // flat mirror planes (camera connections through them)
    float* dSpec = nullptr;            // samples of the scene's sampled spectra
    int nMirrors = 0;
    int nEm = 0;
    double totalPower = 0.0;
    double basePower = 0.0;
    int64_t deviceBytes = 0;
    double buildMs = 0.0;
    // host copies for camera-distance bounds (empty when the tree was built on the
    // device: the bound is then computed there and cached per eye position)
    std::vector<float> pairs;          // 16 floats per pair node (host copy of the built tree)
    std::vector<float> leafTris;       // 9 floats per triangle, leaf order
    std::vector<std::array<double, 12>> emitQuads;   // origin, u, v, normal
    int builder = GP_BVH_HOST_SAH;
    mutable std::map<std::array<double, 3>, double> eyeDist2;

    gp_scene() = default;
    gp_scene(const gp_scene&) = delete;
    gp_scene& operator=(const gp_scene&) = delete;
    ~gp_scene();   // frees the device tables
};

struct gp_film {
    int32_t rx = 0, ry = 0;
    int64_t maxPhotons = 0;
    unsigned long long* d = nullptr;
    double scale = 0.0;                 // contribution -> integer units (0 = not yet fixed)
    unsigned long long qmax = 0;        // largest single splat value
    double splatBudget = 0.0;           // photons*(maxDepth+1) the scale was sized for
    double splatUsed = 0.0;

    gp_film() = default;
    gp_film(const gp_film&) = delete;
    gp_film& operator=(const gp_film&) = delete;
    ~gp_film();    // frees d (a stream-ordered film clears d after its asynchronous free)
};

namespace gp {

gp_status fail(gp_status s, const std::string& msg);
#define GP_HIP(call)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            return ::gp::fail(GP_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

struct DeviceState {
    bool ok = false;
    gp_status status = GP_ERR_NO_DEVICE;
    std::string why;
    int id = 0;
    int cus = 0;
    size_t ldsPerBlock = 0;
    std::string name, arch;
    hipStream_t stream = nullptr;       // engine work
    hipStream_t copyStream = nullptr;   // result downloads that overlap engine work
};
DeviceState& device();
gp_status requireDevice();

// Camera basis as the reference's Camera::lookAt (src/camera.h:110-129), in double.
struct CamBasis {
    double eye[3], u[3], v[3], w[3];
    double tanHalfX, tanHalfY, pixelArea;
    int resX, resY;
};
bool cameraBasis(const gp_camera_desc& c, CamBasis& out);

// Smallest distance from point p to any triangle or emitter quad of the scene.
double sceneMinDistance(const gp_scene& s, const double p[3]);

// Kernel timing slots: schedule, emit, extend, shade, connect (+ mirror), binning.
constexpr int kKernelSlots = 6;

// Wavefront engine: device queues, control block, job table, graphs.
struct Engine {
    std::mutex mu;
    gp_tuning tun{};
    bool tunSet = false;
    unsigned capacity = 0;
    unsigned segCap = 0;                // path slots per queue segment
    int spillDepth = 0;
    size_t spillThreads = 0;
    Queues q{};
    LaunchShape ls{};
    Ctrl* dCtrl = nullptr;
    Ctrl* hCtrl = nullptr;              // pinned
    DevJob* dJobs = nullptr;
    DevJob* hJobs = nullptr;            // pinned
    unsigned long long* dJobAcc = nullptr;       // kMaxJobs * kSegs * kAccRow
    DevCamera* dJobCams = nullptr;               // kMaxJobs * kMaxCams
    DevCamera* hJobCams = nullptr;               // pinned staging for gp_trace
    int camCap = 1;                              // cameras per job the connection queue is sized for
    gp_status resetJobAcc(int slot);             // zero a slot's rows (stream-ordered)
    // ledger[5] and counters[kAccCounters] (the quantities from kAccPhotons on)
    gp_status readJobAcc(int slot, unsigned long long ledger[5], unsigned long long* counters);
    void initCtrl();                            // zero the host control block, set capacities
    void* bufs[40] = {};
    int nbufs = 0;
    int parity = 0;
    hipGraphExec_t graph = nullptr;
    int graphIters = 0;
    int graphLds = 0;
    int graphShape = 0;
    int graphBinned = 0;
    // k_mirror runs in each iteration while a job of a scene with flat mirrors is live
    // (set by the caller before runIterations; the graph is re-captured on a change)
    bool mirrorActive = false;
    bool graphMirror = false;
    bool mirrorQueue = false;           // the mirror-vertex queue has been requested once
    // The control block holds one trace or one server's jobs at a time. `owner` is the
    // server whose jobs are in the engine (live or with chunks in flight); `epoch` counts
    // re-initialisations, so a server sees that a trace used the engine since its last
    // step and starts from a fresh control block.
    const void* owner = nullptr;
    uint64_t epoch = 0;
    // per-kernel HIP-event timing (tuning.profile), slots as gp_trace_stats.kernel_ms
    std::vector<std::array<hipEvent_t, 2>> pending[kKernelSlots];
    double kernelMs[kKernelSlots] = {};
    uint64_t kernelLaunches[kKernelSlots] = {};
    void resetProfile();
    gp_status collectProfile();

    // stackNeed: traversal stack entries the traced scenes can need (gp_scene::stackNeed);
    // mirrors: allocate the mirror-vertex queue (kept from then on, like the camera count)
    // maxTris: the largest triangle count of the traced scenes (automatic ray binning)
    gp_status ensure(unsigned cap, int lds, int stackNeed, int cams = 1, bool mirrors = false, int64_t maxTris = 0);
    void release();
    void releaseQueues();
    void releaseSpill();
    gp_status enqueueIterations(int n);       // plain launches, tracks parity
    gp_status runIterations(int n);           // graph replay when possible
    gp_status readCtrl();                     // synchronous copy of the control block
};
Engine& engine();
gp_tuning defaultTuning();
// Iterations between host reads of the control block: iters_per_graph, rounded down
// to even when graphs are replayed (a graph covers both queue parities).
inline int chunkIterations(const gp_tuning& t) {
    return t.use_graph ? std::max(2, t.iters_per_graph & ~1) : std::max(1, t.iters_per_graph);
}
unsigned segmentBound(unsigned cap, unsigned blocks);
// Whether the wavefront bins its rays (gp_tuning.ray_sort: 1 always, 2 when the largest
// traced scene has kBinAutoTris to kBinAutoMaxTris - 1 triangles).
bool binningFor(const gp_tuning& t, int64_t maxTris);
// Sums the kSegs accumulator rows of one job: ledger[5] (fixed point) and
// counters[kAccCounters] (the quantities from kAccPhotons on).
void sumAccRows(const unsigned long long* rows, unsigned long long ledger[5], unsigned long long* counters);
// Adds summed counters into a stats record.
void addCounters(const unsigned long long* counters, gp_trace_stats& st);

// Job record for slot `slot` over cameras cams[0..nCams) and their films; the camera
// records (to be placed at J.cams) are written to camsOut.
gp_status fillJob(const gp_scene* s, const gp_camera_desc* cams, gp_film* const* films, int nCams,
                  const gp_trace_params& p, int slot, DevJob& J, DevCamera* camsOut);
gp_status prepareFilm(gp_film* film, const gp_scene* s, const CamBasis& cb,
                      const gp_trace_params& p);

// Flat mirror planes of a scene description: mirror triangles grouped by plane, at
// most kMaxMirrorPlanes (the largest total areas), boxes padded.
void buildMirrorPlanes(const gp_scene_desc* d, std::vector<DevMirrorPlane>& out);

}  // namespace gp
