// Plain-old-data layouts shared by the host runtime and the CDNA4 kernels.
// Everything the kernels read lives in HBM in these forms:
//   BVH        four-wide nodes, 128 B each (four slot boxes in SoA form + four links)
//   triangles  3 x float4 per triangle in leaf order (vertices; .w = prim id,
//              material id), plus 1 x float4 geometric normal (.w = material)
//   paths      structure-of-arrays queues of float4/uint4 (80 B per path), split
//              into kSegs segments, each with its own append counter
//   hits       one float2 per path (t, leaf-order triangle index)
//   camera     connections: 3 x float4 per record (48 B)
//   mirrors    vertices to connect through flat mirrors: 4 x float4 per record (64 B)
//   films      int64 fixed-point (X, Y, Z, hits) per pixel, 32 B
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace gp {

constexpr int kMaxHero = 4;           // wavelengths per path (hero + 3)
constexpr int kMaxJobs = 256;         // job slots in one wavefront
constexpr int kLambdaBins = 470;      // 1-nm bins over [360, 830)
constexpr int kCdfGuide = 64;         // guide-table cells over [0, 1) for the CDF inversion
constexpr float kRayEps = 1e-6f;      // origin offset and tmin, as the reference CPU path
constexpr int kLeafMax = 4;           // triangles per BVH leaf
// Queues are split into kSegs segments, each appended to through its own counter
// (its own 128-byte line): workgroup b appends to segment b % kSegs. Same-address
// atomics serialise (~12 ns each on MI355X), so eight counters instead of one cut
// the append cost of a launch eightfold; and because a segment only receives the
// items of every eighth workgroup, its fill is bounded without any assumption about
// where the hardware places workgroups.
constexpr int kSegs = 8;
constexpr int kLine = 32;             // u32 per 128-byte line
constexpr int kAccRow = 16;           // u64 per per-segment ledger/counter row
// Quantities of a per-job accumulator row (u64 index in the row).
constexpr int kAccEnergy = 0;         // emitted, absorbed, sensor, escaped, residual (fixed point)
constexpr int kAccPhotons = 5;        // photons emitted
constexpr int kAccDirect = 6;         // camera connections queued at the emission points
constexpr int kAccSegments = 7;       // closest-hit queries of the paths
constexpr int kAccNodeVisits = 8;     // their four-wide node visits (traversal statistics only)
constexpr int kAccTriTests = 9;       // their triangle tests (traversal statistics only)
constexpr int kAccConnections = 10;   // connection ray queries
constexpr int kAccSplats = 11;        // connections that reached the film
constexpr int kAccAnyNodeVisits = 12; // camera shadow rays' node visits (traversal statistics only)
constexpr int kAccAnyTriTests = 13;   // camera shadow rays' triangle tests (traversal statistics only)
constexpr int kAccCounters = 9;       // counters [kAccPhotons, kAccPhotons + kAccCounters)

// Ray binning (gp_tuning.ray_sort): before the closest-hit kernel, the paths of an
// iteration are reordered by a 12-bit key, the Morton code of the origin's cell in an
// 8 x 8 x 8 grid over the job's triangle-centroid box followed by the direction
// octant, so that the lanes of a wave trace similar rays.
// A counting sort: binCols workgroups (4 per CU, at most kBinCols) each count their
// contiguous share of the queue into a histogram column, the columns are scanned in
// (bin, column) order, and each workgroup scatters its share to the bins' ranges.
constexpr int kBinBits = 12;
constexpr int kBins = 1 << kBinBits;
constexpr int kBinCols = 1024;        // most histogram columns (workgroups of the count/scatter)
constexpr int kBinChunks = 8;         // column chunks per bin in the scan
constexpr int kBinColStep = 8;        // the column count is a multiple of this (and of kBinChunks)
constexpr int kBinThreads = 1024;     // threads of a count/scatter workgroup (one column)
// ray_sort = 2 (automatic) bins while the largest traced scene has at least
// kBinAutoTris and fewer than kBinAutoMaxTris triangles. Below, the traversal is too
// cheap for the sort to pay (c0, 36 triangles: 3.75 G segments/s unbinned, 3.46 G
// binned); at a million triangles and more the closest-hit kernel's postponed-leaf
// traversal gains less from the sort than the sort costs (1M / 2M random soups 0.84 vs
// 0.81 G, the 1.3M-triangle icosphere 2.57 vs 2.39 G, c2's 5.24M 2.51 vs 2.47 G); c1
// (8,064 triangles) gains: 3.10 vs 2.98 G. Between 8K and 1.3M triangles the bound is
// not measured.
constexpr int64_t kBinAutoTris = 4096;
constexpr int64_t kBinAutoMaxTris = 1 << 16;
constexpr int kBinCellsPerAxis = 8;

// Leaf links are negative: ~((first << 3) | (count - 1)).
inline constexpr int leafCode(int first, int count) { return ~((first << 3) | (count - 1)); }

struct DevSpectrum {
    int32_t kind;
    int32_t n;           // sampled: number of samples
    const float* tab;    // sampled: the samples (the scene's device table)
    float p[6];          // sampled: p[0] first sample's wavelength, p[1] samples per nm
    float band_lo, band_hi;
};  // 48 B

struct DevMaterial {
    int32_t type;
    float roughness;
    int32_t pad[2];
    DevSpectrum reflect;
    DevSpectrum ior;
};  // 112 B

struct DevEmitter {
    float origin[3], u[3], v[3], normal[3], beam[3];
    int32_t collimated;
    float power;        // integral(SPD) * area * pi
    int32_t pad;
    const double* cdf;  // kLambdaBins + 1 entries, normalised
    const int32_t* guide;  // kCdfGuide + 1 entries: guide[k] = last bin with cdf <= k / kCdfGuide
};

constexpr int kMaxCams = 16;          // cameras one job splats into (shared photon flight)
constexpr int kMaxMirrorPlanes = 64;  // flat mirror surfaces one scene connects through

// One flat mirror surface: the coplanar mirror triangles of a plane, reduced to the
// plane (unit normal n, dot(n, x) = d) and the padded box of its triangles. The
// connection geometry is solved in double, so the plane is stored in double.
struct DevMirrorPlane {
    double n[3];
    double d;
    double lo[3], hi[3];
};  // 80 B

// Pinhole camera and the film it splats into (Camera::lookAt basis, float).
struct DevCamera {
    float eye[3], u[3], v[3], w[3];
    float tanHalfX, tanHalfY, pixelArea;
    int32_t resX, resY;
    int32_t pad;
    long long* film;             // res*4 int64
    double filmScale;            // contribution -> integer units
    unsigned long long filmQmax; // saturation value of one splat
};

// One job = one (scene, cameras + films, photon budget). A single gp_trace is one job;
// the batched and serving paths run many jobs in the same queues. Every photon of a
// job connects to each of its cameras.
struct DevJob {
    // scene
    const float4* nodes;
    const float4* tris;
    const float4* trin;
    const DevMaterial* mats;
    const DevEmitter* ems;
    const double* emCdf;     // power CDF over emitters (n_em entries)
    int32_t nEm;
    float basePower;         // photon power: the single emitter's, else the total
    // cameras
    const DevCamera* cams;   // nCams entries
    int32_t nCams;
    int32_t nMirrors;        // flat mirror planes of the scene (0: no mirror connections)
    const DevMirrorPlane* mirrors;
    // kSegs rows of kAccRow u64 (layout: the kAcc* indices above); workgroup b adds
    // into row b % kSegs
    unsigned long long* acc;
    double energyScale;
    // path settings
    unsigned long long seedBase, salt;
    int32_t heroC, maxDepth, rrDepth, pad0;
    // ray binning: origin cell = (o - binLo) * binScale, clamped to the grid
    float binLo[3], binScale[3];
};

// Connection tag (float bits of the record's 4th word): job slot in bits 0..15,
// live wavelengths in 16..18, camera in 19..22.
inline constexpr unsigned connTag(unsigned job, unsigned nUp, unsigned cam) {
    return job | (nUp << 16) | (cam << 19);
}

// Device-side wavefront control block.
struct Ctrl {
    unsigned int pathCount[2][kSegs][kLine];   // per segment of path queue 0 / 1 (one line each)
    unsigned int connCount[kSegs][kLine];      // per segment of the connection queue
    unsigned int mirrorCount[kSegs][kLine];    // per segment of the mirror-vertex queue
    unsigned int nEmit;         // photons emitted this iteration
    unsigned int iter;          // iterations run
    int32_t nJobs;              // job slots in use are below this
    unsigned int capacity;      // path queue capacity (total over segments)
    unsigned int segCap;        // path slots per segment
    unsigned int connSegCap;    // connection slots per segment
    unsigned int overflow;      // set by a kernel that found a segment full (never expected)
    unsigned int mirrorSegCap;  // mirror-vertex slots per segment
    unsigned int segBase[kSegs];        // survivors already in each segment (emission goes after)
    unsigned int segEmit[kSegs + 1];    // prefix of this iteration's emission over segments
    unsigned int pad2[7];
    // per job
    unsigned long long nextPhoton[kMaxJobs];   // next photon index to emit (job-relative)
    unsigned long long endPhoton[kMaxJobs];    // one past the last photon
    int32_t lastEmitIter[kMaxJobs];            // iteration that emitted the job's last photon
    int32_t active[kMaxJobs];                  // 1 = emitting / in flight
    int32_t order[kMaxJobs];                   // job slots in emission priority order
    int32_t nOrder;                            // entries of `order` in use
    int32_t pad3[3];
    // this iteration's emission, indexed by position k in `order`
    unsigned int emitPrefix[kMaxJobs + 1];     // prefix of photons emitted before order[k]
    int32_t emitJob[kMaxJobs];                 // = order[k]
    unsigned long long emitStart[kMaxJobs];    // first photon index emitted for order[k]
};

// Path metadata word (uint4.w of the state array).
//   bits  0..15 job slot, 16..23 bounce, 24..26 live wavelengths, 31 dead
inline constexpr unsigned packMeta(unsigned job, unsigned bounce, unsigned nUp) {
    return job | (bounce << 16) | (nUp << 24);
}

}  // namespace gp
