// Host-side launch interface of the wavefront kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "gp_types.hpp"

namespace gp {

// Device pointers of one wavefront's queues. Path queues are double-buffered:
// iteration k reads queue (k & 1) and writes survivors to the other. Each queue is
// kSegs segments of segCap (paths) / connSegCap (connections) slots; a logical index
// maps to (segment, offset) through the prefix of the per-segment counts. With ray
// binning (binned = 1) the emission and shading kernels also write each path's bin key,
// and a third path queue holds the iteration's paths in bin order, packed from slot 0
// (logical index = slot) as 80-byte records (origin, direction, wavelengths, weights,
// state: five float4): the closest-hit and shading kernels read it instead of queue
// (k & 1).
struct Queues {
    float4* ro[2];
    float4* rd[2];
    float4* lam[2];
    float4* beta[2];
    uint4* st[2];
    unsigned short* key[2];   // bin key of every path of queue 0 / 1 (ray binning only)
    float4* pb;               // the binned queue, five float4 per path (ray binning only)
    float2* hit;
    float4* co;      // connection origin.xyz, segment end
    float4* cd;      // connection dir.xyz, pixel index
    float4* cx;      // connection X, Y, Z, (job | nUp << 16)
    // Vertices to connect to the cameras through flat mirrors (allocated once a scene
    // with mirrors is traced; null before): position + (job | nUp << 16), normal,
    // wavelengths, per-wavelength weight (beta * albedo).
    float4* mv[4];
    Ctrl* ctrl;
    const DevJob* jobs;
    int2* spill;     // traversal stack spill, spillDepth entries per thread
    int spillDepth;
    int lds;         // LDS stack entries per lane (8, 16, 24 or 32)
    int useSpill;    // 1: the traced scenes are deeper than the LDS stack (spill variant)
    int connLds;     // k_connect's LDS stack entries (4-byte entries: at least 24 fit at the
                     // occupancy its registers allow)
    int connUseSpill;
    int stats;       // 1: k_extend / k_connect count traversal statistics per job
    int binned;      // 1: ray binning on (the binning kernels run and queue 2 is read)
    int binCols;     // histogram columns (a multiple of kBinColStep, <= kBinCols)
    unsigned* binHist;    // kBinCols x kBins counts, one row per column; the scan turns
                          // them into offsets inside each (bin, chunk)
    unsigned* binChunk;   // kBins x kBinChunks exclusive prefixes of the column chunks
    unsigned* binTotal;   // kBins totals, then (after the scan) kBins first slots
    unsigned segCap, connSegCap, mirrorSegCap;
};

// The closest-hit kernel's grid holds kExtendGridScale times the threads of the other
// traversal kernels (its spill area is sized for that many): smaller shares of the queue
// per wave even out the waves' unequal traversal costs (c2: 203.5 -> 198.8 ms of
// closest-hit time per 2^28 photons at scale 4, 167.0 with ray binning; scale 8 no
// further).
constexpr int kExtendGridScale = 4;

struct LaunchShape {
    int blocksExtend, threadsExtend;
    int blocksShade, threadsShade;
};

void launchSchedule(const Queues& q, int cur, hipStream_t s);
void launchEmit(const Queues& q, int cur, const LaunchShape& ls, hipStream_t s);
void launchExtend(const Queues& q, int cur, const LaunchShape& ls, hipStream_t s);
void launchShade(const Queues& q, int cur, const LaunchShape& ls, hipStream_t s);
void launchConnect(const Queues& q, const LaunchShape& ls, hipStream_t s);
// Ray binning of queue `cur` into queue 2 (four kernels: count, chunk sums, bin scan,
// scatter).
void launchBin(const Queues& q, int cur, hipStream_t s);
void launchMirror(const Queues& q, const LaunchShape& ls, hipStream_t s);

// Batch probes over host-provided rays (already on the device).
void launchProbeClosest(const float4* nodes, const float4* tris, const float* rays, long long n,
                        float tmin, int* prim, float* t, int2* spill, int spillDepth, int lds,
                        int blocks, int threads, hipStream_t s);
void launchProbeAny(const float4* nodes, const float4* tris, const float* rays, const float* tmax,
                    long long n, float tmin, unsigned char* occ, int2* spill, int spillDepth,
                    int lds, int blocks, int threads, hipStream_t s);

// Stream-ordered job-table updates for a running wavefront. The data travel as
// kernel arguments (copied at launch), so the host needs no buffer that outlives the
// call and no synchronisation.
struct JobInstall {
    DevJob job;                     // job.cams points at the slot's camera table
    DevCamera cam;                  // its single camera
    unsigned long long next, end;   // photon range
    int32_t slot;
    int32_t pad[3];
};
struct OrderUpdate {
    int32_t order[kMaxJobs];
    int32_t nOrder, nJobs;
    int32_t clearSlot;               // job slot to deactivate, or -1
    int32_t pad;
};
void launchInstallJob(const Queues& q, const JobInstall& a, hipStream_t s);
void launchSetOrder(const Queues& q, const OrderUpdate& a, hipStream_t s);

// Zero-fill of 64-bit words by a kernel (usable inside captured graphs).
void launchFill64(unsigned long long* p, size_t n, unsigned long long v, hipStream_t s);

}  // namespace gp
