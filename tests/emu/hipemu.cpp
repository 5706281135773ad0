// Test infrastructure only: runtime of the CPU emulation build (see hipemu.hpp).
#include "hipemu.hpp"

#include <dlfcn.h>
#include <condition_variable>
#include <map>
#include <memory>
#include <mutex>
#include <semaphore>
#include <thread>
#include <vector>

thread_local dim3 threadIdx, blockIdx;
dim3 blockDim, gridDim;

namespace {

enum State { Runnable, AtColl, AtSync, Done };
enum Kind { KBallot, KShfl, KSync, KSyncOr };

struct Lane {
    std::binary_semaphore go{0};
    State st = Runnable;
    Kind kind = KBallot;
    void* site = nullptr;
    long long arg = 0;
    int src = 0;
    long long result = 0;
    unsigned tid = 0;
};

struct Block {
    std::vector<std::unique_ptr<Lane>> lanes;
    std::binary_semaphore back{0};
    int nLanes = 0;
};

thread_local Lane* gLane = nullptr;
thread_local Block* gBlock = nullptr;

void yieldLane(Lane& L) {
    gBlock->back.release();
    L.go.acquire();
}

void runBlock(const std::function<void()>& fn, unsigned bx, unsigned nthreads) {
    Block B;
    B.nLanes = (int)nthreads;
    for (unsigned t = 0; t < nthreads; ++t) {
        B.lanes.emplace_back(new Lane());
        B.lanes.back()->tid = t;
    }
    std::vector<std::thread> th;
    th.reserve(nthreads);
    for (unsigned t = 0; t < nthreads; ++t) {
        Lane* L = B.lanes[t].get();
        th.emplace_back([&B, L, &fn, bx]() {
            L->go.acquire();
            threadIdx = dim3(L->tid, 0, 0);
            blockIdx = dim3(bx, 0, 0);
            gLane = L;
            gBlock = &B;
            fn();
            L->st = Done;
            B.back.release();
        });
    }
    while (true) {
        // run every runnable lane until it blocks or finishes
        bool any = false;
        for (auto& lp : B.lanes) {
            Lane& L = *lp;
            if (L.st == Runnable) {
                any = true;
                L.go.release();
                B.back.acquire();
            }
        }
        (void)any;
        bool preempted = false;
        for (auto& lp : B.lanes) preempted = preempted || lp->st == Runnable;
        if (preempted) continue;
        // resolve collectives per wave per site
        bool progress = false;
        const int nw = (B.nLanes + 63) / 64;
        for (int w = 0; w < nw; ++w) {
            std::map<void*, std::vector<Lane*>> groups;
            for (int l = w * 64; l < std::min(B.nLanes, w * 64 + 64); ++l) {
                Lane& L = *B.lanes[l];
                if (L.st == AtColl) groups[L.site].push_back(&L);
            }
            for (auto& kv : groups) {
                auto& g = kv.second;
                const Kind k = g[0]->kind;
                for (Lane* L : g)
                    if (L->kind != k) { fprintf(stderr, "EMU: mixed collectives at one site\n"); abort(); }
                if (k == KBallot) {
                    unsigned long long m = 0;
                    for (Lane* L : g) if (L->arg) m |= 1ull << (L->tid & 63);
                    for (Lane* L : g) { L->result = (long long)m; L->st = Runnable; }
                } else {
                    std::map<unsigned, long long> val;
                    for (Lane* L : g) val[L->tid & 63] = L->arg;
                    for (Lane* L : g) {
                        unsigned src = L->src >= 0 ? (unsigned)L->src : ((L->tid & 63) ^ (unsigned)(-1 - L->src));
                        src &= 63;
                        auto it = val.find(src);
                        if (it == val.end()) {
                            fprintf(stderr, "EMU: shfl from inactive lane %u (block %u)\n", src, blockIdx.x);
                            abort();
                        }
                        L->result = it->second;
                    }
                    for (Lane* L : g) L->st = Runnable;
                }
                progress = true;
            }
        }
        if (progress) continue;
        int done = 0, sync = 0;
        for (auto& lp : B.lanes) {
            if (lp->st == Done) ++done;
            else if (lp->st == AtSync) ++sync;
        }
        if (done == B.nLanes) break;
        if (sync + done == B.nLanes && sync > 0) {
            if (done) { fprintf(stderr, "EMU: barrier with %d exited lanes\n", done); abort(); }
            void* site = nullptr;
            long long orv = 0;
            for (auto& lp : B.lanes) {
                if (lp->st != AtSync) continue;
                if (!site) site = lp->site;
                orv |= lp->arg;
            }
            for (auto& lp : B.lanes)
                if (lp->st == AtSync) { lp->result = orv != 0; lp->st = Runnable; }
            continue;
        }
        fprintf(stderr, "EMU: deadlock (done %d, at barrier %d)\n", done, sync);
        abort();
    }
    for (auto& t : th) t.join();
}

struct Rec {
    std::function<void()> fn;
    dim3 g, b;
};
bool gCapturing = false;
std::vector<Rec> gCapture;

}  // namespace

struct EmuGraph {
    std::vector<Rec> recs;
};
struct EmuStream {
    int id;
};
struct EmuEvent {
    int id;
};

unsigned long long emu_ballot(int pred, void* site) {
    Lane& L = *gLane;
    L.kind = KBallot; L.arg = pred ? 1 : 0; L.site = site; L.st = AtColl;
    yieldLane(L);
    return (unsigned long long)L.result;
}
int emu_shfl_i(int v, int src, void* site) {
    Lane& L = *gLane;
    L.kind = KShfl; L.arg = v; L.src = src; L.site = site; L.st = AtColl;
    yieldLane(L);
    return (int)L.result;
}
unsigned emu_lane() { return gLane->tid & 63; }
// With EMU_PREEMPT=k, a lane gives up the baton at every k-th preemption point, so the
// lanes of a block interleave inside loops that have no collectives (as waves do).
void emu_maybe_yield() {
    static const int every = getenv("EMU_PREEMPT") ? atoi(getenv("EMU_PREEMPT")) : 0;
    if (!every || !gLane) return;
    static thread_local unsigned count = 0;
    if (++count % (unsigned)every) return;
    Lane& L = *gLane;
    L.st = Runnable;
    yieldLane(L);
}
void emu_sync(void* site) {
    Lane& L = *gLane;
    L.kind = KSync; L.arg = 0; L.site = site; L.st = AtSync;
    yieldLane(L);
}
int emu_sync_or(int pred, void* site) {
    Lane& L = *gLane;
    L.kind = KSyncOr; L.arg = pred ? 1 : 0; L.site = site; L.st = AtSync;
    yieldLane(L);
    return (int)L.result;
}

void emu_launch(std::function<void()> fn, dim3 g, dim3 b) {
    if (gCapturing) { gCapture.push_back({fn, g, b}); return; }
    gridDim = g;
    blockDim = b;
    for (unsigned bx = 0; bx < g.x; ++bx) runBlock(fn, bx, b.x);
}

static int gEmuCus() {
    const char* e = getenv("EMU_CUS");
    return e ? atoi(e) : 2;
}
hipError_t hipGetDeviceCount(int* n) { *n = 1; return hipSuccess; }
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipGetDevice(int* d) { *d = 0; return hipSuccess; }
hipError_t hipGetSymbolAddress(void** p, const void* symbol) { *p = const_cast<void*>(symbol); return hipSuccess; }
hipError_t hipGetDeviceProperties(hipDeviceProp_t* p, int) {
    std::memset(p, 0, sizeof *p);
    std::strcpy(p->name, "CPU emulator");
    std::strcpy(p->gcnArchName, "gfx950:sramecc+:xnack-");
    p->multiProcessorCount = gEmuCus();
    p->sharedMemPerBlock = 160 * 1024;
    p->warpSize = 64;
    return hipSuccess;
}
hipError_t hipMemGetInfo(size_t* f, size_t* t) { *f = size_t(8) << 30; *t = size_t(16) << 30; return hipSuccess; }
const char* hipGetErrorString(hipError_t e) { return e == hipSuccess ? "success" : "emulated error"; }
hipError_t hipGetLastError() { return hipSuccess; }
hipError_t hipPeekAtLastError() { return hipSuccess; }
hipError_t hipDeviceSynchronize() { return hipSuccess; }
// fresh device memory holds garbage on the GPU: fill it so reads of unwritten words
// show up (EMU_FILL, default 0xA5)
hipError_t hipMalloc(void** p, size_t n) {
    *p = malloc(n ? n : 1);
    if (!*p) return hipErrorOutOfMemory;
    const char* f = getenv("EMU_FILL");
    std::memset(*p, f ? atoi(f) : 0xA5, n);
    return hipSuccess;
}
hipError_t hipMallocAsync(void** p, size_t n, hipStream_t) { return hipMalloc(p, n); }
hipError_t hipFree(void* p) { free(p); return hipSuccess; }
hipError_t hipFreeAsync(void* p, hipStream_t) { free(p); return hipSuccess; }
hipError_t hipHostMalloc(void** p, size_t n, unsigned) { return hipMalloc(p, n); }
hipError_t hipHostFree(void* p) { free(p); return hipSuccess; }
hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) { if (n) std::memmove(d, s, n); return hipSuccess; }
hipError_t hipMemcpy2D(void* d, size_t dpitch, const void* s, size_t spitch, size_t width, size_t height,
                       hipMemcpyKind) {
    if (width > dpitch || width > spitch) return hipErrorInvalidValue;
    for (size_t r = 0; r < height; ++r)
        std::memmove((char*)d + r * dpitch, (const char*)s + r * spitch, width);
    return hipSuccess;
}
hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind k, hipStream_t) {
    if (gCapturing) { fprintf(stderr, "EMU: memcpy during capture\n"); abort(); }
    return hipMemcpy(d, s, n, k);
}
hipError_t hipMemset(void* d, int v, size_t n) { std::memset(d, v, n); return hipSuccess; }
hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t) {
    if (gCapturing) { fprintf(stderr, "EMU: memset during capture\n"); abort(); }
    std::memset(d, v, n);
    return hipSuccess;
}
hipError_t hipStreamCreate(hipStream_t* s) { *s = new EmuStream{1}; return hipSuccess; }
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) { return hipStreamCreate(s); }
hipError_t hipStreamDestroy(hipStream_t s) { delete s; return hipSuccess; }
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
hipError_t hipEventCreate(hipEvent_t* e) { *e = new EmuEvent{1}; return hipSuccess; }
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }
hipError_t hipEventDestroy(hipEvent_t e) { delete e; return hipSuccess; }
hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipEventQuery(hipEvent_t) { return hipSuccess; }
hipError_t hipEventElapsedTime(float* ms, hipEvent_t, hipEvent_t) { *ms = 0.f; return hipSuccess; }
hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) { gCapturing = true; gCapture.clear(); return hipSuccess; }
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t* g) {
    gCapturing = false;
    *g = new EmuGraph{gCapture};
    gCapture.clear();
    return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, hipGraphNode_t*, char*, size_t) {
    *e = new EmuGraph{g->recs};
    return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t) {
    for (auto& r : e->recs) emu_launch(r.fn, r.g, r.b);
    return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g) { delete g; return hipSuccess; }
hipError_t hipGraphExecDestroy(hipGraphExec_t e) { delete e; return hipSuccess; }
