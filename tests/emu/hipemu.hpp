// Test infrastructure only: a CPU emulation of the HIP subset the library uses, so the
// unchanged kernel sources can be compiled as host C++ and run under AddressSanitizer
// (tests/emu/build.py). The product never links it.
//
// Every GPU lane runs as its own OS thread, but only one runs at a time (a baton), so
// emulated device memory needs no synchronisation. Wave collectives (ballot, any, shfl)
// resolve over the lanes of a 64-lane wave that wait at the same call site (the lanes
// that would be active together on the GPU); barriers resolve over the block. Device
// memory is host malloc filled with garbage (reads of unwritten words show up), so
// AddressSanitizer bounds-checks every kernel access. With EMU_PREEMPT=k a lane gives up
// the baton at every k-th preemption point (each __float_as_int, which the traversal
// loops call per node), so lanes of a block interleave inside loops as waves do.
#pragma once
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <tuple>
#include <type_traits>
#include <utility>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline __attribute__((always_inline))
#define __align__(n) __attribute__((aligned(n)))
#define __launch_bounds__(...)
#define __shared__ static

struct dim3 {
    unsigned x, y, z;
    constexpr dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct float2 { float x, y; };
struct float3 { float x, y, z; };
inline float3 make_float3(float a, float b, float c) { return {a, b, c}; }
struct float4 { float x, y, z, w; };
struct int2 { int x, y; };
struct uint4 { unsigned x, y, z, w; };
inline float2 make_float2(float a, float b) { return {a, b}; }
inline float4 make_float4(float a, float b, float c, float d) { return {a, b, c, d}; }
inline int2 make_int2(int a, int b) { return {a, b}; }
inline uint4 make_uint4(unsigned a, unsigned b, unsigned c, unsigned d) { return {a, b, c, d}; }

extern thread_local dim3 threadIdx, blockIdx;
extern dim3 blockDim, gridDim;

// wave collectives and barriers (emulator)
unsigned long long emu_ballot(int pred, void* site);
int emu_shfl_i(int v, int src, void* site);
unsigned emu_lane();
void emu_sync(void* site);
int emu_sync_or(int pred, void* site);
__attribute__((noinline)) inline unsigned long long __ballot(int pred) { return emu_ballot(pred, __builtin_return_address(0)); }
__attribute__((noinline)) inline int __any(int pred) { return emu_ballot(pred, __builtin_return_address(0)) != 0; }
__attribute__((noinline)) inline int __shfl(int v, int src, int width = 64) { (void)width; return emu_shfl_i(v, src, __builtin_return_address(0)); }
__attribute__((noinline)) inline unsigned __shfl(unsigned v, int src, int width = 64) { (void)width; return (unsigned)emu_shfl_i((int)v, src, __builtin_return_address(0)); }
__attribute__((noinline)) inline unsigned __shfl_xor(unsigned v, int m, int width = 64) { (void)width; return (unsigned)emu_shfl_i((int)v, -1 - m, __builtin_return_address(0)); }
__attribute__((noinline)) inline int __shfl_xor(int v, int m, int width = 64) { (void)width; return emu_shfl_i(v, -1 - m, __builtin_return_address(0)); }
__attribute__((noinline)) inline void __syncthreads() { emu_sync(__builtin_return_address(0)); }
// A wave barrier orders the lanes of one wave (on the GPU they run in lockstep and the
// builtin only fences code motion): emulated as a wave collective.
__attribute__((noinline)) inline void emu_wave_barrier() { (void)emu_ballot(1, __builtin_return_address(0)); }
#define __builtin_amdgcn_wave_barrier() emu_wave_barrier()
__attribute__((noinline)) inline int __syncthreads_or(int p) { return emu_sync_or(p, __builtin_return_address(0)); }
inline unsigned __lane_id() { return emu_lane(); }
inline unsigned emu_mbcnt_lo(unsigned m, unsigned acc) {
    unsigned l = emu_lane();
    unsigned mask = l >= 32 ? 0xffffffffu : ((1u << l) - 1u);
    return acc + (unsigned)__builtin_popcount(m & mask);
}
inline unsigned emu_mbcnt_hi(unsigned m, unsigned acc) {
    unsigned l = emu_lane();
    unsigned mask = l < 32 ? 0u : (l == 32 ? 0u : ((1u << (l - 32)) - 1u));
    return acc + (unsigned)__builtin_popcount(m & mask);
}
#define __builtin_amdgcn_mbcnt_lo emu_mbcnt_lo
#define __builtin_amdgcn_mbcnt_hi emu_mbcnt_hi
inline int __popcll(unsigned long long v) { return __builtin_popcountll(v); }
inline int __clz(int v) { return v == 0 ? 32 : __builtin_clz((unsigned)v); }
inline int __clzll(long long v) { return v == 0 ? 64 : __builtin_clzll((unsigned long long)v); }
inline int __ffsll(long long v) { return __builtin_ffsll(v); }
void emu_maybe_yield();   // preemption point: lets another lane run (interleaving)
inline int __float_as_int(float f) { emu_maybe_yield(); int i; std::memcpy(&i, &f, 4); return i; }
inline float __int_as_float(int i) { float f; std::memcpy(&f, &i, 4); return f; }
inline long long __double2ll_rn(double d) { return std::llrint(d); }
inline float __expf(float x) { return std::exp(x); }
// __hip_atomic_load/store are clang builtins on every target; the scope macros are not
#ifndef __HIP_MEMORY_SCOPE_WORKGROUP
#define __HIP_MEMORY_SCOPE_SINGLETHREAD 1
#define __HIP_MEMORY_SCOPE_WAVEFRONT 2
#define __HIP_MEMORY_SCOPE_WORKGROUP 3
#define __HIP_MEMORY_SCOPE_AGENT 4
#define __HIP_MEMORY_SCOPE_SYSTEM 5
#endif
inline double __longlong_as_double(long long v) { double d; std::memcpy(&d, &v, 8); return d; }
inline long long __double_as_longlong(double d) { long long v; std::memcpy(&v, &d, 8); return v; }
template <class T> inline T atomicAdd(T* p, T v) { T o = *p; *p = o + v; return o; }
template <class T> inline T atomicCAS(T* p, T cmp, T v) { T o = *p; if (o == cmp) *p = v; return o; }
template <class T> inline T atomicMax(T* p, T v) { T o = *p; if (v > o) *p = v; return o; }
template <class T> inline T atomicMin(T* p, T v) { T o = *p; if (v < o) *p = v; return o; }
template <class T> inline T atomicOr(T* p, T v) { T o = *p; *p = o | v; return o; }
template <class T> inline T atomicExch(T* p, T v) { T o = *p; *p = v; return o; }
inline void __threadfence() {}
inline void __threadfence_block() {}
template <class T> inline T min(T a, T b) { return a < b ? a : b; }
template <class T> inline T max(T a, T b) { return a > b ? a : b; }

// ---------------------------------------------------------------- runtime
typedef enum hipError_t {
    hipSuccess = 0, hipErrorInvalidValue = 1, hipErrorOutOfMemory = 2, hipErrorNotReady = 600,
    hipErrorUnknown = 999
} hipError_t;
typedef enum hipMemcpyKind {
    hipMemcpyHostToHost = 0, hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2,
    hipMemcpyDeviceToDevice = 3, hipMemcpyDefault = 4
} hipMemcpyKind;
typedef enum hipStreamCaptureMode { hipStreamCaptureModeGlobal = 0, hipStreamCaptureModeThreadLocal = 1,
                                    hipStreamCaptureModeRelaxed = 2 } hipStreamCaptureMode;
struct EmuStream;
struct EmuEvent;
struct EmuGraph;
typedef EmuStream* hipStream_t;
typedef EmuEvent* hipEvent_t;
typedef EmuGraph* hipGraph_t;
typedef EmuGraph* hipGraphExec_t;
typedef struct hipGraphNode* hipGraphNode_t;
#define hipEventDisableTiming 2
struct hipDeviceProp_t {
    char name[256];
    char gcnArchName[256];
    int multiProcessorCount;
    size_t sharedMemPerBlock;
    int warpSize;
};
hipError_t hipGetDeviceCount(int* n);
hipError_t hipSetDevice(int d);
hipError_t hipGetDevice(int* d);
hipError_t hipGetSymbolAddress(void** p, const void* symbol);
hipError_t hipGetDeviceProperties(hipDeviceProp_t* p, int d);
hipError_t hipMemGetInfo(size_t* f, size_t* t);
const char* hipGetErrorString(hipError_t e);
hipError_t hipGetLastError();
hipError_t hipPeekAtLastError();
hipError_t hipDeviceSynchronize();
hipError_t hipMalloc(void** p, size_t n);
template <class T> hipError_t hipMalloc(T** p, size_t n) { return hipMalloc((void**)p, n); }
hipError_t hipMallocAsync(void** p, size_t n, hipStream_t s);
template <class T> hipError_t hipMallocAsync(T** p, size_t n, hipStream_t s) { return hipMallocAsync((void**)p, n, s); }
hipError_t hipFree(void* p);
hipError_t hipFreeAsync(void* p, hipStream_t s);
hipError_t hipHostMalloc(void** p, size_t n, unsigned flags = 0);
template <class T> hipError_t hipHostMalloc(T** p, size_t n, unsigned flags = 0) { return hipHostMalloc((void**)p, n, flags); }
hipError_t hipHostFree(void* p);
hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind k);
hipError_t hipMemcpy2D(void* d, size_t dpitch, const void* s, size_t spitch, size_t width, size_t height,
                       hipMemcpyKind k);
hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind k, hipStream_t st = nullptr);
hipError_t hipMemset(void* d, int v, size_t n);
hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t st = nullptr);
hipError_t hipStreamCreate(hipStream_t* s);
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned f);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned f);
hipError_t hipEventCreate(hipEvent_t* e);
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned f);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s = nullptr);
hipError_t hipEventSynchronize(hipEvent_t e);
hipError_t hipEventQuery(hipEvent_t e);
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b);
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode m);
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t* g);
hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, hipGraphNode_t* n, char* log, size_t sz);
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t s);
hipError_t hipGraphDestroy(hipGraph_t g);
hipError_t hipGraphExecDestroy(hipGraphExec_t e);

void emu_launch(std::function<void()> fn, dim3 g, dim3 b);

template <class... P, class... A>
void emu_launch_kernel(void (*k)(P...), dim3 g, dim3 b, A&&... a) {
    std::tuple<std::decay_t<P>...> args(static_cast<std::decay_t<P>>(std::forward<A>(a))...);
    emu_launch([k, args]() { std::apply(k, args); }, g, b);
}
#define hipLaunchKernelGGL(K, G, B, SH, ST, ...) emu_launch_kernel(K, dim3(G), dim3(B), ##__VA_ARGS__)
#define HIP_SYMBOL(x) (&(x))
