// Film -> image on the device: scene-linear sRGB (the PFM buffer) and the 8-bit sRGB
// tone map with the p99 auto-exposure anchor, from the int64 fixed-point film.
//
// Follows the reference's filmToLinear / filmToRgb8 (src/main.cpp:10282-10348) and
// xyzToLinearSrgb / srgbGamma (src/color.h:41-53), without the optional denoiser.
// Everything is computed in double with the reference's evaluation order and no FMA
// contraction, starting from the same doubles gp_film_download hands out
// ((double)q * (1 / scale)), so the output equals the CPU reduction of the downloaded
// film bit for bit.
//
// The anchor is an exact order statistic: the element at (size_t)(0.99 (n - 1)) of
// max(r, g, b, 0) over the pixels. It is found by an 8-pass radix select over the
// bit patterns of those non-negative doubles (which order like the values): each
// pass builds a 256-bin histogram of the next byte among the keys that match the
// bytes chosen so far, and one workgroup picks the bin holding the k-th key.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>

#include "gp_runtime.hpp"

namespace gp {
namespace {

constexpr double kAbsExposureGain = 6.0;   // absolute-EV sensor gain (main.cpp:10265)

// sum of ybar over 360..830 nm in 1-nm steps (color.h:32-39)
double cieYIntegralHost() {
    auto lobe = [](double x, double mu, double s1, double s2) {
        double t = (x - mu) * ((x < mu) ? s1 : s2);
        return std::exp(-0.5 * t * t);
    };
    double s = 0;
    for (double w = 360.0; w <= 830.0; w += 1.0)
        s += 0.821 * lobe(w, 568.8, 0.0213, 0.0247) + 0.286 * lobe(w, 530.9, 0.0613, 0.0322);
    return s;
}

__device__ __forceinline__ void linearPixel(const long long* __restrict__ film, size_t i, double inv,
                                            double norm, double& r, double& g, double& b) {
#pragma clang fp contract(off)
    const double x = ((double)film[4 * i + 0] * inv) * norm;
    const double y = ((double)film[4 * i + 1] * inv) * norm;
    const double z = ((double)film[4 * i + 2] * inv) * norm;
    r = 3.2406 * x - 1.5372 * y - 0.4986 * z;
    g = -0.9689 * x + 1.8758 * y + 0.0415 * z;
    b = 0.0557 * x - 0.2040 * y + 1.0570 * z;
}

__device__ __forceinline__ double srgbGamma(double c) {
#pragma clang fp contract(off)
    c = c > 0.0 ? c : 0.0;
    return (c <= 0.0031308) ? 12.92 * c : 1.055 * pow(c, 1.0 / 2.4) - 0.055;
}

// linear sRGB as float (PFM order: the film's own rows) and the anchor keys
__global__ void k_linear(const long long* __restrict__ film, size_t npix, double inv, double norm,
                         float* __restrict__ lin, unsigned long long* __restrict__ keys) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < npix; i += stride) {
        double r, g, b;
        linearPixel(film, i, inv, norm, r, g, b);
        if (lin) {
            lin[3 * i + 0] = (float)r;
            lin[3 * i + 1] = (float)g;
            lin[3 * i + 2] = (float)b;
        }
        if (keys) {
            double m = r > g ? r : g;
            m = m > b ? m : b;
            // every value <= 0 is the same key (+0.0), so -0.0 cannot sort last
            keys[i] = m > 0.0 ? (unsigned long long)__double_as_longlong(m) : 0ull;
        }
    }
}

struct SelectState {
    unsigned long long prefix;   // bytes chosen so far
    unsigned long long mask;     // which bytes are chosen
    unsigned long long k;        // rank still sought among keys matching the prefix
    unsigned long long pad;
};

__global__ void __launch_bounds__(256) k_select_hist(const unsigned long long* __restrict__ keys, size_t n,
                                                     const SelectState* st, int shift,
                                                     unsigned* __restrict__ hist) {
    __shared__ unsigned h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const unsigned long long prefix = st->prefix, mask = st->mask;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const unsigned long long key = keys[i];
        if ((key & mask) == prefix) atomicAdd(&h[(key >> shift) & 255u], 1u);
    }
    __syncthreads();
    if (h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], h[threadIdx.x]);
}

// One workgroup of 256: find the bin holding rank k, narrow the prefix, clear hist.
__global__ void __launch_bounds__(256) k_select_pick(SelectState* st, int shift, unsigned* hist) {
    __shared__ unsigned long long scan[256];
    const int t = threadIdx.x;
    const unsigned c = hist[t];
    scan[t] = c;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const unsigned long long v = t >= off ? scan[t - off] : 0ull;
        __syncthreads();
        scan[t] += v;
        __syncthreads();
    }
    const unsigned long long k = st->k;
    const unsigned long long before = t > 0 ? scan[t - 1] : 0ull;
    __syncthreads();
    if (c > 0 && before <= k && k < scan[t]) {
        st->k = k - before;
        st->prefix |= (unsigned long long)t << shift;
        st->mask |= 255ull << shift;
    }
    hist[t] = 0;
}

// exposure from the anchor (main.cpp:10305-10338); out[0] = eAuto, out[1] = exposure
__global__ void k_exposure(const SelectState* st, double anchor, double expComp, int absolute, double* out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double eAuto;
    if (absolute) {
        eAuto = kAbsExposureGain;
    } else if (anchor > 0.0) {
        eAuto = anchor;
    } else {
        const double p99 = __longlong_as_double((long long)st->prefix);
        eAuto = (p99 > 0) ? 0.9 / p99 : 1.0;
    }
    out[0] = eAuto;
    out[1] = eAuto * (expComp > 0.0 ? expComp : 1.0);
}

// 8-bit sRGB, row 0 = image top (the film's last row)
__global__ void k_rgb8(const long long* __restrict__ film, int W, int H, double inv, double norm,
                       const double* __restrict__ expo, unsigned char* __restrict__ out) {
#pragma clang fp contract(off)
    const double exposure = expo[1];
    const size_t npix = (size_t)W * H;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t d = (size_t)blockIdx.x * blockDim.x + threadIdx.x; d < npix; d += stride) {
        const int y = (int)(d / (size_t)W), x = (int)(d % (size_t)W);
        const size_t src = (size_t)(H - 1 - y) * W + x;
        double c[3];
        linearPixel(film, src, inv, norm, c[0], c[1], c[2]);
        for (int k = 0; k < 3; ++k) {
            double v = srgbGamma(c[k] * exposure) * 255.0 + 0.5;
            v = v < 0.0 ? 0.0 : (v > 255.0 ? 255.0 : v);
            out[3 * d + k] = (unsigned char)v;
        }
    }
}

// Starts a select: the rank sought, empty prefix, zeroed histogram (a kernel rather
// than a host copy and memset, so the allocation-free entry point records only
// kernels on the stream).
__global__ void __launch_bounds__(256) k_select_init(SelectState* st, unsigned* hist, unsigned long long k) {
    hist[threadIdx.x] = 0;
    if (threadIdx.x == 0) *st = SelectState{0, 0, k, 0};
}

struct DevMem {
    void* p = nullptr;
    ~DevMem() { if (p) hipFree(p); }
};

// Scratch memory of the allocation-free tone map (gp_film_tonemap_noalloc): the
// select state, histogram and exposure, then the anchor keys (8 B per pixel) and the
// 8-bit image (3 B per pixel) of a film of at most kNoAllocMaxPixels pixels. Each GPU
// holds its own copy for the life of the process; host code reaches it only through
// the address hipGetSymbolAddress returns on that device (noAllocScratch).
constexpr size_t kNoAllocMaxPixels = size_t(1) << 22;   // 2048 x 2048
constexpr size_t kNoAllocHead = 4096;                   // SelectState, 256 bins, 2 doubles
constexpr size_t kNoAllocKeys = kNoAllocMaxPixels * sizeof(unsigned long long);
constexpr size_t kNoAllocBytes = kNoAllocHead + kNoAllocKeys + kNoAllocMaxPixels * 3;
static_assert(sizeof(SelectState) + 256 * sizeof(unsigned) + 2 * sizeof(double) <= kNoAllocHead, "head");
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kNoAllocMaxDevices = 64;
std::mutex gScratchMu;
void* gScratchPtr[kNoAllocMaxDevices] = {};

// The device address of `scratch` on the current device, looked up once per device
// (with that device current) and kept.
gp_status noAllocScratch(int dev, unsigned char** out) {
    if (dev < 0 || dev >= kNoAllocMaxDevices) return fail(GP_ERR_NO_DEVICE, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(gScratchMu);
    if (!gScratchPtr[dev]) {
        int cur = -1;
        GP_HIP(hipGetDevice(&cur));
        if (cur != dev) GP_HIP(hipSetDevice(dev));
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != dev) hipSetDevice(cur);
        if (e != hipSuccess || !p)
            return fail(GP_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        gScratchPtr[dev] = p;
    }
    *out = (unsigned char*)gScratchPtr[dev];
    return GP_OK;
}

gp_status filmNorm(const gp_film* f, int64_t nPhotons, double& inv, double& norm) {
    if (!f) return fail(GP_ERR_INVALID, "null film");
    if (nPhotons <= 0) return fail(GP_ERR_INVALID, "n_photons must be positive");
    inv = f->scale > 0 ? 1.0 / f->scale : 0.0;
    norm = 1.0 / ((double)nPhotons * cieYIntegralHost());
    return GP_OK;
}

int gridFor(size_t n) {
    size_t b = (n + 255) / 256;
    return (int)(b < 4096 ? (b > 0 ? b : 1) : 4096);
}

}  // namespace
}  // namespace gp

using namespace gp;

extern "C" {

gp_status gp_film_to_linear(gp_film* film, int64_t n_photons, float* rgb) {
    double inv, norm;
    gp_status st = filmNorm(film, n_photons, inv, norm);
    if (st != GP_OK) return st;
    if (!rgb) return fail(GP_ERR_INVALID, "null output");
    if ((st = requireDevice()) != GP_OK) return st;
    const size_t npix = (size_t)film->rx * film->ry;
    hipStream_t s = device().stream;
    DevMem lin;
    GP_HIP(hipMalloc(&lin.p, npix * 3 * sizeof(float)));
    hipLaunchKernelGGL(k_linear, dim3(gridFor(npix)), dim3(256), 0, s, (const long long*)film->d, npix, inv,
                       norm, (float*)lin.p, (unsigned long long*)nullptr);
    GP_HIP(hipGetLastError());
    GP_HIP(hipMemcpyAsync(rgb, lin.p, npix * 3 * sizeof(float), hipMemcpyDeviceToHost, s));
    GP_HIP(hipStreamSynchronize(s));
    return GP_OK;
}

gp_status gp_film_tonemap(gp_film* film, int64_t n_photons, double exposure_comp, int32_t absolute,
                          double* anchor, uint8_t* rgb8, double* exposure_out) {
    double inv, norm;
    gp_status st = filmNorm(film, n_photons, inv, norm);
    if (st != GP_OK) return st;
    if (!rgb8) return fail(GP_ERR_INVALID, "null output");
    if ((st = requireDevice()) != GP_OK) return st;
    const int W = film->rx, H = film->ry;
    const size_t npix = (size_t)W * H;
    hipStream_t s = device().stream;
    const bool select = !absolute && !(anchor && *anchor > 0.0);
    DevMem keys, scratch, img;
    GP_HIP(hipMalloc(&scratch.p, sizeof(SelectState) + 256 * sizeof(unsigned) + 2 * sizeof(double)));
    SelectState* dst = (SelectState*)scratch.p;
    unsigned* hist = (unsigned*)((char*)scratch.p + sizeof(SelectState));
    double* expo = (double*)((char*)hist + 256 * sizeof(unsigned));
    SelectState h0{0, 0, (unsigned long long)(size_t)(0.99 * (double)(npix - 1)), 0};
    GP_HIP(hipMemcpyAsync(dst, &h0, sizeof h0, hipMemcpyHostToDevice, s));
    GP_HIP(hipMemsetAsync(hist, 0, 256 * sizeof(unsigned), s));
    if (select) {
        GP_HIP(hipMalloc(&keys.p, npix * sizeof(unsigned long long)));
        hipLaunchKernelGGL(k_linear, dim3(gridFor(npix)), dim3(256), 0, s, (const long long*)film->d, npix,
                           inv, norm, (float*)nullptr, (unsigned long long*)keys.p);
        for (int pass = 0; pass < 8; ++pass) {
            const int shift = 56 - 8 * pass;
            hipLaunchKernelGGL(k_select_hist, dim3(gridFor(npix)), dim3(256), 0, s,
                               (const unsigned long long*)keys.p, npix, (const SelectState*)dst, shift, hist);
            hipLaunchKernelGGL(k_select_pick, dim3(1), dim3(256), 0, s, dst, shift, hist);
        }
    }
    hipLaunchKernelGGL(k_exposure, dim3(1), dim3(64), 0, s, (const SelectState*)dst,
                       anchor ? *anchor : 0.0, exposure_comp, (int)absolute, expo);
    GP_HIP(hipMalloc(&img.p, npix * 3));
    hipLaunchKernelGGL(k_rgb8, dim3(gridFor(npix)), dim3(256), 0, s, (const long long*)film->d, W, H, inv,
                       norm, (const double*)expo, (unsigned char*)img.p);
    GP_HIP(hipGetLastError());
    double e[2];
    GP_HIP(hipMemcpyAsync(rgb8, img.p, npix * 3, hipMemcpyDeviceToHost, s));
    GP_HIP(hipMemcpyAsync(e, expo, sizeof e, hipMemcpyDeviceToHost, s));
    GP_HIP(hipStreamSynchronize(s));
    if (anchor && !absolute && !(*anchor > 0.0)) *anchor = e[0];
    if (exposure_out) *exposure_out = e[1];
    return GP_OK;
}

gp_status gp_film_tonemap_noalloc(gp_film* film, int64_t n_photons, double exposure_comp, int32_t absolute,
                                  double* anchor, uint8_t* rgb8, double* exposure_out) {
    double inv, norm;
    gp_status st = filmNorm(film, n_photons, inv, norm);
    if (st != GP_OK) return st;
    if (!rgb8) return fail(GP_ERR_INVALID, "null output");
    if ((st = requireDevice()) != GP_OK) return st;
    const int W = film->rx, H = film->ry;
    const size_t npix = (size_t)W * H;
    if (npix > kNoAllocMaxPixels)
        return fail(GP_ERR_CAPACITY, "gp_film_tonemap_noalloc: film of " + std::to_string(npix) +
                                         " pixels exceeds its scratch limit of 4194304 (2048 x 2048)");
    unsigned char* base = nullptr;
    if ((st = noAllocScratch(device().id, &base)) != GP_OK) return st;
    SelectState* dst = (SelectState*)base;
    unsigned* hist = (unsigned*)(base + sizeof(SelectState));
    double* expo = (double*)((char*)hist + 256 * sizeof(unsigned));
    unsigned long long* keys = (unsigned long long*)(base + kNoAllocHead);
    unsigned char* img = base + kNoAllocHead + kNoAllocKeys;
    hipStream_t s = device().stream;
    const bool select = !absolute && !(anchor && *anchor > 0.0);
    // the previous call's state is still in scratch: start the select afresh
    hipLaunchKernelGGL(k_select_init, dim3(1), dim3(256), 0, s, dst, hist,
                       (unsigned long long)(size_t)(0.99 * (double)(npix - 1)));
    if (select) {
        hipLaunchKernelGGL(k_linear, dim3(gridFor(npix)), dim3(256), 0, s, (const long long*)film->d, npix,
                           inv, norm, (float*)nullptr, keys);
        for (int pass = 0; pass < 8; ++pass) {
            const int shift = 56 - 8 * pass;
            hipLaunchKernelGGL(k_select_hist, dim3(gridFor(npix)), dim3(256), 0, s,
                               (const unsigned long long*)keys, npix, (const SelectState*)dst, shift, hist);
            hipLaunchKernelGGL(k_select_pick, dim3(1), dim3(256), 0, s, dst, shift, hist);
        }
    }
    hipLaunchKernelGGL(k_exposure, dim3(1), dim3(64), 0, s, (const SelectState*)dst,
                       anchor ? *anchor : 0.0, exposure_comp, (int)absolute, expo);
    hipLaunchKernelGGL(k_rgb8, dim3(gridFor(npix)), dim3(256), 0, s, (const long long*)film->d, W, H, inv,
                       norm, (const double*)expo, img);
    GP_HIP(hipGetLastError());
    double e[2];
    GP_HIP(hipMemcpyAsync(rgb8, img, npix * 3, hipMemcpyDeviceToHost, s));
    GP_HIP(hipMemcpyAsync(e, expo, sizeof e, hipMemcpyDeviceToHost, s));
    GP_HIP(hipStreamSynchronize(s));
    if (anchor && !absolute && !(*anchor > 0.0)) *anchor = e[0];
    if (exposure_out) *exposure_out = e[1];
    return GP_OK;
}

}  // extern "C"
