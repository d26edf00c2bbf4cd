// The direct form of Richardson-Lucy and of the correlation: the reference's own structure
// (restore/deconvolve.rs:45-149, filter/kernels/convolve.rs:61-118), direct correlations and
// pointwise kernels, for F = f64 and for f32 arrays with an axis longer than the FFT passes take.
// The FFT passes are the f32 hot path; this form gives f64 callers FP64 arithmetic throughout,
// and keeps the f32 entry points open for every 2-D shape the reference accepts.
//
// correlate2d: out[y, x] = sum_{ky, kx} w[ky, kx] * P_mode(in)[y - ly + ky, x - lx + kx],
// accumulated in T, ky-major as in filter/kernels/convolve.rs:61-118, Constant reading `cval`
// outside the array (0 in the Richardson-Lucy iteration, deconvolve.rs:99-103). A 16x16 output
// tile per workgroup reads its input tile with the halo into LDS once when it fits (64 KiB),
// otherwise straight from global memory; arrays 64 or more wide take 64x16 tiles when those fit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>

#include "fft_lds.hpp"
#include "rl_internal.hpp"

namespace lucem {

constexpr int kTile = 16;
constexpr size_t kDirectTileLds = 65536;

template <typename T, bool STAGED>
__global__ __launch_bounds__(kTile* kTile) void correlate2d_kernel(const T* __restrict__ in, const T* __restrict__ w,
                                                                   T* __restrict__ out, int H, int W, int kh, int kw,
                                                                   int ly, int lx, int mode, T cval) {
    extern __shared__ __align__(16) unsigned char tile_raw[];
    T* tile = reinterpret_cast<T*>(tile_raw);
    // output tiles row-major, the blocks striding over them (any array height or width)
    const int nbx = (int)(((int64_t)W + kTile - 1) / kTile);
    const int64_t ntiles = (int64_t)nbx * (((int64_t)H + kTile - 1) / kTile);
    const int tx = threadIdx.x % kTile, ty = threadIdx.x / kTile;
    for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int bx = (int)(t % nbx), by = (int)(t / nbx);
        const int x0 = bx * kTile, y0 = by * kTile;
        const int x = x0 + tx, y = y0 + ty;
        if constexpr (STAGED) {
            const int tw = kTile + kw - 1, th = kTile + kh - 1;
            __syncthreads();  // the previous tile's sums have read the staged values
            for (int i = threadIdx.x; i < tw * th; i += kTile * kTile) {
                const int r = i / tw, c = i - r * tw;
                const int py = pad_index(y0 + r - ly, H, mode), px = pad_index(x0 + c - lx, W, mode);
                tile[i] = (py < 0 || px < 0) ? cval : in[(int64_t)py * W + px];
            }
            __syncthreads();
            if (x < W && y < H) {
                T acc = T(0);
                for (int ky = 0; ky < kh; ky++) {
                    const T* row = tile + (ty + ky) * tw + tx;
                    const T* wr = w + ky * kw;
                    for (int kx = 0; kx < kw; kx++) acc += wr[kx] * row[kx];
                }
                out[(int64_t)y * W + x] = acc;
            }
        } else {
            if (x < W && y < H) {
                T acc = T(0);
                for (int ky = 0; ky < kh; ky++) {
                    const int py = pad_index(y - ly + ky, H, mode);
                    for (int kx = 0; kx < kw; kx++) {
                        const int px = pad_index(x - lx + kx, W, mode);
                        const T s = (py < 0 || px < 0) ? cval : in[(int64_t)py * W + px];
                        acc += w[ky * kw + kx] * s;
                    }
                }
                out[(int64_t)y * W + x] = acc;
            }
        }
    }
}

// The same sum over a 64x16 output tile: each thread keeps four outputs 16 apart along x, so the
// halo is staged once per 1024 outputs instead of per 256. Each output still sums ky-major in
// window order, so the result is the 16x16 form's bit for bit.
constexpr int kWide = 64;

template <typename T>
__global__ __launch_bounds__(kTile* kTile) void correlate2d_wide_kernel(const T* __restrict__ in,
                                                                        const T* __restrict__ w, T* __restrict__ out,
                                                                        int H, int W, int kh, int kw, int ly, int lx,
                                                                        int mode, T cval) {
    extern __shared__ __align__(16) unsigned char tile_raw[];
    T* tile = reinterpret_cast<T*>(tile_raw);
    const int nbx = (int)(((int64_t)W + kWide - 1) / kWide);
    const int64_t ntiles = (int64_t)nbx * (((int64_t)H + kTile - 1) / kTile);
    const int tx = threadIdx.x % kTile, ty = threadIdx.x / kTile;
    const int tw = kWide + kw - 1, th = kTile + kh - 1;
    for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int bx = (int)(t % nbx), by = (int)(t / nbx);
        const int x0 = bx * kWide, y0 = by * kTile;
        const int y = y0 + ty;
        __syncthreads();  // the previous tile's sums have read the staged values
        for (int i = threadIdx.x; i < tw * th; i += kTile * kTile) {
            const int r = i / tw, c = i - r * tw;
            const int py = pad_index(y0 + r - ly, H, mode), px = pad_index(x0 + c - lx, W, mode);
            tile[i] = (py < 0 || px < 0) ? cval : in[(int64_t)py * W + px];
        }
        __syncthreads();
        if (y < H) {
            T acc[4] = {T(0), T(0), T(0), T(0)};
            for (int ky = 0; ky < kh; ky++) {
                const T* row = tile + (ty + ky) * tw + tx;
                const T* wr = w + ky * kw;
                for (int kx = 0; kx < kw; kx++) {
                    const T wt = wr[kx];
#pragma unroll
                    for (int u = 0; u < 4; u++) acc[u] += wt * row[kx + kTile * u];
                }
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int x = x0 + tx + kTile * u;
                if (x < W) out[(int64_t)y * W + x] = acc[u];
            }
        }
    }
}

// rel = obs / (|c| < eps ? eps : c)      restore/kernels/deconvolve.rs:7-22
template <typename T>
__global__ void ratio_kernel(const T* __restrict__ obs, const T* __restrict__ c, T* __restrict__ rel, T eps,
                             int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const T v = c[i];
        const T d = (v < T(0) ? -v : v) < eps ? eps : v;
        rel[i] = obs[i] / d;
    }
}

// dst = src * err, clamped to [-1, 1] when asked   restore/kernels/deconvolve.rs:24-43
template <typename T>
__global__ void update_kernel(const T* __restrict__ src, const T* __restrict__ err, T* __restrict__ dst, int clip,
                              int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        T v = src[i] * err[i];
        if (clip) v = v < T(-1) ? T(-1) : (v > T(1) ? T(1) : v);
        dst[i] = v;
    }
}

template <typename T>
__global__ void fill_kernel(T* __restrict__ p, T v, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// PSF normalisation and flip (deconvolve.rs:73-90) in f64: one workgroup, fixed reduction order.
__global__ __launch_bounds__(256) void psf_prep_f64_kernel(const double* __restrict__ psf, double* __restrict__ h,
                                                           double* __restrict__ rev, int64_t nk) {
    __shared__ double part[256];
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < nk; i += 256) s += psf[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    const double sum = part[0];
    const double norm = fabs(sum) > 1e-12 ? 1.0 / sum : 1.0;
    for (int64_t i = threadIdx.x; i < nk; i += 256) {
        const double v = psf[i] * norm;
        h[i] = v;
        rev[nk - 1 - i] = v;
    }
}

static dim3 grid1(int64_t n) { return dim3((unsigned)((n + 255) / 256)); }

// Blocks of a direct-correlation launch: one per output tile up to 2^20 (256 threads each), which
// then stride over the rest.
static dim3 tile_grid(int64_t ntiles) { return dim3((unsigned)std::min<int64_t>(std::max<int64_t>(ntiles, 1), 1 << 20)); }

// LUCEM_DIRECT_TILE=16 keeps the 16x16 output tiles where the 64x16 ones would run (the two sum
// every output in the same order, so they agree bit for bit; tests compare them).
static bool wide_tiles_allowed() {
    const char* v = std::getenv("LUCEM_DIRECT_TILE");
    return !(v && std::atoi(v) == 16);
}

template <typename T>
static int correlate2d(const T* in, const T* w, T* out, int H, int W, int kh, int kw, int ly, int lx, int mode, T cval,
                       void* stream) {
    hipStream_t st = (hipStream_t)stream;
    const size_t wide_lds = (size_t)(kWide + kw - 1) * (kTile + kh - 1) * sizeof(T);
    if (W >= kWide && wide_lds <= kDirectTileLds && wide_tiles_allowed()) {
        const dim3 wgrid = tile_grid(((int64_t)W + kWide - 1) / kWide * (((int64_t)H + kTile - 1) / kTile));
        hipLaunchKernelGGL(correlate2d_wide_kernel<T>, wgrid, dim3(kTile * kTile), wide_lds, st, in, w, out, H, W, kh,
                           kw, ly, lx, mode, cval);
        return (int)hipGetLastError();
    }
    const dim3 grid = tile_grid(((int64_t)W + kTile - 1) / kTile * (((int64_t)H + kTile - 1) / kTile));
    const size_t lds = (size_t)(kTile + kw - 1) * (kTile + kh - 1) * sizeof(T);
    if (lds <= kDirectTileLds)
        hipLaunchKernelGGL((correlate2d_kernel<T, true>), grid, dim3(kTile * kTile), lds, st, in, w, out, H, W, kh, kw,
                           ly, lx, mode, cval);
    else
        hipLaunchKernelGGL((correlate2d_kernel<T, false>), grid, dim3(kTile * kTile), 0, st, in, w, out, H, W, kh, kw,
                           ly, lx, mode, cval);
    return (int)hipGetLastError();
}

int launch_correlate2d_direct(const double* in, const double* w, double* out, int H, int W, int kh, int kw, int ly,
                              int lx, int mode, double cval, void* stream) {
    return correlate2d<double>(in, w, out, H, W, kh, kw, ly, lx, mode, cval, stream);
}

int launch_correlate2d_direct(const float* in, const float* w, float* out, int H, int W, int kh, int kw, int ly,
                              int lx, int mode, float cval, void* stream) {
    return correlate2d<float>(in, w, out, H, W, kh, kw, ly, lx, mode, cval, stream);
}

template <typename T>
static int ratio(const T* obs, const T* c, T* rel, T eps, int64_t n, void* stream) {
    hipLaunchKernelGGL(ratio_kernel<T>, grid1(n), dim3(256), 0, (hipStream_t)stream, obs, c, rel, eps, n);
    return (int)hipGetLastError();
}
int launch_ratio_direct(const double* obs, const double* c, double* rel, double eps, int64_t n, void* stream) {
    return ratio<double>(obs, c, rel, eps, n, stream);
}
int launch_ratio_direct(const float* obs, const float* c, float* rel, float eps, int64_t n, void* stream) {
    return ratio<float>(obs, c, rel, eps, n, stream);
}

template <typename T>
static int update(const T* src, const T* err, T* dst, int clip, int64_t n, void* stream) {
    hipLaunchKernelGGL(update_kernel<T>, grid1(n), dim3(256), 0, (hipStream_t)stream, src, err, dst, clip, n);
    return (int)hipGetLastError();
}
int launch_update_direct(const double* src, const double* err, double* dst, int clip, int64_t n, void* stream) {
    return update<double>(src, err, dst, clip, n, stream);
}
int launch_update_direct(const float* src, const float* err, float* dst, int clip, int64_t n, void* stream) {
    return update<float>(src, err, dst, clip, n, stream);
}

template <typename T>
static int fill(T* p, T v, int64_t n, void* stream) {
    hipLaunchKernelGGL(fill_kernel<T>, grid1(n), dim3(256), 0, (hipStream_t)stream, p, v, n);
    return (int)hipGetLastError();
}
int launch_fill_direct(double* p, double v, int64_t n, void* stream) { return fill<double>(p, v, n, stream); }
int launch_fill_direct(float* p, float v, int64_t n, void* stream) { return fill<float>(p, v, n, stream); }

int launch_psf_prep_f64(const double* psf, double* h, double* rev, int64_t nk, void* stream) {
    hipLaunchKernelGGL(psf_prep_f64_kernel, dim3(1), dim3(256), 0, (hipStream_t)stream, psf, h, rev, nk);
    return (int)hipGetLastError();
}

}  // namespace lucem
