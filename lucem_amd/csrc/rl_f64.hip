// Double-precision Richardson-Lucy for richardson_lucy_into::<f64> (restore/deconvolve.rs:45-149
// with F = f64): the reference's own structure, direct correlations and pointwise kernels, in
// FP64 on the device. The f32 form is the FFT hot path; this form gives f64 callers the same
// entry point with FP64 arithmetic throughout (the configurations are all f32).
//
// correlate2d_f64: out[y, x] = sum_{ky, kx} w[ky, kx] * P_mode(in)[y - ly + ky, x - lx + kx],
// accumulated ky-major as in filter/kernels/convolve.rs:61-118, Constant reading `cval` outside
// the array (0 in the Richardson-Lucy iteration, deconvolve.rs:99-103). A 16x16 output tile
// per workgroup reads its input tile with the halo into LDS once when it fits (kernels up to
// about 74 x 74), otherwise straight from global memory.
#include <hip/hip_runtime.h>

#include "fft_lds.hpp"
#include "rl_internal.hpp"

namespace lucem {

constexpr int kTile = 16;
constexpr size_t kF64TileLds = 65536;

template <bool STAGED>
__global__ __launch_bounds__(kTile* kTile) void correlate2d_f64_kernel(const double* __restrict__ in,
                                                                       const double* __restrict__ w,
                                                                       double* __restrict__ out, int H, int W, int kh,
                                                                       int kw, int ly, int lx, int mode,
                                                                       double cval) {
    extern __shared__ double tile[];
    const int tx = threadIdx.x % kTile, ty = threadIdx.x / kTile;
    const int x0 = blockIdx.x * kTile, y0 = blockIdx.y * kTile;
    const int x = x0 + tx, y = y0 + ty;
    if constexpr (STAGED) {
        const int tw = kTile + kw - 1, th = kTile + kh - 1;
        for (int i = threadIdx.x; i < tw * th; i += kTile * kTile) {
            const int r = i / tw, c = i - r * tw;
            const int py = pad_index(y0 + r - ly, H, mode), px = pad_index(x0 + c - lx, W, mode);
            tile[i] = (py < 0 || px < 0) ? cval : in[(int64_t)py * W + px];
        }
        __syncthreads();
        if (x < W && y < H) {
            double acc = 0.0;
            for (int ky = 0; ky < kh; ky++) {
                const double* row = tile + (ty + ky) * tw + tx;
                const double* wr = w + ky * kw;
                for (int kx = 0; kx < kw; kx++) acc += wr[kx] * row[kx];
            }
            out[(int64_t)y * W + x] = acc;
        }
    } else {
        if (x < W && y < H) {
            double acc = 0.0;
            for (int ky = 0; ky < kh; ky++) {
                const int py = pad_index(y - ly + ky, H, mode);
                for (int kx = 0; kx < kw; kx++) {
                    const int px = pad_index(x - lx + kx, W, mode);
                    const double s = (py < 0 || px < 0) ? cval : in[(int64_t)py * W + px];
                    acc += w[ky * kw + kx] * s;
                }
            }
            out[(int64_t)y * W + x] = acc;
        }
    }
}

// rel = obs / (|c| < eps ? eps : c)      restore/kernels/deconvolve.rs:7-22
__global__ void ratio_f64_kernel(const double* __restrict__ obs, const double* __restrict__ c, double* __restrict__ rel,
                                 double eps, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const double d = fabs(c[i]) < eps ? eps : c[i];
        rel[i] = obs[i] / d;
    }
}

// dst = src * err, clamped to [-1, 1] when asked   restore/kernels/deconvolve.rs:24-43
__global__ void update_f64_kernel(const double* __restrict__ src, const double* __restrict__ err,
                                  double* __restrict__ dst, int clip, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        double v = src[i] * err[i];
        if (clip) v = fmin(fmax(v, -1.0), 1.0);
        dst[i] = v;
    }
}

__global__ void fill_f64_kernel(double* __restrict__ p, double v, int64_t n) {
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

int launch_correlate2d_f64(const double* in, const double* w, double* out, int H, int W, int kh, int kw, int ly,
                           int lx, int mode, double cval, void* stream) {
    hipStream_t st = (hipStream_t)stream;
    const dim3 grid((unsigned)((W + kTile - 1) / kTile), (unsigned)((H + kTile - 1) / kTile));
    const size_t lds = (size_t)(kTile + kw - 1) * (kTile + kh - 1) * sizeof(double);
    if (lds <= kF64TileLds)
        hipLaunchKernelGGL(correlate2d_f64_kernel<true>, grid, dim3(kTile * kTile), lds, st, in, w, out, H, W, kh, kw,
                           ly, lx, mode, cval);
    else
        hipLaunchKernelGGL(correlate2d_f64_kernel<false>, grid, dim3(kTile * kTile), 0, st, in, w, out, H, W, kh, kw,
                           ly, lx, mode, cval);
    return (int)hipGetLastError();
}

int launch_ratio_f64(const double* obs, const double* c, double* rel, double eps, int64_t n, void* stream) {
    hipLaunchKernelGGL(ratio_f64_kernel, grid1(n), dim3(256), 0, (hipStream_t)stream, obs, c, rel, eps, n);
    return (int)hipGetLastError();
}

int launch_update_f64(const double* src, const double* err, double* dst, int clip, int64_t n, void* stream) {
    hipLaunchKernelGGL(update_f64_kernel, grid1(n), dim3(256), 0, (hipStream_t)stream, src, err, dst, clip, n);
    return (int)hipGetLastError();
}

int launch_fill_f64(double* p, double v, int64_t n, void* stream) {
    hipLaunchKernelGGL(fill_f64_kernel, grid1(n), dim3(256), 0, (hipStream_t)stream, p, v, n);
    return (int)hipGetLastError();
}

int launch_psf_prep_f64(const double* psf, double* h, double* rev, int64_t nk, void* stream) {
    hipLaunchKernelGGL(psf_prep_f64_kernel, dim3(1), dim3(256), 0, (hipStream_t)stream, psf, h, rev, nk);
    return (int)hipGetLastError();
}

}  // namespace lucem
