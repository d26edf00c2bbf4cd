// Wiener and unsupervised Wiener deconvolution (restore/deconvolve.rs:184-315 and :330-509) on
// the circular H x W domain of the image.
//
// The image spectrum comes from the FFT passes of the RL path (row R2C, line FFTs along y). The
// PSF's transfer function is a direct FP64 DFT of the small kernel, placed with its centre
// (kh/2, kw/2) at the origin as fft_shift_2d_kernel does (restore/kernels/deconvolve.rs:76-99):
//   H(u, v) = sum_{ky,kx} psf[ky, kx] exp(-2 pi i (u (ky - kh/2) / Hn + v (kx - kw/2) / Wn)),
// split into a DFT of every PSF row along x and then a sum over the rows, with the phases read
// from FP64 tables indexed by the exact integer products mod Hn and Wn. The Laplacian
// regulariser [0,-1,0; -1,4,-1; 0,-1,0] centred the same way has the transform
// 4 - 2 cos(2 pi u / Hn) - 2 cos(2 pi v / Wn) (deconvolve.rs:266-281; a 1x1 kernel of 1 when an
// axis is shorter than 3).
//
// Only the half spectrum v <= Wn/2 is stored. Every per-frequency term of the unsupervised loop
// is even in (u, v) for a real image and PSF, so sums over the full spectrum are half-spectrum
// sums with weight 2 on the columns that have a mirror (0 < v < Wn/2) and 1 on the others.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "rl_internal.hpp"

namespace lucem {

constexpr int kWienerThreads = 256;

// exp(-2 pi i m / N), m < N, in FP64.
__global__ void dft_table_kernel(double2* __restrict__ tw, int N) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < N) {
        double s, c;
        sincospi(2.0 * (double)m / (double)N, &s, &c);
        tw[m] = make_double2(c, -s);
    }
}

// A[ky, v] = sum_kx psf[ky, kx] exp(-2 pi i v (kx - kw/2) / Wn), v < Lc.
__global__ void psf_row_dft_kernel(const float* __restrict__ psf, int kh, int kw, const double2* __restrict__ tw_w,
                                   int Wn, int Lc, double2* __restrict__ A) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= kh * Lc) return;
    const int ky = idx / Lc, v = idx - ky * Lc;
    const int cx = kw / 2;
    int m = (int)((Wn - ((int64_t)v * cx) % Wn) % Wn);  // v * (0 - cx) mod Wn
    double re = 0.0, im = 0.0;
    const float* row = psf + (int64_t)ky * kw;
    for (int kx = 0; kx < kw; kx++) {
        const double p = (double)row[kx];
        const double2 t = tw_w[m];
        re += p * t.x;
        im += p * t.y;
        m += v;
        if (m >= Wn) m -= Wn;
    }
    A[idx] = make_double2(re, im);
}

__device__ __forceinline__ double laplacian_mag2(const double2* tw_h, const double2* tw_w, int Hn, int Wn, int u,
                                                 int v) {
    if (Hn < 3 || Wn < 3) return 1.0;
    const double l = 4.0 - 2.0 * tw_h[u].x - 2.0 * tw_w[v].x;
    return l * l;
}

// H(u, v) = sum_ky A[ky, v] exp(-2 pi i u (ky - kh/2) / Hn) for u < Hn, v < Lc. Each thread
// takes kOtfRows rows u = u0 + j of one column v, so every A[ky, v] it loads serves them all.
// FILTER: store the Wiener filter conj(H) / max(|H|^2 + balance |L|^2, 1e-12) times `scale` as
// float2 at G[u * pitch + v] (wiener_freq_filter_kernel, restore/kernels/deconvolve.rs:44-73);
// otherwise store H itself (FP64) at Hout[u * Lc + v].
constexpr int kOtfRows = 8;

template <bool FILTER>
__global__ void psf_otf_kernel(const double2* __restrict__ A, int kh, const double2* __restrict__ tw_h,
                               const double2* __restrict__ tw_w, int Hn, int Wn, int Lc, int pitch, double balance,
                               double scale, float2* __restrict__ G, double2* __restrict__ Hout) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t ublocks = (Hn + kOtfRows - 1) / kOtfRows;
    if (idx >= ublocks * Lc) return;
    const int ub = (int)(idx / Lc), v = (int)(idx - (int64_t)ub * Lc);
    const int u0 = ub * kOtfRows;
    const int cy = kh / 2;
    int m[kOtfRows];
    double re[kOtfRows], im[kOtfRows];
#pragma unroll
    for (int j = 0; j < kOtfRows; j++) {
        const int u = min(u0 + j, Hn - 1);
        m[j] = (int)((Hn - ((int64_t)u * cy) % Hn) % Hn);
        re[j] = 0.0;
        im[j] = 0.0;
    }
    for (int ky = 0; ky < kh; ky++) {
        const double2 a = A[(int64_t)ky * Lc + v];
#pragma unroll
        for (int j = 0; j < kOtfRows; j++) {
            const double2 t = tw_h[m[j]];
            re[j] += a.x * t.x - a.y * t.y;
            im[j] += a.x * t.y + a.y * t.x;
            m[j] += min(u0 + j, Hn - 1);
            if (m[j] >= Hn) m[j] -= Hn;
        }
    }
#pragma unroll
    for (int j = 0; j < kOtfRows; j++) {
        const int u = u0 + j;
        if (u >= Hn) break;
        if constexpr (FILTER) {
            const double d =
                fmax(re[j] * re[j] + im[j] * im[j] + balance * laplacian_mag2(tw_h, tw_w, Hn, Wn, u, v), 1.0e-12);
            const double f = scale / d;
            G[(int64_t)u * pitch + v] = make_float2((float)(re[j] * f), (float)(-im[j] * f));
        } else {
            Hout[(int64_t)u * Lc + v] = make_double2(re[j], im[j]);
        }
    }
}

// Fixed-order workgroup sum of NV doubles per thread into part[blockIdx.x * NV + j].
template <int NV>
__device__ __forceinline__ void block_sums(double (&v)[NV], double* __restrict__ part) {
    __shared__ double red[NV][kWienerThreads];
#pragma unroll
    for (int j = 0; j < NV; j++) red[j][threadIdx.x] = v[j];
    __syncthreads();
    for (int w = kWienerThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
#pragma unroll
            for (int j = 0; j < NV; j++) red[j][threadIdx.x] += red[j][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int j = 0; j < NV; j++) part[blockIdx.x * NV + j] = red[j][0];
    }
}

// The same over the partials of `nb` workgroups (one workgroup), result in out[j] of thread 0.
template <int NV>
__device__ __forceinline__ void final_sums(const double* __restrict__ part, int nb, double (&out)[NV]) {
    double v[NV];
#pragma unroll
    for (int j = 0; j < NV; j++) v[j] = 0.0;
    for (int b = threadIdx.x; b < nb; b += kWienerThreads) {
#pragma unroll
        for (int j = 0; j < NV; j++) v[j] += part[b * NV + j];
    }
    __shared__ double red[NV][kWienerThreads];
#pragma unroll
    for (int j = 0; j < NV; j++) red[j][threadIdx.x] = v[j];
    __syncthreads();
    for (int w = kWienerThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
#pragma unroll
            for (int j = 0; j < NV; j++) red[j][threadIdx.x] += red[j][threadIdx.x + w];
        }
        __syncthreads();
    }
#pragma unroll
    for (int j = 0; j < NV; j++) out[j] = red[j][0];
}

// (sum x, sum x^2) partials of the image, for its variance (reduce::mean_var with ddof 0).
__global__ __launch_bounds__(kWienerThreads) void moments_kernel(const float* __restrict__ x, int64_t n,
                                                                  double* __restrict__ part) {
    double v[2] = {0.0, 0.0};
    for (int64_t i = (int64_t)blockIdx.x * kWienerThreads + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kWienerThreads) {
        const double a = (double)x[i];
        v[0] += a;
        v[1] += a * a;
    }
    block_sums<2>(v, part);
}

// state[0] = balance (starts at 0.05, deconvolve.rs:424), state[1] = the image variance.
__global__ __launch_bounds__(kWienerThreads) void uw_init_kernel(const double* __restrict__ part, int nb, double n,
                                                                  double2* __restrict__ mean, int64_t nmean,
                                                                  double* __restrict__ state) {
    if (blockIdx.x == 0) {
        double s[2];
        final_sums<2>(part, nb, s);
        if (threadIdx.x == 0) {
            const double mu = s[0] / n;
            state[0] = 0.05;
            state[1] = fmax(s[1] / n - mu * mu, 0.0);
        }
    }
    for (int64_t i = (int64_t)blockIdx.x * kWienerThreads + threadIdx.x; i < nmean;
         i += (int64_t)gridDim.x * kWienerThreads)
        mean[i] = make_double2(0.0, 0.0);
}

// One sweep of the Gibbs / empirical-Bayes loop (deconvolve.rs:433-458) at the current balance:
// X = conj(H) Y / max(|H|^2 + b |L|^2, 1e-12); partial sums of |Y - H X|^2 / n,
// |L|^2 |X|^2 / n and |H|^2 / denom; X added to the posterior sum after burn-in. `spec` holds the
// image's half spectrum as the FFT passes leave it (2 Y).
__global__ __launch_bounds__(kWienerThreads) void uw_sweep_kernel(const float2* __restrict__ spec, int pitch,
                                                                   const double2* __restrict__ Hc,
                                                                   const double2* __restrict__ tw_h,
                                                                   const double2* __restrict__ tw_w, int Hn, int Wn,
                                                                   int Lc, double n, const double* __restrict__ state,
                                                                   int post, int full, double2* __restrict__ mean,
                                                                   double* __restrict__ part) {
    const double balance = state[0];
    const int64_t total = (int64_t)Hn * Lc;
    double v3[3] = {0.0, 0.0, 0.0};
    for (int64_t i = (int64_t)blockIdx.x * kWienerThreads + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * kWienerThreads) {
        const int u = (int)(i / Lc), v = (int)(i - (int64_t)u * Lc);
        const float2 s = spec[(int64_t)u * pitch + v];
        const double yr = 0.5 * (double)s.x, yi = 0.5 * (double)s.y;
        const double2 h = Hc[i];
        const double hm2 = h.x * h.x + h.y * h.y;
        const double lm2 = laplacian_mag2(tw_h, tw_w, Hn, Wn, u, v);
        const double denom = fmax(hm2 + balance * lm2, 1.0e-12);
        // conj(H) Y (complex_mul_into with conj_b, core/kernels/fft.rs:52-71)
        const double xr = (yr * h.x + yi * h.y) / denom;
        const double xi = (yi * h.x - yr * h.y) / denom;
        const double dr = yr - (h.x * xr - h.y * xi);
        const double di = yi - (h.x * xi + h.y * xr);
        const double w = (full || v == 0 || 2 * v == Wn) ? 1.0 : 2.0;
        v3[0] += w * ((dr * dr + di * di) / n);
        v3[1] += w * (lm2 * ((xr * xr + xi * xi) / n));
        v3[2] += w * (hm2 / denom);
        if (post) {
            double2 a = mean[i];
            a.x += xr;
            a.y += xi;
            mean[i] = a;
        }
    }
    block_sums<3>(v3, part);
}

// The precision update (deconvolve.rs:465-476): sigma_n^2 = rss / max(n - trace, 1) floored at
// 1e-8 max(var, 1e-6), sigma_x^2 = reg / max(trace, 1) floored at 1e-8, and the next balance
// clamp(sigma_n^2 / sigma_x^2 * jitter, 1e-6, 1e3), also appended to the chain.
__global__ __launch_bounds__(kWienerThreads) void uw_update_kernel(const double* __restrict__ part, int nb, double n,
                                                                    double jitter, double* __restrict__ state,
                                                                    double* __restrict__ chain, int iter) {
    double s[3];
    final_sums<3>(part, nb, s);
    if (threadIdx.x == 0) {
        const double rss = s[0], reg = s[1], trace = s[2];
        const double eff_dof = fmax(n - trace, 1.0);
        const double sn2 = fmax(rss / eff_dof, 1.0e-8 * fmax(state[1], 1.0e-6));
        const double sx2 = fmax(reg / fmax(trace, 1.0), 1.0e-8);
        const double b = fmin(fmax(sn2 / sx2 * jitter, 1.0e-6), 1.0e3);
        state[0] = b;
        chain[iter] = b;
    }
}

// Sizes that are not 7-smooth run on the general complex FFT over split re/im planes: the
// filter is applied in place, the spectrum moves between split planes and the interleaved
// layout of the unsupervised sweep, and the output is the real part (clipped when asked).
__global__ void apply_filter_kernel(float* __restrict__ re, float* __restrict__ im, const float2* __restrict__ G,
                                    int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const float2 g = G[i];
        const float a = re[i], b = im[i];
        re[i] = a * g.x - b * g.y;
        im[i] = a * g.y + b * g.x;
    }
}

__global__ void pack_spectrum_kernel(const float* __restrict__ re, const float* __restrict__ im, float scale,
                                     float2* __restrict__ spec, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) spec[i] = make_float2(re[i] * scale, im[i] * scale);
}

__global__ void unpack_spectrum_kernel(const float2* __restrict__ spec, float* __restrict__ re,
                                       float* __restrict__ im, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const float2 v = spec[i];
        re[i] = v.x;
        im[i] = v.y;
    }
}

__global__ void real_part_kernel(const float* __restrict__ re, float* __restrict__ out, int64_t n, int clip) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        float v = re[i];
        if (clip) v = fminf(fmaxf(v, -1.f), 1.f);
        out[i] = v;
    }
}

// Posterior mean -> the half spectrum the inverse passes take: mean * scale as float2.
__global__ void uw_finish_kernel(const double2* __restrict__ mean, double scale, int Hn, int Lc, int pitch,
                                 float2* __restrict__ spec) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)Hn * Lc) return;
    const int u = (int)(i / Lc), v = (int)(i - (int64_t)u * Lc);
    const double2 a = mean[i];
    spec[(int64_t)u * pitch + v] = make_float2((float)(a.x * scale), (float)(a.y * scale));
}

static unsigned blocks_for(int64_t n, int64_t cap = 1 << 30) {
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + kWienerThreads - 1) / kWienerThreads, cap));
}

int launch_wiener_otf(const float* psf, int kh, int kw, int Hn, int Wn, int Lc, int pitch, double2* tw_h,
                      double2* tw_w, double2* A, int filter, double balance, double scale, float2* G, double2* Hout,
                      void* stream) {
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(dft_table_kernel, dim3(blocks_for(Hn)), dim3(kWienerThreads), 0, st, tw_h, Hn);
    hipLaunchKernelGGL(dft_table_kernel, dim3(blocks_for(Wn)), dim3(kWienerThreads), 0, st, tw_w, Wn);
    hipLaunchKernelGGL(psf_row_dft_kernel, dim3(blocks_for((int64_t)kh * Lc)), dim3(kWienerThreads), 0, st, psf, kh,
                       kw, (const double2*)tw_w, Wn, Lc, A);
    const dim3 grid(blocks_for((int64_t)((Hn + kOtfRows - 1) / kOtfRows) * Lc));
    if (filter)
        hipLaunchKernelGGL(psf_otf_kernel<true>, grid, dim3(kWienerThreads), 0, st, (const double2*)A, kh,
                           (const double2*)tw_h, (const double2*)tw_w, Hn, Wn, Lc, pitch, balance, scale, G, Hout);
    else
        hipLaunchKernelGGL(psf_otf_kernel<false>, grid, dim3(kWienerThreads), 0, st, (const double2*)A, kh,
                           (const double2*)tw_h, (const double2*)tw_w, Hn, Wn, Lc, pitch, balance, scale, G, Hout);
    return (int)hipGetLastError();
}

int uw_blocks(int64_t n) { return (int)blocks_for(n, kUwMaxBlocks); }

int launch_uw_init(const float* x, int64_t n, double* part, int nb, double2* mean, int64_t nmean, double* state,
                   void* stream) {
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(moments_kernel, dim3(nb), dim3(kWienerThreads), 0, st, x, n, part);
    hipLaunchKernelGGL(uw_init_kernel, dim3(blocks_for(nmean, kUwMaxBlocks)), dim3(kWienerThreads), 0, st,
                       (const double*)part, nb, (double)n, mean, nmean, state);
    return (int)hipGetLastError();
}

int launch_uw_iteration(const float2* spec, int pitch, const double2* Hc, const double2* tw_h, const double2* tw_w,
                        int Hn, int Wn, int Lc, double n, double* state, int post, int full, double2* mean, double* part,
                        int nb, double jitter, double* chain, int iter, void* stream) {
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(uw_sweep_kernel, dim3(nb), dim3(kWienerThreads), 0, st, spec, pitch, Hc, tw_h, tw_w, Hn, Wn, Lc,
                       n, (const double*)state, post, full, mean, part);
    hipLaunchKernelGGL(uw_update_kernel, dim3(1), dim3(kWienerThreads), 0, st, (const double*)part, nb, n, jitter,
                       state, chain, iter);
    return (int)hipGetLastError();
}

int launch_uw_finish(const double2* mean, double scale, int Hn, int Lc, int pitch, float2* spec, void* stream) {
    hipLaunchKernelGGL(uw_finish_kernel, dim3(blocks_for((int64_t)Hn * Lc)), dim3(kWienerThreads), 0,
                       (hipStream_t)stream, mean, scale, Hn, Lc, pitch, spec);
    return (int)hipGetLastError();
}

int launch_apply_filter(float* re, float* im, const float2* G, int64_t n, void* stream) {
    hipLaunchKernelGGL(apply_filter_kernel, dim3(blocks_for(n)), dim3(kWienerThreads), 0, (hipStream_t)stream, re, im,
                       G, n);
    return (int)hipGetLastError();
}

int launch_pack_spectrum(const float* re, const float* im, float scale, float2* spec, int64_t n, void* stream) {
    hipLaunchKernelGGL(pack_spectrum_kernel, dim3(blocks_for(n)), dim3(kWienerThreads), 0, (hipStream_t)stream, re, im,
                       scale, spec, n);
    return (int)hipGetLastError();
}

int launch_unpack_spectrum(const float2* spec, float* re, float* im, int64_t n, void* stream) {
    hipLaunchKernelGGL(unpack_spectrum_kernel, dim3(blocks_for(n)), dim3(kWienerThreads), 0, (hipStream_t)stream, spec,
                       re, im, n);
    return (int)hipGetLastError();
}

__global__ void zero_f32_kernel(float* __restrict__ p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

int launch_zero_f32(float* p, int64_t n, void* stream) {
    hipLaunchKernelGGL(zero_f32_kernel, dim3(blocks_for(n)), dim3(kWienerThreads), 0, (hipStream_t)stream, p, n);
    return (int)hipGetLastError();
}

int launch_real_part(const float* re, float* out, int64_t n, int clip, void* stream) {
    hipLaunchKernelGGL(real_part_kernel, dim3(blocks_for(n)), dim3(kWienerThreads), 0, (hipStream_t)stream, re, out, n,
                       clip);
    return (int)hipGetLastError();
}

}  // namespace lucem
