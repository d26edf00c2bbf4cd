// General complex FFT along one axis of split re/im arrays: the fast replacement for the
// reference's O(N^2) DFT (crates/lucem-ops/src/core/fft.rs:13-100, core/kernels/fft.rs:11-47).
//
// Element (o, i, j) of an array viewed as [outer, n, inner] sits at (o * n + i) * inner + j; the
// transform runs over i. A workgroup holds B lines in LDS, interleaved as Z[i * B + b]:
//   inner > 1 (CONTIG false): B adjacent j of one o, so each global access covers B floats;
//   inner = 1 (CONTIG true):  B adjacent o, each line contiguous.
// 7-smooth lengths run the LDS Stockham FFT directly. Other lengths use Bluestein's identity
// X_k = c_k sum_m (x_m c_m) conj(c_{k-m}), c_m = exp(-i pi m^2 / n), as a circular convolution of
// power-of-two length M >= 2n - 1: forward FFT, product with the precomputed transform of the
// chirp, inverse FFT, all in LDS. The inverse transform is conj(DFT(conj(x))) / n. Every line is
// read into LDS before any is written, so the output may alias the input.
#include <hip/hip_runtime.h>

#include "fft_lds.hpp"
#include "rl_internal.hpp"

namespace lucem {

template <int T, int E, bool BLU, bool CONTIG>
__global__ __launch_bounds__(T) void axis_fft_kernel(AxisFft d) {
    extern __shared__ float2 Z[];
    const int B = d.B;
    const int n = d.n;
    const int L = d.f.L;  // n, or M for Bluestein
    const int tid = threadIdx.x;
    const int total = L * B;
    const Divisor bd = Divisor::of(B);
    const Divisor ld = Divisor::of(L);
    int64_t base, line0, col0;  // first element; first line (CONTIG) or slab; first column
    int nlines;
    if constexpr (CONTIG) {
        const int64_t o0 = (int64_t)blockIdx.x * B;
        base = o0 * n;
        line0 = o0;
        col0 = 0;
        nlines = (int)(d.outer - o0 < B ? d.outer - o0 : B);
    } else {
        const int64_t o = blockIdx.x / d.nkb;
        const int64_t j0 = ((int64_t)blockIdx.x - o * d.nkb) * B;
        base = o * n * d.inner + j0;
        line0 = o;
        col0 = j0;
        nlines = (int)(d.inner - j0 < B ? d.inner - j0 : B);
    }
    const float sgn = d.inverse ? -1.f : 1.f;  // conjugate on the way in and out for the inverse
    const float2* twl = stage_twiddles(Z, total, d.f, d.tw);
    {
        float2 v[E];
#pragma unroll
        for (int s = 0; s < E; s++) {
            const int e = tid + s * T;
            v[s] = make_float2(0.f, 0.f);
            if (e < total) {
                int i, b;
                if constexpr (CONTIG) {
                    b = ld.div(e);
                    i = e - b * L;
                } else {
                    i = bd.div(e);
                    b = e - i * B;
                }
                if (i < n && b < nlines) {
                    const int64_t a = CONTIG ? base + (int64_t)b * n + i : base + (int64_t)i * d.inner + b;
                    float2 x = make_float2(d.re_in[a], sgn * d.im_in[a]);
                    if constexpr (BLU) x = cmul(x, d.chirp[i]);
                    v[s] = x;
                }
            }
        }
#pragma unroll
        for (int s = 0; s < E; s++) {
            const int e = tid + s * T;
            if (e < total) {
                int i, b;
                if constexpr (CONTIG) {
                    b = ld.div(e);
                    i = e - b * L;
                } else {
                    i = bd.div(e);
                    b = e - i * B;
                }
                Z[i * B + b] = v[s];
            }
        }
    }
    __syncthreads();
    lds_fft<false, T, E>(Z, bd, d.f, d.tw, twl);
    if constexpr (BLU) {
        for (int e = tid; e < total; e += T) Z[e] = cmul(Z[e], d.chirp_hat[bd.div(e)]);
        __syncthreads();
        lds_fft<true, T, E>(Z, bd, d.f, d.tw, twl);
    }
    const Divisor nd = Divisor::of(n);
    const int nstore = n * B;
    const bool tr = d.tr_n1 > 0;  // four-step second pass: transposed store, k-major for coalescing
    for (int e = tid; e < nstore; e += T) {
        int k, b;
        if (CONTIG && !tr) {
            b = nd.div(e);
            k = e - b * n;
        } else {
            k = bd.div(e);
            b = e - k * B;
        }
        if (b < nlines) {
            float2 z = Z[k * B + b];
            if constexpr (BLU) z = cmul(z, d.chirp[k]);
            if (d.post_tw) {  // four-step first pass (a strided view): twiddle of column j at output k
                const int64_t j2 = (col0 + b) / d.post_inner;
                z = cmul(z, d.post_tw[(j2 * k) % d.post_n]);
            }
            int64_t a;
            if (tr) {
                const int64_t line = CONTIG ? line0 + b : line0;
                const int64_t j = CONTIG ? 0 : col0 + b;
                const int64_t ob = line / d.tr_n1, k1 = line - ob * d.tr_n1;
                a = (ob * d.tr_n + k1 + d.tr_n1 * (int64_t)k) * d.inner + j;
            } else {
                a = CONTIG ? base + e : base + (int64_t)k * d.inner + b;
            }
            d.re_out[a] = z.x * d.scale;
            d.im_out[a] = sgn * z.y * d.scale;
        }
    }
}

// Bluestein in several passes: element e of [outer, M, inner] has axis index m = (e / inner) % M.
__global__ void bluestein_pre_kernel(const float* __restrict__ re, const float* __restrict__ im,
                                     float* __restrict__ ure, float* __restrict__ uim,
                                     const float2* __restrict__ chirp, int64_t outer, int64_t n, int64_t M,
                                     int64_t inner, int inverse) {
    const int64_t total = outer * M * inner;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = e % inner, r = e / inner, m = r % M, o = r / M;
        float2 v = make_float2(0.f, 0.f);
        if (m < n) {
            const int64_t a = (o * n + m) * inner + j;
            v = cmul(make_float2(re[a], inverse ? -im[a] : im[a]), chirp[m]);
        }
        ure[e] = v.x;
        uim[e] = v.y;
    }
}

__global__ void bluestein_mid_kernel(float* __restrict__ ure, float* __restrict__ uim, const float2* __restrict__ hat,
                                     int64_t outer, int64_t M, int64_t inner) {
    const int64_t total = outer * M * inner;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t m = (e / inner) % M;
        const float2 v = cmul(make_float2(ure[e], uim[e]), hat[m]);
        ure[e] = v.x;
        uim[e] = v.y;
    }
}

__global__ void bluestein_post_kernel(const float* __restrict__ ure, const float* __restrict__ uim,
                                      float* __restrict__ re, float* __restrict__ im,
                                      const float2* __restrict__ chirp, int64_t outer, int64_t n, int64_t M,
                                      int64_t inner, int inverse, float scale) {
    const int64_t total = outer * n * inner;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = e % inner, r = e / inner, k = r % n, o = r / n;
        const int64_t a = (o * M + k) * inner + j;
        const float2 v = cmul(make_float2(ure[a], uim[a]), chirp[k]);
        re[e] = v.x * scale;
        im[e] = (inverse ? -v.y : v.y) * scale;
    }
}

// (ar + i ai) * (br + i bi), or times conj(b): complex_mul_kernel (core/kernels/fft.rs:49-71).
__global__ void complex_mul_kernel(const float* __restrict__ ar, const float* __restrict__ ai,
                                   const float* __restrict__ br, const float* __restrict__ bi, float* __restrict__ orr,
                                   float* __restrict__ oi, int64_t n, int conj_b) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const float a = ar[i], b = ai[i], c = br[i], dd = conj_b ? -bi[i] : bi[i];
        orr[i] = a * c - b * dd;
        oi[i] = a * dd + b * c;
    }
}

__global__ void copy_f32_kernel(const float* __restrict__ src, float* __restrict__ dst, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

template <int T>
static int launch_axis_t(const AxisFft& d, int blocks, size_t lds, hipStream_t st) {
    constexpr int E = kElemsPerThread;
    const bool contig = d.inner == 1;
    if (d.bluestein) {
        if (contig)
            hipLaunchKernelGGL((axis_fft_kernel<T, E, true, true>), dim3(blocks), dim3(T), lds, st, d);
        else
            hipLaunchKernelGGL((axis_fft_kernel<T, E, true, false>), dim3(blocks), dim3(T), lds, st, d);
    } else {
        if (contig)
            hipLaunchKernelGGL((axis_fft_kernel<T, E, false, true>), dim3(blocks), dim3(T), lds, st, d);
        else
            hipLaunchKernelGGL((axis_fft_kernel<T, E, false, false>), dim3(blocks), dim3(T), lds, st, d);
    }
    return (int)hipGetLastError();
}

size_t axis_fft_lds_bytes(int L, int B) { return (size_t)lds_elems(L * B, L) * sizeof(float2); }

int launch_axis_fft(const AxisFft& d, int blocks, void* stream) {
    const size_t lds = axis_fft_lds_bytes(d.f.L, d.B);
    hipStream_t st = (hipStream_t)stream;
    switch (d.wide) {
        case 2: return launch_axis_t<kThreadsXL>(d, blocks, lds, st);
        case 1: return launch_axis_t<kThreadsWide>(d, blocks, lds, st);
        default: return launch_axis_t<kThreads>(d, blocks, lds, st);
    }
}

static dim3 stride_grid(int64_t total) {
    const int64_t blocks = (total + 255) / 256;
    return dim3((unsigned)(blocks < 65536 ? blocks : 65536));
}

int launch_bluestein_pre(const float* re, const float* im, float* ure, float* uim, const float2* chirp, int64_t outer,
                         int64_t n, int64_t M, int64_t inner, int inverse, void* stream) {
    hipLaunchKernelGGL(bluestein_pre_kernel, stride_grid(outer * M * inner), dim3(256), 0, (hipStream_t)stream, re, im,
                       ure, uim, chirp, outer, n, M, inner, inverse);
    return (int)hipGetLastError();
}

int launch_bluestein_mid(float* ure, float* uim, const float2* hat, int64_t outer, int64_t M, int64_t inner,
                         void* stream) {
    hipLaunchKernelGGL(bluestein_mid_kernel, stride_grid(outer * M * inner), dim3(256), 0, (hipStream_t)stream, ure,
                       uim, hat, outer, M, inner);
    return (int)hipGetLastError();
}

int launch_bluestein_post(const float* ure, const float* uim, float* re, float* im, const float2* chirp, int64_t outer,
                          int64_t n, int64_t M, int64_t inner, int inverse, float scale, void* stream) {
    hipLaunchKernelGGL(bluestein_post_kernel, stride_grid(outer * n * inner), dim3(256), 0, (hipStream_t)stream, ure,
                       uim, re, im, chirp, outer, n, M, inner, inverse, scale);
    return (int)hipGetLastError();
}

int launch_complex_mul(const float* ar, const float* ai, const float* br, const float* bi, float* orr, float* oi,
                       int64_t n, int conj_b, void* stream) {
    const unsigned g = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(complex_mul_kernel, dim3(g), dim3(256), 0, (hipStream_t)stream, ar, ai, br, bi, orr, oi, n,
                       conj_b);
    return (int)hipGetLastError();
}

int launch_copy_f32(const float* src, float* dst, int64_t n, void* stream) {
    const unsigned g = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(copy_f32_kernel, dim3(g), dim3(256), 0, (hipStream_t)stream, src, dst, n);
    return (int)hipGetLastError();
}

template <class K>
static int allow_axis_lds(K kernel) {
    return (int)hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)(lds_elems(kLdsComplexXL, kLdsComplexXL) * sizeof(float2)));
}

int allow_big_lds_axis_kernels() {
    constexpr int E = kElemsPerThread;
    int err = 0;
    for (int e : {allow_axis_lds(axis_fft_kernel<kThreadsWide, E, false, false>),
                  allow_axis_lds(axis_fft_kernel<kThreadsWide, E, false, true>),
                  allow_axis_lds(axis_fft_kernel<kThreadsWide, E, true, false>),
                  allow_axis_lds(axis_fft_kernel<kThreadsWide, E, true, true>),
                  allow_axis_lds(axis_fft_kernel<kThreadsXL, E, false, false>),
                  allow_axis_lds(axis_fft_kernel<kThreadsXL, E, false, true>),
                  allow_axis_lds(axis_fft_kernel<kThreadsXL, E, true, false>),
                  allow_axis_lds(axis_fft_kernel<kThreadsXL, E, true, true>)})
        if (e && !err) err = e;
    return err;
}

}  // namespace lucem
