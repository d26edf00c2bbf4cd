// Device helpers of the correlation entry points (filter/convolve.rs:40-166): the direct 1-D
// correlation along an axis, and the Constant-mode cval of the FFT form. The FFT passes pad with
// zeros, so that form correlates in - cval and adds cval * sum(w) back:
// P_cval(in) = P_0(in - cval) + cval everywhere on the padded domain.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "fft_lds.hpp"
#include "rl_internal.hpp"

namespace lucem {

__global__ void shift_kernel(const float* __restrict__ in, float* __restrict__ out, float c, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = in[i] - c;
}

// sum of the weights in f64, one workgroup of 256 threads
__global__ __launch_bounds__(256) void weight_sum_kernel(const float* __restrict__ w, int64_t nk,
                                                         double* __restrict__ sum) {
    __shared__ double part[256];
    double acc = 0.0;
    for (int64_t i = threadIdx.x; i < nk; i += 256) acc += (double)w[i];
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) *sum = part[0];
}

__global__ void add_weight_sum_kernel(float* __restrict__ out, const double* __restrict__ sum, double cval,
                                      int64_t n) {
    const float add = (float)(cval * *sum);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] += add;
}

static dim3 stride_grid(int64_t n) {
    const int64_t blocks = (n + 255) / 256;
    return dim3((unsigned)(blocks < 65536 ? blocks : 65536));
}

int launch_cval_shift(const float* in, float* out, float cval, int64_t n, void* stream) {
    hipLaunchKernelGGL(shift_kernel, stride_grid(n), dim3(256), 0, (hipStream_t)stream, in, out, cval, n);
    return (int)hipGetLastError();
}

int launch_weight_sum(const float* w, int64_t nk, double* sum, void* stream) {
    hipLaunchKernelGGL(weight_sum_kernel, dim3(1), dim3(256), 0, (hipStream_t)stream, w, nk, sum);
    return (int)hipGetLastError();
}

int launch_add_weight_sum(float* out, const double* sum, double cval, int64_t n, void* stream) {
    hipLaunchKernelGGL(add_weight_sum_kernel, stride_grid(n), dim3(256), 0, (hipStream_t)stream, out, sum, cval, n);
    return (int)hipGetLastError();
}


// pad_index (core/kernels/pad.rs:11-34) in 64-bit arithmetic, for axes up to the element limit;
// -1 for Constant outside [0, len).
__device__ __forceinline__ int64_t pad_index64(int64_t p, int64_t len, int mode) {
    if (p >= 0 && p < len) return p;
    switch (mode) {
        case 1: return -1;
        case 2: return p < 0 ? 0 : len - 1;
        case 4: return ((p % len) + len) % len;
        case 0: {
            const int64_t period = 2 * len;
            const int64_t r = ((p % period) + period) % period;
            return r < len ? r : period - 1 - r;
        }
        default: {
            const int64_t period = len > 1 ? 2 * len - 2 : 1;
            const int64_t r = ((p % period) + period) % period;
            return r < len ? r : period - r;
        }
    }
}

template <typename T>
__device__ __forceinline__ T padded_sample(const T* __restrict__ line, int64_t pos, int64_t len, int64_t stride,
                                           T cval, int mode) {
    const int64_t q = pad_index64(pos, len, mode);
    return q < 0 ? cval : line[q * stride];
}

// correlate_1d (filter/kernels/convolve.rs:8-55): along an axis seen as [outer, len, inner],
// out[o, i, k] = sum_t w[t] * P(in)[o, i + t - left, k], accumulated in T in window order. Every
// form below reads the same padded samples and sums them in the same order, so they agree bit
// for bit.
//
// The general form: one output per thread, each tap one global load (used for short lines and
// narrow inner extents, and for windows too long for the LDS forms).
template <typename T>
__global__ void correlate1d_kernel(const T* __restrict__ in, T* __restrict__ out, const T* __restrict__ w,
                                   int64_t total, int64_t len, int64_t inner, int64_t left, T cval, int size,
                                   int mode) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = (e / inner) % len;
        const int64_t base = e - i * inner;
        const int64_t start = i - left;
        T acc = T(0);
        if (start >= 0 && start + size <= len) {
            const T* p = in + base + start * inner;
            for (int t = 0; t < size; t++) acc += w[t] * p[(int64_t)t * inner];
        } else {
            for (int t = 0; t < size; t++) acc += w[t] * padded_sample(in + base, start + t, len, inner, cval, mode);
        }
        out[e] = acc;
    }
}

constexpr int kRowMaxOut = 1024;  // outputs per block along a contiguous line (up to 4 per thread)
constexpr int kColLanes = 64; // inner positions per block (one per lane of a wave)
constexpr size_t kCorr1dLds = 65536;

// The last axis (inner = 1): a block stages the weights and the 256 U + size - 1 padded samples of
// one line segment in LDS with coalesced loads, then each thread sums U windows from LDS (at its
// index and every 256 further on). U = 4, 2 or 1 as the line holds 1024, 512 or 256 outputs.
template <typename T, int U>
__global__ __launch_bounds__(256) void correlate1d_row_kernel(const T* __restrict__ in, T* __restrict__ out,
                                                              const T* __restrict__ w, int64_t len, uint32_t nseg,
                                                              int64_t left, T cval, int size, int mode) {
    extern __shared__ __align__(16) unsigned char corr1d_raw[];
    T* tile = reinterpret_cast<T*>(corr1d_raw);
    constexpr int kOut = 256 * U;
    T* wl = tile + (kOut + size - 1);
    const uint32_t seg = blockIdx.x % nseg, line = blockIdx.x / nseg;
    const T* src = in + (int64_t)line * len;
    const int64_t i0 = (int64_t)seg * kOut;
    for (int t = threadIdx.x; t < size; t += 256) wl[t] = w[t];
    const int span = kOut + size - 1;
    for (int j = threadIdx.x; j < span; j += 256) {
        const int64_t pos = i0 - left + j;
        tile[j] = (pos >= 0 && pos < len) ? src[pos] : padded_sample(src, pos, len, (int64_t)1, cval, mode);
    }
    __syncthreads();
    T acc[U];
#pragma unroll
    for (int u = 0; u < U; u++) acc[u] = T(0);
    const T* p = tile + threadIdx.x;
    for (int t = 0; t < size; t++) {
        const T wt = wl[t];
#pragma unroll
        for (int u = 0; u < U; u++) acc[u] += wt * p[t + 256 * u];
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
        const int64_t i = i0 + threadIdx.x + 256 * u;
        if (i < len) out[(int64_t)line * len + i] = acc[u];
    }
}

// A strided axis (inner >= 64): a block of four waves covers 64 consecutive inner positions (one
// per lane, so every global access is a coalesced 64-wide row) and ROWS outputs along the axis
// (64, or 32 where 64 would leave fewer than 8 blocks per CU). It stages the ROWS + size - 1
// padded rows in LDS, rows interleaved over the waves, then each wave sums ROWS / 4 outputs from
// LDS (a lane reads its own column: no bank conflicts).
template <typename T, int kColRows>
__global__ __launch_bounds__(256) void correlate1d_col_kernel(const T* __restrict__ in, T* __restrict__ out,
                                                              const T* __restrict__ w, int64_t len, int64_t inner,
                                                              uint32_t nkb, uint32_t nib, int64_t left, T cval,
                                                              int size, int mode) {
    extern __shared__ __align__(16) unsigned char corr1d_raw[];
    T* tile = reinterpret_cast<T*>(corr1d_raw);
    T* wl = tile + (int64_t)(kColRows + size - 1) * kColLanes;
    uint32_t b = blockIdx.x;
    const uint32_t kb = b % nkb;
    b /= nkb;
    const uint32_t ib = b % nib, o = b / nib;
    const int lane = threadIdx.x & (kColLanes - 1), wave = threadIdx.x / kColLanes;
    const int64_t k = (int64_t)kb * kColLanes + lane;
    const T* src = in + (int64_t)o * len * inner + k;
    T* dst = out + (int64_t)o * len * inner + k;
    const int64_t i0 = (int64_t)ib * kColRows;
    for (int t = threadIdx.x; t < size; t += 256) wl[t] = w[t];
    const int rows = kColRows + size - 1;
    if (k < inner) {
        for (int r = wave; r < rows; r += 4) {
            const int64_t pos = i0 - left + r;
            tile[r * kColLanes + lane] =
                (pos >= 0 && pos < len) ? src[pos * inner] : padded_sample(src, pos, len, inner, cval, mode);
        }
    }
    __syncthreads();
    if (k < inner) {
        for (int r = wave; r < kColRows; r += 4) {
            const int64_t i = i0 + r;
            if (i >= len) break;
            const T* p = tile + r * kColLanes + lane;
            T acc = T(0);
            for (int t = 0; t < size; t++) acc += wl[t] * p[t * kColLanes];
            dst[i * inner] = acc;
        }
    }
}

template <typename T>
static int correlate1d(const T* in, T* out, const T* w, int64_t total, int64_t len, int64_t inner, int64_t left,
                       T cval, int size, int mode, void* stream) {
    hipStream_t st = (hipStream_t)stream;
    const int64_t outer = total / (len * inner);
    const int64_t max_blocks = ((int64_t)1 << 31) / 256;
    if (inner == 1 && len >= 256) {
        const int u = len >= 1024 ? 4 : (len >= 512 ? 2 : 1);
        const size_t lds = (size_t)(256 * u + 2 * size - 1) * sizeof(T);
        const int64_t nseg = (len + 256 * u - 1) / (256 * u);
        if (lds <= kCorr1dLds && nseg * outer <= max_blocks) {
            auto kernel = u == 4 ? correlate1d_row_kernel<T, 4>
                                 : (u == 2 ? correlate1d_row_kernel<T, 2> : correlate1d_row_kernel<T, 1>);
            hipLaunchKernelGGL(kernel, dim3((unsigned)(nseg * outer)), dim3(256), lds, st, in, out, w, len,
                               (uint32_t)nseg, left, cval, size, mode);
            return (int)hipGetLastError();
        }
    } else if (inner >= kColLanes) {
        const int64_t nkb = (inner + kColLanes - 1) / kColLanes;
        const int rows = nkb * ((len + 63) / 64) * outer >= 2048 ? 64 : 32;
        const size_t lds = ((size_t)(rows + size - 1) * kColLanes + size) * sizeof(T);
        const int64_t nib = (len + rows - 1) / rows;
        if (lds <= kCorr1dLds && nkb * nib * outer <= max_blocks) {
            auto kernel = rows == 64 ? correlate1d_col_kernel<T, 64> : correlate1d_col_kernel<T, 32>;
            hipLaunchKernelGGL(kernel, dim3((unsigned)(nkb * nib * outer)), dim3(256), lds, st, in, out, w, len,
                               inner, (uint32_t)nkb, (uint32_t)nib, left, cval, size, mode);
            return (int)hipGetLastError();
        }
    }
    hipLaunchKernelGGL(correlate1d_kernel<T>, stride_grid(total), dim3(256), 0, st, in, out, w, total, len, inner, left,
                       cval, size, mode);
    return (int)hipGetLastError();
}

int launch_correlate1d(const float* in, float* out, const float* w, int64_t total, int64_t len, int64_t inner,
                       int64_t left, float cval, int size, int mode, void* stream) {
    return correlate1d<float>(in, out, w, total, len, inner, left, cval, size, mode, stream);
}

int launch_correlate1d(const double* in, double* out, const double* w, int64_t total, int64_t len, int64_t inner,
                       int64_t left, double cval, int size, int mode, void* stream) {
    return correlate1d<double>(in, out, w, total, len, inner, left, cval, size, mode, stream);
}

}  // namespace lucem
