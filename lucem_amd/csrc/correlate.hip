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


// correlate_1d (filter/kernels/convolve.rs:8-55): along an axis seen as [outer, len, inner],
// out[o, i, k] = sum_t w[t] * P(in)[o, i + t - left, k], accumulated in T in window order, with
// the interior fast path of the reference (no boundary arithmetic when the window fits).
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
            for (int t = 0; t < size; t++) {
                const int q = pad_index((int)(start + t), (int)len, mode);
                const T sample = q < 0 ? cval : in[base + (int64_t)q * inner];
                acc += w[t] * sample;
            }
        }
        out[e] = acc;
    }
}

template <typename T>
static int correlate1d(const T* in, T* out, const T* w, int64_t total, int64_t len, int64_t inner, int64_t left,
                       T cval, int size, int mode, void* stream) {
    hipLaunchKernelGGL(correlate1d_kernel<T>, stride_grid(total), dim3(256), 0, (hipStream_t)stream, in, out, w, total,
                       len, inner, left, cval, size, mode);
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
