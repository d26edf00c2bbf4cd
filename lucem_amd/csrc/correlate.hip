// Device helpers of the correlation entry points (filter/convolve.rs:126-166) for Constant mode
// with a nonzero cval. The FFT passes pad with zeros, so the correlation runs on in - cval and
// cval * sum(w) is added back: P_cval(in) = P_0(in - cval) + cval everywhere on the padded domain.
#include <hip/hip_runtime.h>

#include <cstdint>

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

}  // namespace lucem
