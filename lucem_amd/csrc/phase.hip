// Kernels of phase correlation (crates/lucem-ops/src/register/phase.rs:52-250) on the device:
// the cross-power spectrum (register/kernels/phase.rs:10-40, in FP32 as the reference), the first
// index of the largest |cc|^2 (core/reduce.rs:138-171, first-index tie rule), sums of a spectrum
// against plane waves at given positions (the local inverse DFT of the subpixel refinement and
// the correlation at one index), and the sum of squares of an image. The sums are FP64 and each
// workgroup writes its own partial, summed on the host in a fixed order.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "rl_internal.hpp"

namespace lucem {

constexpr int kPhaseThreads = 256;

// A * conj(B), divided by its magnitude when `normalize` (0 where the magnitude is <= eps).
__global__ void cross_power_kernel(const float* __restrict__ ar, const float* __restrict__ ai,
                                   const float* __restrict__ br, const float* __restrict__ bi, float* __restrict__ pr,
                                   float* __restrict__ pi, int64_t n, int normalize, float eps) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float a = ar[i], b = ai[i], c = br[i], d = bi[i];
        const float re = a * c + b * d;
        const float im = b * c - a * d;
        if (normalize) {
            const float mag = sqrtf(re * re + im * im);
            const float scale = mag > eps ? 1.0f / mag : 0.0f;
            pr[i] = re * scale;
            pi[i] = im * scale;
        } else {
            pr[i] = re;
            pi[i] = im;
        }
    }
}

__device__ __forceinline__ bool better(float v, int64_t i, float bv, int64_t bi) {
    return v > bv || (v == bv && i < bi);
}

// Per workgroup: the largest re^2 + im^2 and its first index.
__global__ __launch_bounds__(kPhaseThreads) void argmax_mag2_kernel(const float* __restrict__ re,
                                                                    const float* __restrict__ im, int64_t n,
                                                                    float* __restrict__ part_v,
                                                                    int64_t* __restrict__ part_i) {
    __shared__ float sv[kPhaseThreads];
    __shared__ int64_t si[kPhaseThreads];
    float bv = -1.0f;
    int64_t bi = INT64_MAX;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float v = re[i] * re[i] + im[i] * im[i];
        if (better(v, i, bv, bi)) {
            bv = v;
            bi = i;
        }
    }
    sv[threadIdx.x] = bv;
    si[threadIdx.x] = bi;
    __syncthreads();
    for (int s = kPhaseThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s && better(sv[threadIdx.x + s], si[threadIdx.x + s], sv[threadIdx.x], si[threadIdx.x])) {
            sv[threadIdx.x] = sv[threadIdx.x + s];
            si[threadIdx.x] = si[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part_v[blockIdx.x] = sv[0];
        part_i[blockIdx.x] = si[0];
    }
}

// Partial sums S_c = sum_idx z[idx] exp(+2 pi i sum_d f_d(k_d) s_cd) over this workgroup's
// elements, for the C positions s_c (3 doubles each, axes z y x of a right-aligned shape), with
// f(k) = k / n for k <= n / 2, else (k - n) / n.
__global__ __launch_bounds__(kPhaseThreads) void plane_wave_sums_kernel(const float* __restrict__ re,
                                                                        const float* __restrict__ im, int64_t n0,
                                                                        int64_t n1, int64_t n2,
                                                                        const double* __restrict__ pos, int C,
                                                                        double2* __restrict__ part) {
    __shared__ double sr[kPhaseThreads], si[kPhaseThreads];
    const int64_t n = n0 * n1 * n2;
    const double two_pi = 6.283185307179586476925286766559;
    for (int c = 0; c < C; c++) {
        const double s0 = pos[3 * c], s1 = pos[3 * c + 1], s2 = pos[3 * c + 2];
        double accr = 0.0, acci = 0.0;
        for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
             idx += (int64_t)gridDim.x * blockDim.x) {
            const int64_t k2 = idx % n2, r = idx / n2, k1 = r % n1, k0 = r / n1;
            const double f0 = (double)(k0 <= n0 / 2 ? k0 : k0 - n0) / (double)n0;
            const double f1 = (double)(k1 <= n1 / 2 ? k1 : k1 - n1) / (double)n1;
            const double f2 = (double)(k2 <= n2 / 2 ? k2 : k2 - n2) / (double)n2;
            double sn, cs;
            sincos(two_pi * (f0 * s0 + f1 * s1 + f2 * s2), &sn, &cs);
            const double zr = re[idx], zi = im[idx];
            accr += zr * cs - zi * sn;
            acci += zr * sn + zi * cs;
        }
        sr[threadIdx.x] = accr;
        si[threadIdx.x] = acci;
        __syncthreads();
        for (int s = kPhaseThreads / 2; s > 0; s >>= 1) {
            if ((int)threadIdx.x < s) {
                sr[threadIdx.x] += sr[threadIdx.x + s];
                si[threadIdx.x] += si[threadIdx.x + s];
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) part[(int64_t)c * gridDim.x + blockIdx.x] = make_double2(sr[0], si[0]);
        __syncthreads();
    }
}

// Per workgroup: sum of x^2 in FP64.
__global__ __launch_bounds__(kPhaseThreads) void sum_squares_kernel(const float* __restrict__ x, int64_t n,
                                                                    double* __restrict__ part) {
    __shared__ double s[kPhaseThreads];
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double v = x[i];
        acc += v * v;
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kPhaseThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

int phase_blocks(int64_t n) {
    const int64_t b = (n + kPhaseThreads - 1) / kPhaseThreads;
    return (int)(b < 1024 ? (b < 1 ? 1 : b) : 1024);
}

int launch_cross_power(const float* ar, const float* ai, const float* br, const float* bi, float* pr, float* pi,
                       int64_t n, int normalize, void* stream) {
    const int64_t b = (n + 255) / 256;
    hipLaunchKernelGGL(cross_power_kernel, dim3((unsigned)(b < 65536 ? b : 65536)), dim3(256), 0, (hipStream_t)stream,
                       ar, ai, br, bi, pr, pi, n, normalize, 1.0e-12f);
    return (int)hipGetLastError();
}

int launch_argmax_mag2(const float* re, const float* im, int64_t n, float* part_v, int64_t* part_i, void* stream) {
    hipLaunchKernelGGL(argmax_mag2_kernel, dim3(phase_blocks(n)), dim3(kPhaseThreads), 0, (hipStream_t)stream, re, im,
                       n, part_v, part_i);
    return (int)hipGetLastError();
}

int launch_plane_wave_sums(const float* re, const float* im, const int64_t dims3[3], const double* pos, int C,
                           double2* part, void* stream) {
    const int64_t n = dims3[0] * dims3[1] * dims3[2];
    hipLaunchKernelGGL(plane_wave_sums_kernel, dim3(phase_blocks(n)), dim3(kPhaseThreads), 0, (hipStream_t)stream, re,
                       im, dims3[0], dims3[1], dims3[2], pos, C, part);
    return (int)hipGetLastError();
}

int launch_sum_squares(const float* x, int64_t n, double* part, void* stream) {
    hipLaunchKernelGGL(sum_squares_kernel, dim3(phase_blocks(n)), dim3(kPhaseThreads), 0, (hipStream_t)stream, x, n,
                       part);
    return (int)hipGetLastError();
}

}  // namespace lucem
