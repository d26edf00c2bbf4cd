// Frequency-domain filters (crates/lucem-ops/src/filter/spectral.rs:20-212) on split re/im
// spectra: per-axis transfer functions of the signed frequency f = k / n (k <= n / 2, else
// (k - n) / n), evaluated in FP64 and rounded to FP32 as the reference's uploaded tables are, and
// applied in FP32 one axis at a time in the reference's order (filter/kernels/spectral.rs:10-51).
// Butterworth's weight is the N-D radial one of spectral.rs:186-200.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "rl_internal.hpp"

namespace lucem {

__device__ __forceinline__ double signed_freq(int64_t k, int64_t n) {
    return (double)(k <= n / 2 ? k : k - n) / (double)n;
}

// kind 0: exp(-2 pi^2 p^2 f^2) (Gaussian, p = sigma); 1: sinc(p f) (box, p = size);
// 2: exp(-2 pi i p f) (shift, p = shift), on the axis viewed as [outer, len, inner].
__global__ void spectral_axis_kernel(float* __restrict__ re, float* __restrict__ im, int64_t total, int64_t len,
                                     int64_t inner, int kind, double p) {
    const double pi = 3.14159265358979323846264338327950288;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const double f = signed_freq((i / inner) % len, len);
        if (kind == 2) {
            double s, c;
            sincos(-2.0 * pi * p * f, &s, &c);
            const float cf = (float)c, sf = (float)s;
            const float r = re[i], m = im[i];
            re[i] = r * cf - m * sf;
            im[i] = r * sf + m * cf;
        } else {
            double w;
            if (kind == 0) {
                w = exp(-2.0 * pi * pi * p * p * f * f);
            } else {
                const double x = pi * p * f;
                w = fabs(x) < 1e-12 ? 1.0 : sin(x) / x;
            }
            const float wf = (float)w;
            re[i] *= wf;
            im[i] *= wf;
        }
    }
}

// Butterworth: w = lp or 1 - lp, lp = 1 / (1 + dist^(2 order)), dist = |f / cutoff| over the
// axes (right-aligned dims n0 n1 n2 n3), applied to the spectrum in FP32.
__global__ void butterworth_weight_kernel(float* __restrict__ re, float* __restrict__ im, int64_t n0, int64_t n1,
                                          int64_t n2, int64_t n3, double cutoff, double order, int high_pass) {
    const int64_t total = n0 * n1 * n2 * n3;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        int64_t r = i;
        const int64_t k3 = r % n3;
        r /= n3;
        const int64_t k2 = r % n2;
        r /= n2;
        const int64_t k1 = r % n1;
        const int64_t k0 = r / n1;
        double r2 = 0.0;
        const int64_t ks[4] = {k0, k1, k2, k3}, ns[4] = {n0, n1, n2, n3};
        for (int d = 0; d < 4; d++) {
            if (ns[d] == 1) continue;  // a padding axis of the right-aligned shape
            const double q = signed_freq(ks[d], ns[d]) / cutoff;
            r2 += q * q;
        }
        const double lp = 1.0 / (1.0 + pow(sqrt(r2), 2.0 * order));
        const float w = (float)(high_pass ? 1.0 - lp : lp);
        re[i] *= w;
        im[i] *= w;
    }
}

static dim3 grid_for(int64_t total) {
    const int64_t b = (total + 255) / 256;
    return dim3((unsigned)(b < 65536 ? (b < 1 ? 1 : b) : 65536));
}

int launch_spectral_axis(float* re, float* im, int64_t total, int64_t len, int64_t inner, int kind, double p,
                         void* stream) {
    hipLaunchKernelGGL(spectral_axis_kernel, grid_for(total), dim3(256), 0, (hipStream_t)stream, re, im, total, len,
                       inner, kind, p);
    return (int)hipGetLastError();
}

int launch_butterworth_weight(float* re, float* im, const int64_t dims4[4], double cutoff, double order, int high_pass,
                              void* stream) {
    const int64_t total = dims4[0] * dims4[1] * dims4[2] * dims4[3];
    hipLaunchKernelGGL(butterworth_weight_kernel, grid_for(total), dim3(256), 0, (hipStream_t)stream, re, im,
                       dims4[0], dims4[1], dims4[2], dims4[3], cutoff, order, high_pass);
    return (int)hipGetLastError();
}

}  // namespace lucem
