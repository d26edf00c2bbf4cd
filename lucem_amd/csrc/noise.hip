// Synthetic noise on the device: add_noise (crates/lucem-ops/src/intensity/noise.rs:35-143) with
// the counter-based generator of intensity/kernels/noise.rs:1-135 and core/kernels/random.rs:7-12.
//
// Every pixel draws from its own counter: uniform(seed, c) = pcg_hash(c ^ seed) / 2^32 (the PCG
// output hash of Jarzynski and Olano, "Hash Functions for GPU Rendering", 2020), normal draws by
// Box-Muller from the counters 2i and 2i + 1, and the Poisson draws for lambda < 20 by Knuth's
// product of uniforms from the counters 131 i + k. The integer arithmetic wraps at 32 bits, as in
// the reference, so the uniforms and every decision taken on them match it bit for bit.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/lucem_hip.h"
#include "rl_internal.hpp"

namespace lucem {

__device__ __forceinline__ uint32_t pcg_hash(uint32_t input) {
    const uint32_t state = input * 747796405u + 2891336453u;
    const uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}

__device__ __forceinline__ float uniform_f32(uint32_t seed, uint32_t c) {
    return (float)pcg_hash(c ^ seed) * (1.0f / 4294967296.0f);
}

__device__ __forceinline__ float normal_f32(uint32_t seed, uint32_t idx) {
    const float u1 = fmaxf(uniform_f32(seed, idx * 2u), 1e-7f);
    const float u2 = uniform_f32(seed, idx * 2u + 1u);
    return sqrtf(-2.0f * logf(u1)) * cosf(6.2831855f * u2);
}

__device__ __forceinline__ float clip01(float v) { return fminf(fmaxf(v, 0.0f), 1.0f); }

template <int KIND>
__global__ void add_noise_kernel(const float* __restrict__ in, float* __restrict__ out, uint32_t seed, float a,
                                 float b, int clip, uint32_t total) {
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const float x = in[idx];
    float val;
    if constexpr (KIND == LUCEM_NOISE_GAUSSIAN) {  // x + mean + std * z
        val = x + a + b * normal_f32(seed, idx);
        if (clip) val = clip01(val);
    } else if constexpr (KIND == LUCEM_NOISE_SPECKLE) {  // x + x * (mean + std * z)
        val = x + x * (a + b * normal_f32(seed, idx));
        if (clip) val = clip01(val);
    } else if constexpr (KIND == LUCEM_NOISE_POISSON) {  // a = photons per unit intensity
        const float lambda = fmaxf(x, 0.0f) * a;
        if (lambda <= 0.0f) {
            val = 0.0f;
        } else if (lambda < 20.0f) {
            const float l = expf(-lambda);
            uint32_t k = 0;
            float p = 1.0f;
            while (k < 128u) {
                k += 1u;
                p *= fmaxf(uniform_f32(seed, idx * 131u + k), 1e-7f);
                if (p <= l) break;
            }
            val = (float)(k - 1u) / a;
        } else {
            const float z = normal_f32(seed, idx);
            val = fmaxf(floorf(lambda + sqrtf(lambda) * z + 0.5f), 0.0f) / a;
        }
        if (clip) val = clip01(val);
    } else {  // impulse: a = amount, b = salt_vs_pepper; salt 1, pepper 0
        const float u = uniform_f32(seed, idx * 2u);
        const float v = uniform_f32(seed, idx * 2u + 1u);
        val = x;
        if (u < a) val = v < b ? 1.0f : 0.0f;
    }
    out[idx] = val;
}

int launch_add_noise(const float* in, float* out, int kind, uint32_t seed, float a, float b, int clip, uint32_t total,
                     void* stream) {
    const dim3 grid((unsigned)(((uint64_t)total + 255) / 256));
    hipStream_t st = (hipStream_t)stream;
    switch (kind) {
        case LUCEM_NOISE_GAUSSIAN:
            hipLaunchKernelGGL(add_noise_kernel<LUCEM_NOISE_GAUSSIAN>, grid, dim3(256), 0, st, in, out, seed, a, b,
                               clip, total);
            break;
        case LUCEM_NOISE_SPECKLE:
            hipLaunchKernelGGL(add_noise_kernel<LUCEM_NOISE_SPECKLE>, grid, dim3(256), 0, st, in, out, seed, a, b,
                               clip, total);
            break;
        case LUCEM_NOISE_POISSON:
            hipLaunchKernelGGL(add_noise_kernel<LUCEM_NOISE_POISSON>, grid, dim3(256), 0, st, in, out, seed, a, b,
                               clip, total);
            break;
        default:
            hipLaunchKernelGGL(add_noise_kernel<LUCEM_NOISE_SALT>, grid, dim3(256), 0, st, in, out, seed, a, b, clip,
                               total);
            break;
    }
    return (int)hipGetLastError();
}

}  // namespace lucem
