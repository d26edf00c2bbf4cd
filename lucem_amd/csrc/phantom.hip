// Synthetic phantoms rendered on the device: background plus filled ellipses (2-D) or spheres
// (3-D), the ground truth of the synthetic Richardson-Lucy workloads. Each pixel sums, in the
// order of the object list and in float64, the value of every object it lies inside, and rounds
// the sum to float32 once. The inside tests use the generator's float64 operations in its order
// with contraction off, so the result equals the host generator's float64 image rounded to
// float32, bit for bit.
//
// A workgroup renders a 16x16 (2-D) or 4x8x8 (3-D) block of pixels. It walks the object list in
// chunks of 256: each thread tests one object's box against the block, the objects that overlap
// are compacted in list order into LDS (wave ballots and counts), and every pixel then tests
// only those.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/lucem_hip.h"
#include "rl_internal.hpp"

namespace lucem {

constexpr int kObjChunk = 256;

__device__ __forceinline__ bool inside_ellipse(const lucem_ellipse& e, int y, int x) {
#pragma clang fp contract(off)
    const double dy = (double)y - e.cy, dx = (double)x - e.cx;
    const double u = (dx * e.cos_theta + dy * e.sin_theta) / e.a;
    const double v = (-dx * e.sin_theta + dy * e.cos_theta) / e.b;
    return u * u + v * v <= 1.0;
}

__device__ __forceinline__ bool inside_sphere(const lucem_sphere& s, int z, int y, int x) {
#pragma clang fp contract(off)
    const double dz = (double)z - s.cz, dy = (double)y - s.cy, dx = (double)x - s.cx;
    return dz * dz + dy * dy + dx * dx <= s.radius_sq;
}

// Position of this thread's object among the chunk's hits (list order), and the chunk's count.
__device__ __forceinline__ int compact_slot(bool hit, int* wcount, int* total) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const unsigned long long m = __ballot(hit);
    if (lane == 0) wcount[wave] = __popcll(m);
    __syncthreads();
    int off = 0, all = 0;
#pragma unroll
    for (int w = 0; w < kObjChunk / 64; w++) {
        if (w < wave) off += wcount[w];
        all += wcount[w];
    }
    *total = all;
    return off + __popcll(m & ((1ull << lane) - 1ull));
}

__global__ __launch_bounds__(256) void render_ellipses_kernel(float* __restrict__ out, int H, int W,
                                                              const lucem_ellipse* __restrict__ obj, int64_t count,
                                                              double background) {
    __shared__ lucem_ellipse sel[kObjChunk];
    __shared__ int wcount[kObjChunk / 64];
    const int tid = threadIdx.x;
    const int x0 = blockIdx.x * 16, y0 = blockIdx.y * 16;
    const int x = x0 + (tid & 15), y = y0 + (tid >> 4);
    double acc = background;
    for (int64_t base = 0; base < count; base += kObjChunk) {
        const int64_t i = base + tid;
        bool hit = false;
        lucem_ellipse e{};
        if (i < count) {
            e = obj[i];
            // every pixel inside lies within max(a, b) of the centre; one pixel of margin
            const double r = fmax(e.a, e.b) + 1.0;
            hit = e.cy + r >= y0 && e.cy - r <= y0 + 15 && e.cx + r >= x0 && e.cx - r <= x0 + 15;
        }
        int total;
        const int slot = compact_slot(hit, wcount, &total);
        if (hit) sel[slot] = e;
        __syncthreads();
        if (x < W && y < H)
            for (int j = 0; j < total; j++)
                if (inside_ellipse(sel[j], y, x)) acc += sel[j].value;
        __syncthreads();
    }
    if (x < W && y < H) out[(int64_t)y * W + x] = (float)acc;
}

__global__ __launch_bounds__(256) void render_spheres_kernel(float* __restrict__ out, int D, int H, int W,
                                                             const lucem_sphere* __restrict__ obj, int64_t count,
                                                             double background) {
    __shared__ lucem_sphere sel[kObjChunk];
    __shared__ int wcount[kObjChunk / 64];
    const int tid = threadIdx.x;
    const int x0 = blockIdx.x * 8, y0 = blockIdx.y * 8, z0 = blockIdx.z * 4;
    const int x = x0 + (tid & 7), y = y0 + ((tid >> 3) & 7), z = z0 + (tid >> 6);
    double acc = background;
    for (int64_t base = 0; base < count; base += kObjChunk) {
        const int64_t i = base + tid;
        bool hit = false;
        lucem_sphere s{};
        if (i < count) {
            s = obj[i];
            const double r = s.radius + 1.0;
            hit = s.cz + r >= z0 && s.cz - r <= z0 + 3 && s.cy + r >= y0 && s.cy - r <= y0 + 7 && s.cx + r >= x0 &&
                  s.cx - r <= x0 + 7;
        }
        int total;
        const int slot = compact_slot(hit, wcount, &total);
        if (hit) sel[slot] = s;
        __syncthreads();
        if (x < W && y < H && z < D)
            for (int j = 0; j < total; j++)
                if (inside_sphere(sel[j], z, y, x)) acc += sel[j].value;
        __syncthreads();
    }
    if (x < W && y < H && z < D) out[((int64_t)z * H + y) * W + x] = (float)acc;
}

int launch_render_ellipses(float* out, int H, int W, const lucem_ellipse* obj, int64_t count, double background,
                           void* stream) {
    const dim3 grid((unsigned)((W + 15) / 16), (unsigned)((H + 15) / 16));
    hipLaunchKernelGGL(render_ellipses_kernel, grid, dim3(256), 0, (hipStream_t)stream, out, H, W, obj, count,
                       background);
    return (int)hipGetLastError();
}

int launch_render_spheres(float* out, int D, int H, int W, const lucem_sphere* obj, int64_t count, double background,
                          void* stream) {
    const dim3 grid((unsigned)((W + 7) / 8), (unsigned)((H + 7) / 8), (unsigned)((D + 3) / 4));
    hipLaunchKernelGGL(render_spheres_kernel, grid, dim3(256), 0, (hipStream_t)stream, out, D, H, W, obj, count,
                       background);
    return (int)hipGetLastError();
}

}  // namespace lucem
