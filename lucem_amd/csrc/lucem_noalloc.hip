// Allocation-free variant of the plain 2-D entry point, for callers that cannot allocate device
// memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// lucem_richardson_lucy_into_noalloc_f32 runs exactly the launches of
// lucem_richardson_lucy_into_f32 (the tiled form, the FFT form, or the direct form past the FFT
// length limit),
// but every buffer beyond the caller's input, PSF and output is carved from `scratch`, one
// fixed-size device array per GPU that the library holds until the process ends. The FFT twiddle
// tables are built into it by a kernel, the launch shapes are the heuristics or the ones this
// client has already measured for the geometry (no measuring), and profiling records nothing, so
// the call creates no event, stream or buffer: it calls no hipMalloc, hipFree or allocating
// library routine.
//
// Every call on one GPU shares that GPU's scratch and overwrites it: calls that overlap on one GPU
// (two streams, two threads) race. Every scratch value a call reads is written earlier in the same
// call, so nothing has to start at zero.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>

#include "../../include/lucem_hip.h"
#include "lucem_internal_host.hpp"
#include "rl_internal.hpp"

using namespace lucem;

namespace {

constexpr size_t kScratchBytes = (size_t)LUCEM_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Device address of `scratch` on each device (the host-side name is a stand-in, not an address)
// and the status of resolving it.
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};
hipError_t g_err[kMaxDevices] = {};
bool g_done[kMaxDevices] = {};

// exp(-2 pi i m / L) in double, rounded once to float: the table twiddles_for builds on the host.
__global__ void twiddle_table_kernel(float2* out, int L) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= L) return;
    const double two_pi = 6.283185307179586476925286766559;
    const double a = two_pi * (double)m / (double)L;
    out[m] = make_float2((float)cos(a), (float)(-sin(a)));
}

// Bump allocation from the scratch base, 256-byte aligned.
struct Carver {
    char* base;
    size_t used = 0;
    template <class T>
    T* take(size_t bytes) {
        T* p = (T*)(base + used);
        used += (bytes + 255) & ~(size_t)255;
        return p;
    }
};

}  // namespace

void lucem::noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return;
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_done[device]) return;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_err[device] = e;
    g_ptr[device] = e == hipSuccess ? p : nullptr;
    g_done[device] = true;
}

// Scratch bytes the call needs for geometry g: the tiled form's buffers and twiddle tables, the
// FFT form's, or the direct form's five images and two PSF copies.
static size_t scratch_need(const Geometry& g, bool direct, const TileGrid* tiled) {
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    if (tiled) {
        const TileSizes z = tile_sizes(g, *tiled);
        return 2 * al(z.est) + 2 * al(z.otf) + 2 * al(z.psf) + al(z.pa) + al(sizeof(float2) * (size_t)tiled->L0) +
               al(sizeof(float2) * (size_t)tiled->L1);
    }
    if (direct) {
        const size_t n = sizeof(float) * (size_t)g.n[1] * g.n[2], nk = sizeof(float) * (size_t)g.k[1] * g.k[2];
        return 5 * al(n) + 2 * al(nk);
    }
    const RlSizes z = rl_sizes(g);
    return al(z.est) + al(z.sa) + 2 * al(z.otf) + 2 * al(z.psf) + al(z.pa) + al(sizeof(float2) * (size_t)g.L[1]) +
           al(sizeof(float2) * (size_t)g.L[2]);
}

extern "C" uint64_t lucem_noalloc_scratch_bytes(void) { return (uint64_t)kScratchBytes; }

extern "C" lucem_status lucem_richardson_lucy_into_noalloc_f32(lucem_client* client, const float* input,
                                                               const uint64_t* in_dims, uint32_t in_ndim,
                                                               const float* psf, const uint64_t* psf_dims,
                                                               uint32_t psf_ndim, float* output,
                                                               const uint64_t* out_dims, uint32_t out_ndim,
                                                               const lucem_rl_settings* settings) {
    // the checks of lucem_richardson_lucy_into_f32, in the same order
    bool empty = false;
    Geometry g;
    lucem_status st = validate_rl(input, in_dims, in_ndim, psf, psf_dims, psf_ndim, output, out_dims, out_ndim,
                                  settings, 2, &empty, &g);
    bool direct = false;
    if (st == LUCEM_UNSUPPORTED_SHAPE && in_ndim == 2) {
        st = validate_rl(input, in_dims, in_ndim, psf, psf_dims, psf_ndim, output, out_dims, out_ndim, settings, 2,
                         &empty, &g, sizeof(float), false);
        direct = true;
    }
    if (st != LUCEM_OK || empty) return st;
    if (!client) return fail(LUCEM_INVALID_HANDLE, "client is NULL");
    if (!input || !psf || !output) return fail(LUCEM_INVALID_HANDLE, "input, psf or output pointer is NULL");
    TileGrid grid{};
    const bool tiled = !direct && g.ndim == 2 && use_tiled(g, client->num_cus, &grid);
    const size_t need = scratch_need(g, direct, tiled ? &grid : nullptr);
    if (need > kScratchBytes)
        return fail(LUCEM_UNSUPPORTED_SHAPE, "input " + dims_str(in_dims, in_ndim) + " with psf " +
                                                 dims_str(psf_dims, psf_ndim) + " needs " + std::to_string(need) +
                                                 " bytes of scratch; the allocation-free entry point holds " +
                                                 std::to_string(kScratchBytes) + " bytes per device");
    const int dev = client->device;
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (dev < 0 || dev >= kMaxDevices || !g_done[dev])
            return fail(LUCEM_HIP_ERROR, "the allocation-free scratch of device " + std::to_string(dev) +
                                             " was not set up (the client was not created by lucem_client_create)");
        if (g_err[dev] != hipSuccess) return hip_fail(g_err[dev], "hipGetSymbolAddress(scratch)");
        base = g_ptr[dev];
    }
    hipError_t e = hipSetDevice(dev);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    NoProfileScope quiet;
    Carver cv{(char*)base};
    if (direct) {
        const size_t n = sizeof(float) * (size_t)g.n[1] * g.n[2], nk = sizeof(float) * (size_t)g.k[1] * g.k[2];
        float* bufs[7];
        for (int i = 0; i < 5; i++) bufs[i] = cv.take<float>(n);
        for (int i = 5; i < 7; i++) bufs[i] = cv.take<float>(nk);
        return rl_direct_launch_f32(client, input, psf, output, g, *settings, bufs);
    }
    if (tiled) {
        const TileSizes z = tile_sizes(g, grid);
        TileBufs b{};
        b.est_a = cv.take<float>(z.est);
        b.est_b = cv.take<float>(z.est);
        b.otf1 = cv.take<float2>(z.otf);
        b.otf2 = cv.take<float2>(z.otf);
        b.h = cv.take<float>(z.psf);
        b.rev = cv.take<float>(z.psf);
        b.pa = cv.take<float2>(z.pa);
        float2* tw0 = cv.take<float2>(sizeof(float2) * (size_t)grid.L0);
        float2* tw1 = cv.take<float2>(sizeof(float2) * (size_t)grid.L1);
        hipLaunchKernelGGL(twiddle_table_kernel, dim3((grid.L0 + 255) / 256), dim3(256), 0, client->stream, tw0, grid.L0);
        hipLaunchKernelGGL(twiddle_table_kernel, dim3((grid.L1 + 255) / 256), dim3(256), 0, client->stream, tw1, grid.L1);
        if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "twiddle table");
        return rl_tiled_launch(client, g, *settings, input, psf, output, b, tw0, tw1, grid);
    }
    const RlSizes z = rl_sizes(g);
    RlBuffers b{};
    b.est = cv.take<float>(z.est);
    b.sa = cv.take<float2>(z.sa);
    b.otf1 = cv.take<float2>(z.otf);
    b.otf2 = cv.take<float2>(z.otf);
    b.h = cv.take<float>(z.psf);
    b.rev = cv.take<float>(z.psf);
    b.pa = cv.take<float2>(z.pa);
    float2* tw_y = cv.take<float2>(sizeof(float2) * (size_t)g.L[1]);
    float2* tw_x = cv.take<float2>(sizeof(float2) * (size_t)g.L[2]);
    hipLaunchKernelGGL(twiddle_table_kernel, dim3((g.L[1] + 255) / 256), dim3(256), 0, client->stream, tw_y, g.L[1]);
    hipLaunchKernelGGL(twiddle_table_kernel, dim3((g.L[2] + 255) / 256), dim3(256), 0, client->stream, tw_x, g.L[2]);
    if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "twiddle table");
    const float2* tws[3] = {nullptr, tw_y, tw_x};
    return rl_launch(client, g, *settings, input, psf, output, b, tws, tuned_lookup(client, g));
}
