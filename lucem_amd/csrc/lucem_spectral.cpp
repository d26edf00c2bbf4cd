// C ABI of the frequency-domain filters of crates/lucem-ops/src/filter/spectral.rs, callers of
// the DFT this library replaces:
//
//   fourier_gaussian_into   spectral.rs:20-62    spectrum *= exp(-2 pi^2 sigma_d^2 f_d^2) per axis
//   fourier_box_into        spectral.rs:64-108   spectrum *= sinc(size_d f_d) per axis
//   fourier_shift_into      spectral.rs:110-165  spectrum *= exp(-2 pi i shift_d f_d) per axis
//   butterworth             spectral.rs:166-212  forward FFT, radial Butterworth weight, inverse
//                                                FFT, real part
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/lucem_hip.h"
#include "lucem_internal_host.hpp"
#include "rl_internal.hpp"

using namespace lucem;

namespace {

lucem_status check(int err, const char* what) { return err == hipSuccess ? LUCEM_OK : hip_fail(err, what); }

bool same(const uint64_t* a, uint32_t na, const uint64_t* b, uint32_t nb) {
    if (na != nb) return false;
    for (uint32_t i = 0; i < na; i++)
        if (a[i] != b[i]) return false;
    return true;
}

// spectral.rs: same_shape(re, im), broadcast_per_axis (core/check.rs:92-108), empty => OK.
lucem_status spectral_axes(lucem_client* client, float* re, const uint64_t* re_dims, uint32_t re_ndim, float* im,
                           const uint64_t* im_dims, uint32_t im_ndim, const double* values, uint32_t num_values,
                           const char* name, int kind) {
    if ((re_ndim && !re_dims) || (im_ndim && !im_dims)) return fail(LUCEM_INVALID_HANDLE, "dims pointer is NULL");
    if (!same(re_dims, re_ndim, im_dims, im_ndim))
        return fail(LUCEM_SHAPE_MISMATCH, "re shape " + dims_str(re_dims, re_ndim) + " does not match im shape " +
                                              dims_str(im_dims, im_ndim));
    if (num_values && !values) return fail(LUCEM_INVALID_HANDLE, "values pointer is NULL");
    const uint32_t nd = re_ndim;
    std::vector<double> per;
    if (nd > 0) {
        if (num_values == 1) {
            per.assign(nd, values[0]);
        } else if (num_values == nd) {
            per.assign(values, values + nd);
        } else {
            return fail(LUCEM_INVALID_PARAMETER, std::string("invalid parameter `") + name + "`: expected 1 or " +
                                                     std::to_string(nd) + " values, got " +
                                                     std::to_string(num_values));
        }
    }
    uint64_t n = 1;
    bool overflow = false;
    for (uint32_t i = 0; i < nd; i++) {
        if (re_dims[i] && n > UINT64_MAX / re_dims[i]) overflow = true;
        n *= re_dims[i];
    }
    if (n == 0) return LUCEM_OK;
    if (overflow || n > 0xFFFFFFFFull)
        return fail(LUCEM_TOO_MANY_ELEMENTS,
                    "dims " + dims_str(re_dims, nd) + " exceed the device element limit of 4294967295");
    if (!client) return fail(LUCEM_INVALID_HANDLE, "client is NULL");
    if (!re || !im) return fail(LUCEM_INVALID_HANDLE, "re or im pointer is NULL");
    hipError_t e = hipSetDevice(client->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    for (uint32_t ax = 0; ax < nd; ax++) {
        int64_t inner = 1;
        for (uint32_t a = ax + 1; a < nd; a++) inner *= (int64_t)re_dims[a];
        lucem_status st = check(launch_spectral_axis(re, im, (int64_t)n, (int64_t)re_dims[ax], inner, kind, per[ax],
                                                     client->stream),
                                "spectral axis");
        if (st) return st;
    }
    return LUCEM_OK;
}

}  // namespace

extern "C" lucem_status lucem_fourier_gaussian_into_f32(lucem_client* client, float* re, const uint64_t* re_dims,
                                                        uint32_t re_ndim, float* im, const uint64_t* im_dims,
                                                        uint32_t im_ndim, const double* sigmas, uint32_t num_sigmas) {
    return spectral_axes(client, re, re_dims, re_ndim, im, im_dims, im_ndim, sigmas, num_sigmas, "sigmas", 0);
}

extern "C" lucem_status lucem_fourier_box_into_f32(lucem_client* client, float* re, const uint64_t* re_dims,
                                                   uint32_t re_ndim, float* im, const uint64_t* im_dims,
                                                   uint32_t im_ndim, const double* sizes, uint32_t num_sizes) {
    return spectral_axes(client, re, re_dims, re_ndim, im, im_dims, im_ndim, sizes, num_sizes, "sizes", 1);
}

extern "C" lucem_status lucem_fourier_shift_into_f32(lucem_client* client, float* re, const uint64_t* re_dims,
                                                     uint32_t re_ndim, float* im, const uint64_t* im_dims,
                                                     uint32_t im_ndim, const double* shifts, uint32_t num_shifts) {
    return spectral_axes(client, re, re_dims, re_ndim, im, im_dims, im_ndim, shifts, num_shifts, "shifts", 2);
}

extern "C" lucem_status lucem_butterworth_f32(lucem_client* client, const float* input, const uint64_t* in_dims,
                                              uint32_t in_ndim, float* output, const uint64_t* out_dims,
                                              uint32_t out_ndim, double cutoff_ratio, int32_t high_pass, double order,
                                              lucem_workspace* ws) {
    // spectral.rs:172-177: cutoff_ratio in (0, 0.5]
    if (!(cutoff_ratio > 0.0 && cutoff_ratio <= 0.5)) {
        char buf[64];
        std::snprintf(buf, sizeof buf, "%g", cutoff_ratio);
        return fail(LUCEM_INVALID_PARAMETER,
                    std::string("invalid parameter `cutoff_ratio`: cutoff_ratio must be in (0.0, 0.5], got ") + buf);
    }
    if ((in_ndim && !in_dims) || (out_ndim && !out_dims)) return fail(LUCEM_INVALID_HANDLE, "dims pointer is NULL");
    if (!same(in_dims, in_ndim, out_dims, out_ndim))
        return fail(LUCEM_SHAPE_MISMATCH, "input shape " + dims_str(in_dims, in_ndim) + " does not match output shape " +
                                              dims_str(out_dims, out_ndim));
    uint64_t n = 1;
    bool overflow = false;
    for (uint32_t i = 0; i < in_ndim; i++) {
        if (in_dims[i] && n > UINT64_MAX / in_dims[i]) overflow = true;
        n *= in_dims[i];
    }
    if (n == 0) return LUCEM_OK;
    if (in_ndim > 4)
        return fail(LUCEM_UNSUPPORTED_SHAPE, "butterworth takes 1-D to 4-D arrays, got " + std::to_string(in_ndim) +
                                                 "-D shape " + dims_str(in_dims, in_ndim));
    if (overflow || n > 0xFFFFFFFFull)
        return fail(LUCEM_TOO_MANY_ELEMENTS,
                    "dims " + dims_str(in_dims, in_ndim) + " exceed the device element limit of 4294967295");
    if (!client) return fail(LUCEM_INVALID_HANDLE, "client is NULL");
    if (!input || !output) return fail(LUCEM_INVALID_HANDLE, "input or output pointer is NULL");
    if (!ws) ws = client->internal_ws.get();
    if (ws->client != client) return fail(LUCEM_INVALID_HANDLE, "workspace belongs to another client");
    hipError_t e = hipSetDevice(client->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    float *zero, *fr, *fi, *oi;
    lucem_status st;
    const size_t fb = sizeof(float) * n;
    if ((st = ws_slot(ws, "bw_zero", fb, (void**)&zero)) || (st = ws_slot(ws, "bw_re", fb, (void**)&fr)) ||
        (st = ws_slot(ws, "bw_im", fb, (void**)&fi)) || (st = ws_slot(ws, "bw_out_im", fb, (void**)&oi)))
        return st;
    if ((st = check(launch_zero_f32(zero, (int64_t)n, client->stream), "zero"))) return st;
    const uint64_t* d = in_dims;
    const uint32_t nd = in_ndim;
    // spectral.rs:182-185: the forward transform of (input, 0)
    if ((st = lucem_fft_nd_into_f32(client, input, d, nd, zero, d, nd, fr, d, nd, fi, d, nd, 0, ws))) return st;
    int64_t dims4[4] = {1, 1, 1, 1};
    for (uint32_t i = 0; i < nd; i++) dims4[4 - nd + i] = (int64_t)d[i];
    // spectral.rs:186-205: the radial weight, then the product (freq * (w + 0i))
    if ((st = check(launch_butterworth_weight(fr, fi, dims4, cutoff_ratio, order, high_pass ? 1 : 0, client->stream),
                    "butterworth weight")))
        return st;
    // spectral.rs:207-210: the inverse transform; the output is its real part
    return lucem_fft_nd_into_f32(client, fr, d, nd, fi, d, nd, output, d, nd, oi, d, nd, 1, ws);
}
