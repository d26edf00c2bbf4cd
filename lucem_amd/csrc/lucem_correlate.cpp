// C ABI of correlate_2d_into (crates/lucem-ops/src/filter/convolve.rs:113-166), the correlation
// that the Richardson-Lucy iteration calls twice (restore/deconvolve.rs:120, :133).
//
// f32, windows of more than 169 taps: one linear correlation through the passes of the RL path. The weights' transfer function
// is built as for the PSF (without the normalisation), the input is gathered with the boundary
// mode and transformed (ROW_LOAD), the line passes multiply by the transfer function, and the
// last row pass (ROW_UPDATE with a constant estimate of 1) stores the cropped result:
// 1.0f * c == c, so the output is the correlation itself. The origin moves the gather start
// (k / 2 + origin per axis). Constant mode with cval != 0 correlates in - cval and adds
// cval * sum(w) (the zero-padded transform sees P_cval(in) - cval).
// f64, f32 windows of at most 169 taps, and f32 arrays with an axis longer than the FFT passes
// take: the direct sum in the reference's order (rl_direct.hip).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/lucem_hip.h"
#include "lucem_internal_host.hpp"
#include "rl_internal.hpp"

using namespace lucem;

extern "C" void lucem_correlate_settings_default(lucem_correlate_settings* out) {
    if (!out) return;
    std::memset(out, 0, sizeof *out);
    out->mode = LUCEM_PAD_REFLECT;
}

namespace {

// check::origin (core/check.rs:59-72)
lucem_status check_origin(int32_t origin, uint64_t size) {
    if (size == 0) return fail(LUCEM_INVALID_WEIGHTS, "window: size must be at least 1");
    const int64_t lo = -(int64_t)(size / 2), hi = (int64_t)((size - 1) / 2);
    if (origin < lo || origin > hi)
        return fail(LUCEM_INVALID_ORIGIN, "origin " + std::to_string(origin) + " is out of bounds for window of size " +
                                              std::to_string(size) + " (must be in " + std::to_string(lo) +
                                              "..=" + std::to_string(hi) + ")");
    return LUCEM_OK;
}

// The checks of correlate_2d_into in its order (convolve.rs:133-141); nd allows 3-D arrays with
// weights of the same rank. Then this library's own limits: pad mode id, element count, FFT length.
lucem_status validate_correlate(const void* input, const uint64_t* in_dims, uint32_t in_ndim, const uint64_t* w_dims,
                                uint32_t w_ndim, const void* output, const uint64_t* out_dims, uint32_t out_ndim,
                                const lucem_correlate_settings* s, bool nd, size_t elem, bool fft, bool* empty,
                                Geometry* g) {
    *empty = false;
    if (!s) return fail(LUCEM_INVALID_HANDLE, "settings is NULL");
    if ((in_ndim && !in_dims) || (out_ndim && !out_dims) || (w_ndim && !w_dims))
        return fail(LUCEM_INVALID_HANDLE, "dims pointer is NULL");
    const char* name = nd ? "correlate_nd" : "correlate_2d";
    const std::string expected = nd ? "2 or 3" : "2";
    if (in_ndim < 2 || in_ndim > (nd ? 3u : 2u))
        return fail(LUCEM_INVALID_DIMENSION, std::string(name) + " expects " + expected + "-D array, got " +
                                                 std::to_string(in_ndim) + "-D shape " + dims_str(in_dims, in_ndim));
    if (w_ndim != in_ndim)
        return fail(LUCEM_INVALID_DIMENSION, "weights expects " + std::to_string(in_ndim) + "-D array, got " +
                                                 std::to_string(w_ndim) + "-D shape " + dims_str(w_dims, w_ndim));
    bool same = in_ndim == out_ndim;
    for (uint32_t i = 0; same && i < in_ndim; i++) same = in_dims[i] == out_dims[i];
    if (!same)
        return fail(LUCEM_SHAPE_MISMATCH, "input shape " + dims_str(in_dims, in_ndim) + " does not match output shape " +
                                              dims_str(out_dims, out_ndim));
    uint64_t n = 1;
    bool overflow = false;
    for (uint32_t i = 0; i < in_ndim; i++) {
        if (in_dims[i] && n > UINT64_MAX / in_dims[i]) overflow = true;
        n *= in_dims[i];
    }
    if (n > 0 && !overflow && input && output) {
        const char* a = (const char*)input;
        const char* b = (const char*)output;
        const uint64_t bytes = n * elem;
        if (a < b + bytes && b < a + bytes)
            return fail(LUCEM_ALIASED_BUFFERS, "input and output arrays must not share the same device buffer");
    }
    for (uint32_t i = 0; i < in_ndim; i++) {
        lucem_status st = check_origin(s->origin[i], w_dims[i]);
        if (st) return st;
    }
    if (s->mode > 4)
        return fail(LUCEM_UNKNOWN_PAD_MODE, "boundary pad mode id " + std::to_string(s->mode) +
                                                " is not supported (expected 0..4: reflect, constant, nearest, "
                                                "mirror, wrap)");
    if (n == 0) {
        *empty = true;
        return LUCEM_OK;
    }
    if (overflow || n > 0xFFFFFFFFull)
        return fail(LUCEM_TOO_MANY_ELEMENTS,
                    "dims " + dims_str(in_dims, in_ndim) + " exceed the device element limit of 4294967295");
    int64_t n3[3] = {1, 1, 1}, k3[3] = {1, 1, 1};
    int o3[3] = {0, 0, 0};
    for (uint32_t i = 0; i < in_ndim; i++) {
        n3[3 - in_ndim + i] = (int64_t)in_dims[i];
        k3[3 - in_ndim + i] = (int64_t)w_dims[i];
        o3[3 - in_ndim + i] = s->origin[i];
    }
    if (!fft) {
        // the direct kernels index an axis's padded extent n + k - 1 with 32-bit integers
        for (int a = 0; a < 3; a++)
            if (n3[a] + k3[a] - 1 > 0x7FFFFFFF)
                return fail(LUCEM_UNSUPPORTED_SHAPE, "input " + dims_str(in_dims, in_ndim) + " with weights " +
                                                         dims_str(w_dims, w_ndim) +
                                                         ": an axis with its window is longer than 2147483647");
        std::memset(g, 0, sizeof *g);
        g->ndim = (int)in_ndim;
        for (int a = 0; a < 3; a++) {
            g->n[a] = (int)n3[a];
            g->k[a] = (int)k3[a];
        }
    } else if (!make_geometry(n3, k3, (int)in_ndim, g)) {
        return fail(LUCEM_UNSUPPORTED_SHAPE, "input " + dims_str(in_dims, in_ndim) + " with weights " +
                                                 dims_str(w_dims, w_ndim) +
                                                 ": an axis needs an FFT longer than 16384 (n + k - 1 per axis)");
    }
    g->mode = (int)s->mode;
    for (int a = 0; a < 3; a++) g->origin[a] = o3[a];
    return LUCEM_OK;
}

lucem_status check(int err, const char* what) { return err == hipSuccess ? LUCEM_OK : hip_fail(err, what); }

// Largest window (kh * kw taps) that a 2-D f32 correlation sums directly rather than through the
// FFT passes: measured on 2048^2, the direct sum is faster up to 13 x 13 (0.10 against 0.12 ms)
// and slower from 15 x 15 (0.13 against 0.11 ms; the FFT passes cost the same for any window).
// LUCEM_CORRELATE_DIRECT_MAX overrides it (0: always the FFT passes).
int64_t direct_window_max() {
    if (const char* v = std::getenv("LUCEM_CORRELATE_DIRECT_MAX")) return std::atoll(v);
    return 169;
}

lucem_status correlate_run(lucem_client* c, lucem_workspace* ws, const float* input, const float* w, float* output,
                           const Geometry& g, double cval) {
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    const int64_t n = (int64_t)g.n[0] * g.n[1] * g.n[2];
    const int64_t nk = (int64_t)g.k[0] * g.k[1] * g.k[2];
    const int64_t rows = (int64_t)g.n[0] * g.n[1];
    const int64_t krows = (int64_t)g.k[0] * g.k[1];
    const size_t c8 = sizeof(float2);
    const bool shift = g.mode == LUCEM_PAD_CONSTANT && cval != 0.0;
    float* est;
    float* shifted = nullptr;
    double* wsum = nullptr;
    float2 *sa, *sb = nullptr, *otf, *pa, *pb = nullptr;
    lucem_status st;
    // the large slots share the RL path's names, so one workspace serves both without growing twice
    if ((st = ws_slot(ws, "rl_est", sizeof(float) * n, (void**)&est))) return st;
    if ((st = ws_slot(ws, "rl_spec_a", c8 * rows * g.pitch, (void**)&sa))) return st;
    if (g.ndim == 3 && (st = ws_slot(ws, "rl_spec_b", c8 * (int64_t)g.n[0] * g.L[1] * g.pitch, (void**)&sb)))
        return st;
    if ((st = ws_slot(ws, "rl_otf_conv", c8 * (int64_t)g.L[0] * g.L[1] * g.pitch, (void**)&otf))) return st;
    if ((st = ws_slot(ws, "rl_psf_spec_a", c8 * krows * g.pitch, (void**)&pa))) return st;
    if (g.ndim == 3 &&
        (st = ws_slot(ws, "rl_psf_spec_b", c8 * (int64_t)g.k[0] * g.L[1] * g.pitch, (void**)&pb)))
        return st;
    if (shift && ((st = ws_slot(ws, "cor_shifted", sizeof(float) * n, (void**)&shifted)) ||
                  (st = ws_slot(ws, "cor_weight_sum", sizeof(double), (void**)&wsum))))
        return st;
    const float2* tws[3] = {nullptr, nullptr, nullptr};
    for (int a = (g.ndim == 2 ? 1 : 0); a < 3; a++)
        if ((st = twiddles_for(c, g.L[a], &tws[a]))) return st;

    if ((st = run_otf(c, g, w, pa, pb, otf, tws))) return st;
    ShapePolicy sp{c->num_cus, 0, 0, nullptr};
    if ((st = tuned_shapes(c, g, est, sa, sb, otf, tws, &sp.table))) return st;
    if (shift) {
        if ((st = check(launch_cval_shift(input, shifted, (float)cval, n, c->stream), "cval shift"))) return st;
        if ((st = check(launch_weight_sum(w, nk, wsum, c->stream), "weight sum"))) return st;
    }
    int row_blocks = 0;
    RowPass r = make_row_desc(g, sp, tws[2], 1.0f, nullptr, est, output, sa, &row_blocks);
    r.op = ROW_LOAD;
    r.src = shift ? shifted : input;
    if ((st = run_row(c, r, row_blocks, "cor_row_load"))) return st;
    if ((st = run_corr(c, g, sp, sa, sb, otf, tws))) return st;
    r.op = ROW_UPDATE;
    r.first = 1;
    r.last = 1;
    r.clip = 0;
    r.init = 1.0f;
    if ((st = run_row(c, r, row_blocks, "cor_row_store"))) return st;
    if (shift && (st = check(launch_add_weight_sum(output, wsum, cval, n, c->stream), "cval add"))) return st;
    return LUCEM_OK;
}

lucem_status correlate_entry(lucem_client* client, const float* input, const uint64_t* in_dims, uint32_t in_ndim,
                             const float* w, const uint64_t* w_dims, uint32_t w_ndim, float* output,
                             const uint64_t* out_dims, uint32_t out_ndim, const lucem_correlate_settings* s,
                             lucem_workspace* ws, bool nd) {
    bool empty = false;
    Geometry g;
    lucem_status st = validate_correlate(input, in_dims, in_ndim, w_dims, w_ndim, output, out_dims, out_ndim, s, nd,
                                         sizeof(float), true, &empty, &g);
    bool direct = false;
    if (st == LUCEM_UNSUPPORTED_SHAPE && in_ndim == 2) {
        // an axis longer than the FFT passes take: the direct sum
        st = validate_correlate(input, in_dims, in_ndim, w_dims, w_ndim, output, out_dims, out_ndim, s, nd,
                                sizeof(float), false, &empty, &g);
        direct = true;
    } else if (st == LUCEM_OK && !empty && in_ndim == 2 && (int64_t)g.k[1] * g.k[2] <= direct_window_max()) {
        direct = true;  // a small window: the direct sum is cheaper than three FFT passes
    }
    if (st != LUCEM_OK || empty) return st;
    if (!client) return fail(LUCEM_INVALID_HANDLE, "client is NULL");
    if (!input || !w || !output) return fail(LUCEM_INVALID_HANDLE, "input, weights or output pointer is NULL");
    if (direct) {
        hipError_t e = hipSetDevice(client->device);
        if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
        const int kh = g.k[1], kw = g.k[2];
        return check(launch_correlate2d_direct(input, w, output, g.n[1], g.n[2], kh, kw, kh / 2 + g.origin[1],
                                               kw / 2 + g.origin[2], g.mode, (float)s->cval, client->stream),
                     "correlate (direct)");
    }
    if (!ws) ws = client->internal_ws.get();
    if (ws->client != client) return fail(LUCEM_INVALID_HANDLE, "workspace belongs to another client");
    return correlate_run(client, ws, input, w, output, g, s->cval);
}

}  // namespace

extern "C" lucem_status lucem_correlate_2d_into_f32(lucem_client* client, const float* input, const uint64_t* in_dims,
                                                    uint32_t in_ndim, const float* weights, const uint64_t* w_dims,
                                                    uint32_t w_ndim, float* output, const uint64_t* out_dims,
                                                    uint32_t out_ndim, const lucem_correlate_settings* settings,
                                                    lucem_workspace* ws) {
    return correlate_entry(client, input, in_dims, in_ndim, weights, w_dims, w_ndim, output, out_dims, out_ndim,
                           settings, ws, false);
}

extern "C" lucem_status lucem_correlate_nd_into_f32(lucem_client* client, const float* input, const uint64_t* in_dims,
                                                    uint32_t in_ndim, const float* weights, const uint64_t* w_dims,
                                                    uint32_t w_ndim, float* output, const uint64_t* out_dims,
                                                    uint32_t out_ndim, const lucem_correlate_settings* settings,
                                                    lucem_workspace* ws) {
    return correlate_entry(client, input, in_dims, in_ndim, weights, w_dims, w_ndim, output, out_dims, out_ndim,
                           settings, ws, true);
}

extern "C" lucem_status lucem_correlate_2d_into_f64(lucem_client* client, const double* input, const uint64_t* in_dims,
                                                    uint32_t in_ndim, const double* weights, const uint64_t* w_dims,
                                                    uint32_t w_ndim, double* output, const uint64_t* out_dims,
                                                    uint32_t out_ndim, const lucem_correlate_settings* settings) {
    bool empty = false;
    Geometry g;
    lucem_status st = validate_correlate(input, in_dims, in_ndim, w_dims, w_ndim, output, out_dims, out_ndim, settings,
                                         false, sizeof(double), false, &empty, &g);
    if (st != LUCEM_OK || empty) return st;
    if (!client) return fail(LUCEM_INVALID_HANDLE, "client is NULL");
    if (!input || !weights || !output) return fail(LUCEM_INVALID_HANDLE, "input, weights or output pointer is NULL");
    hipError_t e = hipSetDevice(client->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    const int kh = g.k[1], kw = g.k[2];
    return check(launch_correlate2d_direct(input, weights, output, g.n[1], g.n[2], kh, kw, kh / 2 + g.origin[1],
                                           kw / 2 + g.origin[2], g.mode, settings->cval, client->stream),
                 "correlate (f64)");
}

// gabor_kernel_2d (filter/spectral.rs:216-246): the complex Gabor kernel in FP64 on the host,
// rounded to f32 as the reference's upload casts it.
static void gabor_kernel(const lucem_gabor_settings& s, int* kh, int* kw, std::vector<float>* re,
                         std::vector<float>* im) {
    const double pi = 3.14159265358979323846264338327950288;
    const double ct = std::cos(s.theta), st = std::sin(s.theta);
    // fmax drops a NaN as the reference's f64::max does
    const long rx = (long)std::fmax(std::ceil(s.n_stds * std::hypot(s.sigma_x * ct, s.sigma_y * st)), 1.0);
    const long ry = (long)std::fmax(std::ceil(s.n_stds * std::hypot(s.sigma_x * st, s.sigma_y * ct)), 1.0);
    *kh = (int)(2 * ry + 1);
    *kw = (int)(2 * rx + 1);
    const double norm = 1.0 / (2.0 * pi * s.sigma_x * s.sigma_y);
    re->clear();
    im->clear();
    for (long y = -ry; y <= ry; y++)
        for (long x = -rx; x <= rx; x++) {
            const double xr = (double)x * ct + (double)y * st;
            const double yr = -(double)x * st + (double)y * ct;
            const double env =
                norm * std::exp(-0.5 * (xr * xr / (s.sigma_x * s.sigma_x) + yr * yr / (s.sigma_y * s.sigma_y)));
            const double phase = 2.0 * pi * s.frequency * xr + s.offset;
            re->push_back((float)(env * std::cos(phase)));
            im->push_back((float)(env * std::sin(phase)));
        }
}

extern "C" void lucem_gabor_settings_default(lucem_gabor_settings* out) {
    if (!out) return;
    std::memset(out, 0, sizeof *out);
    out->frequency = 0.1;
    out->sigma_x = 1.0;
    out->sigma_y = 1.0;
    out->n_stds = 3.0;
    out->mode = LUCEM_PAD_REFLECT;
}

extern "C" lucem_status lucem_gabor_2d_f32(lucem_client* client, const float* input, const uint64_t* dims,
                                           uint32_t ndim, float* out_re, float* out_im,
                                           const lucem_gabor_settings* settings, lucem_workspace* ws) {
    if (!settings) return fail(LUCEM_INVALID_HANDLE, "settings is NULL");
    if (ndim && !dims) return fail(LUCEM_INVALID_HANDLE, "dims pointer is NULL");
    {
        // a kernel radius beyond 2^20 is refused (the reference would try to allocate it)
        const double r = std::fabs(settings->n_stds) * (std::fabs(settings->sigma_x) + std::fabs(settings->sigma_y));
        if (r > 1048576.0)
            return fail(LUCEM_INVALID_PARAMETER, "invalid parameter `n_stds`: the Gabor kernel would be too large");
    }
    int kh = 0, kw = 0;
    std::vector<float> kre, kim;
    gabor_kernel(*settings, &kh, &kw, &kre, &kim);
    const uint64_t kd[2] = {(uint64_t)kh, (uint64_t)kw};
    lucem_correlate_settings cs;
    lucem_correlate_settings_default(&cs);
    cs.mode = settings->mode;
    cs.cval = settings->cval;
    // spectral.rs:262-266: two correlate_2d calls with the kernel's real and imaginary parts (the
    // checks are theirs); the kernels are uploaded once the arguments pass
    bool empty = false;
    Geometry g;
    lucem_status st = validate_correlate(input, dims, ndim, kd, 2, out_re, dims, ndim, &cs, false, sizeof(float), false,
                                         &empty, &g);
    if (st == LUCEM_OK && !empty)
        st = validate_correlate(input, dims, ndim, kd, 2, out_im, dims, ndim, &cs, false, sizeof(float), false, &empty,
                                &g);
    if (st != LUCEM_OK || empty) return st;
    if (!client) return fail(LUCEM_INVALID_HANDLE, "client is NULL");
    if (!ws) ws = client->internal_ws.get();
    if (ws->client != client) return fail(LUCEM_INVALID_HANDLE, "workspace belongs to another client");
    hipError_t e = hipSetDevice(client->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    float *dre, *dim;
    const size_t kb = sizeof(float) * kre.size();
    if ((st = ws_slot(ws, "gabor_re", kb, (void**)&dre)) || (st = ws_slot(ws, "gabor_im", kb, (void**)&dim))) return st;
    if ((e = hipMemcpyAsync(dre, kre.data(), kb, hipMemcpyHostToDevice, client->stream)) != hipSuccess ||
        (e = hipMemcpyAsync(dim, kim.data(), kb, hipMemcpyHostToDevice, client->stream)) != hipSuccess ||
        (e = hipStreamSynchronize(client->stream)) != hipSuccess)
        return hip_fail(e, "gabor kernel upload");
    if ((st = correlate_entry(client, input, dims, ndim, dre, kd, 2, out_re, dims, ndim, &cs, ws, false))) return st;
    return correlate_entry(client, input, dims, ndim, dim, kd, 2, out_im, dims, ndim, &cs, ws, false);
}

extern "C" void lucem_correlate1d_settings_default(lucem_correlate1d_settings* out) {
    if (!out) return;
    std::memset(out, 0, sizeof *out);
    out->axis = -1;
    out->mode = LUCEM_PAD_REFLECT;
}

// correlate_1d_into (filter/convolve.rs:40-75): same shape, not aliased, origin, axis, empty.
template <typename T>
static lucem_status correlate1d_entry(lucem_client* client, const T* input, const uint64_t* dims, uint32_t ndim,
                                      const T* weights, uint64_t num_weights, T* output, const uint64_t* out_dims,
                                      uint32_t out_ndim, const lucem_correlate1d_settings* s) {
    if (!s) return fail(LUCEM_INVALID_HANDLE, "settings is NULL");
    if ((ndim && !dims) || (out_ndim && !out_dims)) return fail(LUCEM_INVALID_HANDLE, "dims pointer is NULL");
    bool same = ndim == out_ndim;
    for (uint32_t i = 0; same && i < ndim; i++) same = dims[i] == out_dims[i];
    if (!same)
        return fail(LUCEM_SHAPE_MISMATCH, "input shape " + dims_str(dims, ndim) + " does not match output shape " +
                                              dims_str(out_dims, out_ndim));
    uint64_t n = 1;
    bool overflow = false;
    for (uint32_t i = 0; i < ndim; i++) {
        if (dims[i] && n > UINT64_MAX / dims[i]) overflow = true;
        n *= dims[i];
    }
    if (n > 0 && !overflow && input && output) {
        const char* a = (const char*)input;
        const char* b = (const char*)output;
        const uint64_t bytes = n * sizeof(T);
        if (a < b + bytes && b < a + bytes)
            return fail(LUCEM_ALIASED_BUFFERS, "input and output arrays must not share the same device buffer");
    }
    lucem_status st = check_origin(s->origin, num_weights);
    if (st) return st;
    const int64_t nd = (int64_t)ndim;
    const int64_t ax = s->axis < 0 ? s->axis + nd : s->axis;
    if (ax < 0 || ax >= nd)
        return fail(LUCEM_AXIS_OUT_OF_BOUNDS,
                    "axis " + std::to_string(s->axis) + " is out of bounds for " + std::to_string(nd) + "-D array");
    if (s->mode > 4)
        return fail(LUCEM_UNKNOWN_PAD_MODE, "boundary pad mode id " + std::to_string(s->mode) +
                                                " is not supported (expected 0..4: reflect, constant, nearest, "
                                                "mirror, wrap)");
    if (n == 0) return LUCEM_OK;
    if (overflow || n > 0xFFFFFFFFull)
        return fail(LUCEM_TOO_MANY_ELEMENTS,
                    "dims " + dims_str(dims, ndim) + " exceed the device element limit of 4294967295");
    if (dims[ax] > 0x7FFFFFFFull || num_weights > 0x7FFFFFFFull)
        return fail(LUCEM_UNSUPPORTED_SHAPE, "axis or window longer than 2^31 - 1");
    if (!client) return fail(LUCEM_INVALID_HANDLE, "client is NULL");
    if (!input || !weights || !output) return fail(LUCEM_INVALID_HANDLE, "input, weights or output pointer is NULL");
    hipError_t e = hipSetDevice(client->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    int64_t inner = 1;
    for (int64_t a = ax + 1; a < nd; a++) inner *= (int64_t)dims[a];
    const int64_t left = (int64_t)(num_weights / 2) + s->origin;
    return check(launch_correlate1d(input, output, weights, (int64_t)n, (int64_t)dims[ax], inner, left, (T)s->cval,
                                    (int)num_weights, (int)s->mode, client->stream),
                 "correlate 1-D");
}

extern "C" lucem_status lucem_correlate_1d_into_f32(lucem_client* client, const float* input, const uint64_t* dims,
                                                    uint32_t ndim, const float* weights, uint64_t num_weights,
                                                    float* output, const uint64_t* out_dims, uint32_t out_ndim,
                                                    const lucem_correlate1d_settings* settings) {
    return correlate1d_entry<float>(client, input, dims, ndim, weights, num_weights, output, out_dims, out_ndim,
                                    settings);
}

extern "C" lucem_status lucem_correlate_1d_into_f64(lucem_client* client, const double* input, const uint64_t* dims,
                                                    uint32_t ndim, const double* weights, uint64_t num_weights,
                                                    double* output, const uint64_t* out_dims, uint32_t out_ndim,
                                                    const lucem_correlate1d_settings* settings) {
    return correlate1d_entry<double>(client, input, dims, ndim, weights, num_weights, output, out_dims, out_ndim,
                                     settings);
}
