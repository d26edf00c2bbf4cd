// C ABI of the Wiener forms on the FFT engine of the RL path.
//
//   wiener_deconvolve_into   crates/lucem-ops/src/restore/deconvolve.rs:213-315 (checks :221-249)
//   wiener_deconvolve        deconvolve.rs:318-327
//   unsupervised_wiener      deconvolve.rs:357-509 (checks :363-383)
//
// Both work on the circular H x W domain of the image (no padding), as the reference's DFTs do.
// The PSF's transfer function is a direct FP64 DFT of the kernel (wiener.hip). When both image
// sizes are 7-smooth (at most 16384) the image's half spectrum comes from the row R2C and the
// line FFT along y, and the result from the inverse line FFT and the row C2R, whose store
// applies the optional clip. Other sizes take the general complex FFT of lucem_fft.cpp over the
// full spectrum (Bluestein lines where a length is not 7-smooth, passes for long lines).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/lucem_hip.h"
#include "lucem_internal_host.hpp"
#include "rl_internal.hpp"

using namespace lucem;

namespace {

struct WienerGeom {
    int H, W, kh, kw, Lc, pitch;
    bool general;  // full spectrum on the general complex FFT (Lc = pitch = W)
};

bool fft_length_ok(int v) {
    if (v < 1 || v > LUCEM_MAX_FFT_LEN) return false;
    for (int p : {2, 3, 5, 7})
        while (v % p == 0) v /= p;
    return v == 1;
}

bool dims_ok(const uint64_t* d, uint32_t nd) { return nd == 0 || d != nullptr; }

uint64_t count(const uint64_t* d, uint32_t nd, bool* overflow) {
    uint64_t n = 1;
    for (uint32_t i = 0; i < nd; i++) {
        if (d[i] && n > UINT64_MAX / d[i]) *overflow = true;
        n *= d[i];
    }
    return n;
}

bool overlaps(const void* a, const void* b, uint64_t bytes) {
    const char* x = (const char*)a;
    const char* y = (const char*)b;
    return x < y + bytes && y < x + bytes;
}

// The checks that follow the reference's own (sizes this library needs), shared by both forms.
lucem_status finish_geometry(const uint64_t* in_dims, const uint64_t* psf_dims, uint64_t n, bool overflow,
                             WienerGeom* g) {
    if (psf_dims[0] > in_dims[0] || psf_dims[1] > in_dims[1])
        return fail(LUCEM_INVALID_WEIGHTS, "psf: PSF dimensions must not exceed image dimensions");
    if (overflow || n > 0xFFFFFFFFull)
        return fail(LUCEM_TOO_MANY_ELEMENTS,
                    "dims " + dims_str(in_dims, 2) + " exceed the device element limit of 4294967295");
    const int H = (int)std::min<uint64_t>(in_dims[0], 1u << 30), W = (int)std::min<uint64_t>(in_dims[1], 1u << 30);
    // the general FFT: 7-smooth lengths that split into two of at most 16384, others up to 2^27
    auto general_ok = [](int v) { return fft_length_ok(v) || v <= (1 << 27); };
    if (!general_ok(H) || !general_ok(W))
        return fail(LUCEM_UNSUPPORTED_SHAPE, "image " + dims_str(in_dims, 2) + ": an axis is longer than the FFT takes");
    g->H = H;
    g->W = W;
    g->kh = (int)psf_dims[0];
    g->kw = (int)psf_dims[1];
    g->general = !fft_length_ok(H) || !fft_length_ok(W);
    g->Lc = g->general ? W : W / 2 + 1;
    g->pitch = g->general ? W : (g->Lc + 15) & ~15;
    return LUCEM_OK;
}

lucem_status validate_wiener(const void* input, const uint64_t* in_dims, uint32_t in_ndim, const uint64_t* psf_dims,
                             uint32_t psf_ndim, const void* output, const uint64_t* out_dims, uint32_t out_ndim,
                             const lucem_wiener_settings* s, bool* empty, WienerGeom* g) {
    *empty = false;
    if (!s) return fail(LUCEM_INVALID_HANDLE, "settings is NULL");
    if (!dims_ok(in_dims, in_ndim) || !dims_ok(out_dims, out_ndim) || !dims_ok(psf_dims, psf_ndim))
        return fail(LUCEM_INVALID_HANDLE, "dims pointer is NULL");
    bool same = in_ndim == out_ndim;
    for (uint32_t i = 0; same && i < in_ndim; i++) same = in_dims[i] == out_dims[i];
    if (!same)
        return fail(LUCEM_SHAPE_MISMATCH, "input shape " + dims_str(in_dims, in_ndim) + " does not match output shape " +
                                              dims_str(out_dims, out_ndim));
    bool overflow = false;
    const uint64_t n = count(in_dims, in_ndim, &overflow);
    if (n > 0 && !overflow && input && output && overlaps(input, output, n * sizeof(float)))
        return fail(LUCEM_ALIASED_BUFFERS, "input and output arrays must not share the same device buffer");
    if (in_ndim != 2)
        return fail(LUCEM_INVALID_DIMENSION, "wiener_deconvolve input expects 2-D array, got " +
                                                 std::to_string(in_ndim) + "-D shape " + dims_str(in_dims, in_ndim));
    if (psf_ndim != 2)
        return fail(LUCEM_INVALID_DIMENSION, "wiener_deconvolve psf expects 2-D array, got " +
                                                 std::to_string(psf_ndim) + "-D shape " + dims_str(psf_dims, psf_ndim));
    if (psf_dims[0] * psf_dims[1] == 0) return fail(LUCEM_INVALID_WEIGHTS, "psf: PSF cannot be empty");
    if (!std::isfinite(s->balance) || s->balance < 0.0) {
        char buf[96];
        std::snprintf(buf, sizeof buf, "%g", s->balance);
        return fail(LUCEM_INVALID_PARAMETER,
                    std::string("invalid parameter `balance`: balance must be non-negative and finite, got ") + buf);
    }
    if (n == 0) {
        *empty = true;
        return LUCEM_OK;
    }
    return finish_geometry(in_dims, psf_dims, n, overflow, g);
}

lucem_status validate_unsupervised(const void* input, const uint64_t* dims, uint32_t ndim, const uint64_t* psf_dims,
                                   uint32_t psf_ndim, const void* output, const lucem_unsupervised_wiener_settings* s,
                                   uint64_t chain_capacity, bool* empty, WienerGeom* g) {
    *empty = false;
    if (!s) return fail(LUCEM_INVALID_HANDLE, "settings is NULL");
    if (!dims_ok(dims, ndim) || !dims_ok(psf_dims, psf_ndim)) return fail(LUCEM_INVALID_HANDLE, "dims pointer is NULL");
    if (ndim != 2)
        return fail(LUCEM_INVALID_DIMENSION, "unsupervised_wiener input expects 2-D array, got " +
                                                 std::to_string(ndim) + "-D shape " + dims_str(dims, ndim));
    if (psf_ndim != 2)
        return fail(LUCEM_INVALID_DIMENSION, "unsupervised_wiener psf expects 2-D array, got " +
                                                 std::to_string(psf_ndim) + "-D shape " + dims_str(psf_dims, psf_ndim));
    if (psf_dims[0] * psf_dims[1] == 0) return fail(LUCEM_INVALID_WEIGHTS, "psf: PSF cannot be empty");
    if (s->num_iter == 0 || s->burnin >= s->num_iter)
        return fail(LUCEM_INVALID_PARAMETER, "invalid parameter `num_iter`: num_iter (" + std::to_string(s->num_iter) +
                                                 ") must be > burnin (" + std::to_string(s->burnin) + ") and >= 1");
    bool overflow = false;
    const uint64_t n = count(dims, ndim, &overflow);
    if (n == 0) {
        *empty = true;
        return LUCEM_OK;
    }
    if (!overflow && input && output && overlaps(input, output, n * sizeof(float)))
        return fail(LUCEM_ALIASED_BUFFERS, "input and output arrays must not share the same device buffer");
    if (chain_capacity < s->num_iter)
        return fail(LUCEM_INVALID_PARAMETER, "invalid parameter `balance_chain`: room for " +
                                                 std::to_string(chain_capacity) + " values, num_iter is " +
                                                 std::to_string(s->num_iter));
    if (s->num_iter > 0x7FFFFFFFull)
        return fail(LUCEM_INVALID_PARAMETER, "invalid parameter `num_iter`: at most 2147483647 iterations");
    return finish_geometry(dims, psf_dims, n, overflow, g);
}

lucem_status check_launch(int err, const char* what) {
    if (err != hipSuccess) return hip_fail(err, what);
    return LUCEM_OK;
}

// Row pass over the image rows (length W, no padding) and the line pass along y (length H) of
// the half spectra in `spec` ([H][pitch]).
RowPass image_rows(const lucem_client* c, const WienerGeom& g, const float2* tw_w, float2* spec, int* blocks) {
    const ShapePolicy sp{c->num_cus, 0, 0, nullptr};
    RowPass r{};
    r.nx = g.W;
    r.rows = g.H;
    r.left = 0;
    r.mode = LUCEM_PAD_WRAP;
    r.Lc = g.Lc;
    r.pitch = g.pitch;
    int var = 0;
    r.P = choose_P(g.W, g.H, sp, &r.wide, &var);
    r.f = make_plan(g.W, var);
    r.init = 1.0f;  // ROW_UPDATE with `first`: the result is 1 * (the inverse transform)
    r.spec = spec;
    r.tw = tw_w;
    *blocks = (((g.H + 1) / 2) + r.P - 1) / r.P;
    return r;
}

ColPass image_lines(const lucem_client* c, const WienerGeom& g, int op, int kind, const float2* tw_h, float2* spec,
                    const float2* otf, int* blocks) {
    const ShapePolicy sp{c->num_cus, 0, 0, nullptr};
    ColPass y{};
    y.op = op;
    y.n_src = g.H;
    y.left = 0;
    y.mode = LUCEM_PAD_WRAP;
    y.n_dst = g.H;
    y.Lc = g.Lc;
    y.tw = tw_h;
    y.src = spec;
    y.dst = spec;
    y.src_outer = y.dst_outer = 0;
    y.src_axis = y.dst_axis = g.pitch;
    y.otf = otf;
    y.otf_outer = 0;
    y.otf_axis = g.pitch;
    int var = 0;
    y.B = choose_B(g.H, g.Lc, 1, kind, sp, &y.wide, &var);
    y.f = make_plan(g.H, var);
    y.nkb = (g.Lc + y.B - 1) / y.B;
    *blocks = y.nkb;
    return y;
}

lucem_status prepare(lucem_client* c, lucem_workspace* ws, const WienerGeom& g, float2** spec, double2** tw,
                     double2** psf_rows, const float2* tws[2]) {
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    lucem_status st;
    if ((st = ws_slot(ws, "wiener_spec", sizeof(float2) * (size_t)g.H * g.pitch, (void**)spec))) return st;
    if ((st = ws_slot(ws, "wiener_dft_tables", sizeof(double2) * ((size_t)g.H + g.W), (void**)tw))) return st;
    if ((st = ws_slot(ws, "wiener_psf_rows", sizeof(double2) * (size_t)g.kh * g.Lc, (void**)psf_rows))) return st;
    if (g.general) return LUCEM_OK;  // the general FFT plans its own lines
    if ((st = twiddles_for(c, g.H, &tws[0]))) return st;
    return twiddles_for(c, g.W, &tws[1]);
}

// General route: (re, im) = the image's full spectrum, through lucem_fft_nd_into_f32.
lucem_status image_spectrum_general(lucem_client* c, lucem_workspace* ws, const float* input, const WienerGeom& g,
                                    float** re, float** im) {
    const int64_t n = (int64_t)g.H * g.W;
    lucem_status st;
    if ((st = ws_slot(ws, "wiener_re", sizeof(float) * n, (void**)re))) return st;
    if ((st = ws_slot(ws, "wiener_im", sizeof(float) * n, (void**)im))) return st;
    if ((st = check_launch(launch_zero_f32(*im, n, c->stream), "wiener zero"))) return st;
    const uint64_t d[2] = {(uint64_t)g.H, (uint64_t)g.W};
    return lucem_fft_nd_into_f32(c, input, d, 2, *im, d, 2, *re, d, 2, *im, d, 2, 0, ws);
}

// General route: output = real part of the inverse transform of (re, im), clipped when asked.
lucem_status image_from_spectrum_general(lucem_client* c, lucem_workspace* ws, float* re, float* im, float* output,
                                         const WienerGeom& g, bool clip) {
    const uint64_t d[2] = {(uint64_t)g.H, (uint64_t)g.W};
    lucem_status st = lucem_fft_nd_into_f32(c, re, d, 2, im, d, 2, re, d, 2, im, d, 2, 1, ws);
    if (st) return st;
    return check_launch(launch_real_part(re, output, (int64_t)g.H * g.W, clip ? 1 : 0, c->stream), "wiener output");
}

lucem_status wiener_run(lucem_client* c, lucem_workspace* ws, const float* input, const float* psf, float* output,
                        const WienerGeom& g, const lucem_wiener_settings& s) {
    float2 *spec, *filt;
    double2 *tw, *psf_rows;
    const float2* tws[2];
    lucem_status st = prepare(c, ws, g, &spec, &tw, &psf_rows, tws);
    if (st) return st;
    if ((st = ws_slot(ws, "wiener_filter", sizeof(float2) * (size_t)g.H * g.pitch, (void**)&filt))) return st;
    // The row R2C carries a factor 2 and the unnormalised inverse transforms a factor H * W; the
    // general FFT's inverse is normalised.
    const double scale = g.general ? 1.0 : 0.5 / ((double)g.H * (double)g.W);
    if ((st = check_launch(launch_wiener_otf(psf, g.kh, g.kw, g.H, g.W, g.Lc, g.pitch, tw, tw + g.H, psf_rows, 1,
                                             s.balance, scale, filt, nullptr, c->stream),
                           "wiener filter")))
        return st;
    if (g.general) {
        float *re, *im;
        if ((st = image_spectrum_general(c, ws, input, g, &re, &im))) return st;
        if ((st = check_launch(launch_apply_filter(re, im, filt, (int64_t)g.H * g.W, c->stream), "wiener filter")))
            return st;
        return image_from_spectrum_general(c, ws, re, im, output, g, s.clip != 0);
    }
    int rb = 0, cb = 0;
    RowPass r = image_rows(c, g, tws[1], spec, &rb);
    r.op = ROW_LOAD;
    r.src = input;
    if ((st = run_row(c, r, rb, "wiener_row_fwd"))) return st;
    const ColPass y = image_lines(c, g, COL_FUSED, SHAPE_COL_FUSED, tws[0], spec, filt, &cb);
    if ((st = run_col(c, y, cb, "wiener_line_fused_y"))) return st;
    r.op = ROW_UPDATE;
    r.first = 1;
    r.last = 1;
    r.clip = s.clip ? 1 : 0;
    r.out = output;
    return run_row(c, r, rb, "wiener_row_inv");
}

// The jitter of every iteration (deconvolve.rs:429-431, :472-475): a 64-bit LCG seeded with
// seed + 0x9E3779B97F4A7C15, u = (state >> 33) / 2^31 - 0.5, jitter = exp(u * 2 / sqrt(n)).
std::vector<double> jitters(uint64_t seed, uint64_t iters, double n) {
    std::vector<double> out(iters);
    uint64_t state = seed + 0x9E3779B97F4A7C15ull;
    for (uint64_t i = 0; i < iters; i++) {
        state = state * 6364136223846793005ull + 1ull;
        const double u = (double)(state >> 33) / (double)(1ull << 31) - 0.5;
        out[i] = std::exp(u * (2.0 / std::sqrt(n)));
    }
    return out;
}

lucem_status unsupervised_run(lucem_client* c, lucem_workspace* ws, const float* input, const float* psf,
                              float* output, const WienerGeom& g, const lucem_unsupervised_wiener_settings& s,
                              double* chain_host) {
    float2* spec;
    double2 *tw, *psf_rows, *Hc, *mean;
    double *part, *state, *chain;
    const float2* tws[2];
    lucem_status st = prepare(c, ws, g, &spec, &tw, &psf_rows, tws);
    if (st) return st;
    const int64_t nh = (int64_t)g.H * g.Lc;
    if ((st = ws_slot(ws, "uw_otf", sizeof(double2) * nh, (void**)&Hc))) return st;
    if ((st = ws_slot(ws, "uw_mean", sizeof(double2) * nh, (void**)&mean))) return st;
    if ((st = ws_slot(ws, "uw_partials", sizeof(double) * 3 * kUwMaxBlocks, (void**)&part))) return st;
    if ((st = ws_slot(ws, "uw_state", sizeof(double) * 2, (void**)&state))) return st;
    if ((st = ws_slot(ws, "uw_chain", sizeof(double) * s.num_iter, (void**)&chain))) return st;
    const double n = (double)g.H * (double)g.W;
    const int64_t npix = (int64_t)g.H * g.W;

    // transfer function (FP64), image moments, balance 0.05, posterior sum cleared
    if ((st = check_launch(launch_wiener_otf(psf, g.kh, g.kw, g.H, g.W, g.Lc, g.pitch, tw, tw + g.H, psf_rows, 0, 0.0,
                                             0.0, nullptr, Hc, c->stream),
                           "unsupervised_wiener transfer function")))
        return st;
    if ((st = check_launch(launch_uw_init(input, npix, part, uw_blocks(npix), mean, nh, state, c->stream),
                           "unsupervised_wiener init")))
        return st;
    // the image's spectrum as 2 Y (half spectrum from the R2C passes, or the full one packed)
    int rb = 0, cb = 0;
    RowPass r{};
    float *re = nullptr, *im = nullptr;
    if (g.general) {
        if ((st = image_spectrum_general(c, ws, input, g, &re, &im))) return st;
        if ((st = check_launch(launch_pack_spectrum(re, im, 2.0f, spec, npix, c->stream), "unsupervised_wiener pack")))
            return st;
    } else {
        r = image_rows(c, g, tws[1], spec, &rb);
        r.op = ROW_LOAD;
        r.src = input;
        if ((st = run_row(c, r, rb, "uw_row_fwd"))) return st;
        const ColPass yf = image_lines(c, g, COL_FWD, SHAPE_COL_FWD, tws[0], spec, nullptr, &cb);
        if ((st = run_col(c, yf, cb, "uw_line_fwd_y"))) return st;
    }

    const std::vector<double> jit = jitters(s.seed, s.num_iter, n);
    const int nb = uw_blocks(nh);
    for (uint64_t it = 0; it < s.num_iter; it++) {
        if ((st = check_launch(launch_uw_iteration(spec, g.pitch, Hc, tw, tw + g.H, g.H, g.W, g.Lc, n, state,
                                                   it >= s.burnin ? 1 : 0, g.general ? 1 : 0, mean, part, nb,
                                                   jit[it], chain, (int)it, c->stream),
                               "unsupervised_wiener iteration")))
            return st;
    }
    // posterior mean / count (and 1 / n for the unnormalised inverse passes), then back
    const double inv = 1.0 / ((double)(s.num_iter - s.burnin) * (g.general ? 1.0 : n));
    if ((st = check_launch(launch_uw_finish(mean, inv, g.H, g.Lc, g.pitch, spec, c->stream), "unsupervised_wiener mean")))
        return st;
    if (g.general) {
        if ((st = check_launch(launch_unpack_spectrum(spec, re, im, npix, c->stream), "unsupervised_wiener unpack")))
            return st;
        if ((st = image_from_spectrum_general(c, ws, re, im, output, g, s.clip != 0))) return st;
    } else {
        const ColPass yi = image_lines(c, g, COL_INV, SHAPE_COL_INV, tws[0], spec, nullptr, &cb);
        if ((st = run_col(c, yi, cb, "uw_line_inv_y"))) return st;
        r.op = ROW_UPDATE;
        r.first = 1;
        r.last = 1;
        r.clip = s.clip ? 1 : 0;
        r.out = output;
        if ((st = run_row(c, r, rb, "uw_row_inv"))) return st;
    }
    // the chain is returned on the host, as the reference returns its Vec<f64>
    hipError_t e = hipMemcpyAsync(chain_host, chain, sizeof(double) * s.num_iter, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_fail(e, "unsupervised_wiener balance chain");
    return LUCEM_OK;
}

lucem_status wiener_entry(lucem_client* client, const float* input, const uint64_t* in_dims, uint32_t in_ndim,
                          const float* psf, const uint64_t* psf_dims, uint32_t psf_ndim, float* output,
                          const uint64_t* out_dims, uint32_t out_ndim, const lucem_wiener_settings* settings,
                          lucem_workspace* ws) {
    bool empty = false;
    WienerGeom g{};
    lucem_status st = validate_wiener(input, in_dims, in_ndim, psf_dims, psf_ndim, output, out_dims, out_ndim,
                                      settings, &empty, &g);
    if (st != LUCEM_OK || empty) return st;
    if (!client) return fail(LUCEM_INVALID_HANDLE, "client is NULL");
    if (!input || !psf || !output) return fail(LUCEM_INVALID_HANDLE, "input, psf or output pointer is NULL");
    if (!ws) ws = client->internal_ws.get();
    if (ws->client != client) return fail(LUCEM_INVALID_HANDLE, "workspace belongs to another client");
    return wiener_run(client, ws, input, psf, output, g, *settings);
}

}  // namespace

extern "C" void lucem_wiener_settings_default(lucem_wiener_settings* out) {
    if (!out) return;
    std::memset(out, 0, sizeof *out);
    out->balance = 0.1;
    out->clip = 1;
}

extern "C" void lucem_unsupervised_wiener_settings_default(lucem_unsupervised_wiener_settings* out) {
    if (!out) return;
    std::memset(out, 0, sizeof *out);
    out->num_iter = 30;
    out->burnin = 10;
    out->clip = 1;
    out->seed = 0;
}

extern "C" lucem_status lucem_wiener_deconvolve_into_f32(lucem_client* client, const float* input,
                                                         const uint64_t* in_dims, uint32_t in_ndim, const float* psf,
                                                         const uint64_t* psf_dims, uint32_t psf_ndim, float* output,
                                                         const uint64_t* out_dims, uint32_t out_ndim,
                                                         const lucem_wiener_settings* settings, lucem_workspace* ws) {
    return wiener_entry(client, input, in_dims, in_ndim, psf, psf_dims, psf_ndim, output, out_dims, out_ndim,
                        settings, ws);
}

extern "C" lucem_status lucem_wiener_deconvolve_f32(lucem_client* client, const float* input, const uint64_t* dims,
                                                    uint32_t ndim, const float* psf, const uint64_t* psf_dims,
                                                    uint32_t psf_ndim, float* output,
                                                    const lucem_wiener_settings* settings) {
    return wiener_entry(client, input, dims, ndim, psf, psf_dims, psf_ndim, output, dims, ndim, settings, nullptr);
}

extern "C" lucem_status lucem_unsupervised_wiener_f32(lucem_client* client, const float* input, const uint64_t* dims,
                                                      uint32_t ndim, const float* psf, const uint64_t* psf_dims,
                                                      uint32_t psf_ndim, float* output,
                                                      const lucem_unsupervised_wiener_settings* settings,
                                                      double* balance_chain, uint64_t chain_capacity,
                                                      uint64_t* chain_len, lucem_workspace* ws) {
    if (chain_len) *chain_len = 0;
    bool empty = false;
    WienerGeom g{};
    lucem_status st = validate_unsupervised(input, dims, ndim, psf_dims, psf_ndim, output, settings, chain_capacity,
                                            &empty, &g);
    if (st != LUCEM_OK || empty) return st;
    if (!client) return fail(LUCEM_INVALID_HANDLE, "client is NULL");
    if (!input || !psf || !output || !balance_chain)
        return fail(LUCEM_INVALID_HANDLE, "input, psf, output or balance_chain pointer is NULL");
    if (!ws) ws = client->internal_ws.get();
    if (ws->client != client) return fail(LUCEM_INVALID_HANDLE, "workspace belongs to another client");
    st = unsupervised_run(client, ws, input, psf, output, g, *settings, balance_chain);
    if (st == LUCEM_OK && chain_len) *chain_len = settings->num_iter;
    return st;
}
