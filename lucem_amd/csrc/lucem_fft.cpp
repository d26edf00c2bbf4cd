// C ABI of the general complex FFT on split re/im arrays (crates/lucem-ops/src/core/fft.rs):
//
//   fft_1d_into        core/fft.rs:13-51   one axis, forward or inverse (1/n)
//   fft_nd_into        core/fft.rs:53-100  every axis in order, through two workspace slots
//   complex_mul_into   core/fft.rs:102-135 pointwise product, optionally with conj(b)
//
// Each axis transform is one launch of the LDS FFT kernel of fft_axis.hip when a workgroup holds
// the line: the Stockham FFT for 7-smooth lengths up to 16384, Bluestein's convolution
// (power-of-two length up to 16384) for other lengths up to 8192. Longer lines run in passes:
// 7-smooth lengths as a four-step transform of two single-pass lengths, others as Bluestein's
// convolution whose power-of-two transforms are four-step. The reference evaluates every output
// as an O(n) sum with its own sin/cos (core/kernels/fft.rs:11-47).
#include <hip/hip_runtime.h>

#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/lucem_hip.h"
#include "lucem_internal_host.hpp"
#include "rl_internal.hpp"

using namespace lucem;

namespace {

bool smooth7(int v) {
    for (int p : {2, 3, 5, 7})
        while (v % p == 0) v /= p;
    return v == 1;
}

// In-place iterative radix-2 FFT in FP64 (exp(-2 pi i k m / M)), M a power of two.
void fft_pow2(std::vector<std::complex<double>>& a) {
    const size_t M = a.size();
    for (size_t i = 1, j = 0; i < M; i++) {
        size_t bit = M >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(a[i], a[j]);
    }
    const double pi = 3.14159265358979323846264338327950288;
    for (size_t len = 2; len <= M; len <<= 1) {
        for (size_t k = 0; k < len / 2; k++) {
            const std::complex<double> w = std::polar(1.0, -2.0 * pi * (double)k / (double)len);
            for (size_t s = 0; s < M; s += len) {
                const std::complex<double> u = a[s + k], v = a[s + k + len / 2] * w;
                a[s + k] = u + v;
                a[s + k + len / 2] = u - v;
            }
        }
    }
}

// Bluestein tables of length n, cached per client: chirp c_m = exp(-i pi m^2 / n) (m^2 reduced
// mod 2n exactly), then DFT_M of b (b_m = conj(c_|m|) for |m| < n, wrapped mod M) times 1/M.
lucem_status chirp_tables(lucem_client* c, int n, int M, const float2** chirp, const float2** hat) {
    auto it = c->chirps.find(n);
    if (it == c->chirps.end()) {
        const double pi = 3.14159265358979323846264338327950288;
        std::vector<float2> host((size_t)n + M);
        std::vector<std::complex<double>> cm(n), b(M, 0.0);
        for (int m = 0; m < n; m++) {
            const int64_t r = ((int64_t)m * m) % (2 * (int64_t)n);
            cm[m] = std::polar(1.0, -pi * (double)r / (double)n);
            host[m] = make_float2((float)cm[m].real(), (float)cm[m].imag());
        }
        b[0] = std::conj(cm[0]);
        for (int m = 1; m < n; m++) b[m] = b[M - m] = std::conj(cm[m]);
        fft_pow2(b);
        for (int m = 0; m < M; m++) {
            const std::complex<double> v = b[m] / (double)M;
            host[(size_t)n + m] = make_float2((float)v.real(), (float)v.imag());
        }
        float2* dev = nullptr;
        hipError_t e = hipMalloc(&dev, sizeof(float2) * host.size());
        if (e != hipSuccess) return hip_fail(e, "hipMalloc(chirp tables)");
        e = upload_sync(dev, host.data(), sizeof(float2) * host.size(), c->stream);
        if (e != hipSuccess) {
            (void)hipFree(dev);
            return hip_fail(e, "hipMemcpy(chirp tables)");
        }
        it = c->chirps.emplace(n, dev).first;
    }
    *chirp = it->second;
    *hat = it->second + n;
    return LUCEM_OK;
}

std::string shape_str(const uint64_t* d, uint32_t nd) { return dims_str(d, nd); }

bool same_dims(const uint64_t* a, uint32_t na, const uint64_t* b, uint32_t nb) {
    if (na != nb) return false;
    for (uint32_t i = 0; i < na; i++)
        if (a[i] != b[i]) return false;
    return true;
}

lucem_status shape_check(const char* an, const uint64_t* a, uint32_t na, const char* bn, const uint64_t* b,
                         uint32_t nb) {
    if ((na && !a) || (nb && !b)) return fail(LUCEM_INVALID_HANDLE, "dims pointer is NULL");
    if (same_dims(a, na, b, nb)) return LUCEM_OK;
    return fail(LUCEM_SHAPE_MISMATCH, std::string(an) + " shape " + shape_str(a, na) + " does not match " + bn +
                                          " shape " + shape_str(b, nb));
}

uint64_t elems(const uint64_t* d, uint32_t nd, bool* overflow) {
    uint64_t n = 1;
    for (uint32_t i = 0; i < nd; i++) {
        if (d[i] && n > UINT64_MAX / d[i]) *overflow = true;
        n *= d[i];
    }
    return n;
}

bool overlap(const void* a, const void* b, uint64_t bytes) {
    if (!a || !b) return false;
    const char* x = (const char*)a;
    const char* y = (const char*)b;
    return x < y + bytes && y < x + bytes;
}

lucem_status element_limit(uint64_t n, bool overflow, const uint64_t* d, uint32_t nd) {
    if (overflow || n > 0xFFFFFFFFull)
        return fail(LUCEM_TOO_MANY_ELEMENTS, "dims " + shape_str(d, nd) + " exceed the device element limit of 4294967295");
    return LUCEM_OK;
}

}  // namespace

// Descriptor and grid of one axis transform of an array viewed as [outer, n, inner].
lucem_status lucem::axis_fft_desc(lucem_client* c, int n, int64_t outer, int64_t inner, bool inverse, AxisFft* d,
                                  int* blocks) {
    std::memset(d, 0, sizeof *d);
    int L = n;
    if (n > LUCEM_MAX_FFT_LEN || !smooth7(n)) {
        int M = 1;
        while (M < 2 * n - 1) M <<= 1;
        if (M > LUCEM_MAX_FFT_LEN)
            return fail(LUCEM_UNSUPPORTED_SHAPE, "FFT length " + std::to_string(n) +
                                                     ": 7-smooth lengths up to 16384 and others up to 8192 are supported");
        L = M;
        d->bluestein = 1;
        lucem_status st = chirp_tables(c, n, M, &d->chirp, &d->chirp_hat);
        if (st) return st;
    }
    lucem_status st = twiddles_for(c, L, &d->tw);
    if (st) return st;
    d->n = n;
    d->f = make_plan(L);
    d->inner = inner;
    d->outer = outer;
    d->inverse = inverse ? 1 : 0;
    d->scale = inverse ? (float)(1.0 / (double)n) : 1.0f;
    // Contiguous lines fill the smallest thread tier that holds them. Strided lines want at least
    // 8 adjacent columns per workgroup (32-byte accesses): a wider tier when the short one holds
    // fewer, as the RL line passes do.
    d->wide = tier_for_length(L);
    const int64_t line_dim = inner == 1 ? outer : inner;
    const int64_t want = std::min<int64_t>(8, line_dim);
    if (inner > 1)
        while (d->wide < 2 && tier_threads(d->wide) * kElemsPerThread / L < want) d->wide++;
    const int cap = tier_threads(d->wide) * kElemsPerThread;
    int B = (int)std::max<int64_t>(1, std::min<int64_t>(cap / L, line_dim));
    const int64_t lines = outer * inner;
    while (B > 4 && lines / B < c->num_cus) B--;
    d->B = B;
    int64_t nb;
    if (inner == 1) {
        nb = (outer + B - 1) / B;
    } else {
        d->nkb = (inner + B - 1) / B;
        nb = outer * d->nkb;
    }
    if (nb > 0x7FFFFFFF) return fail(LUCEM_TOO_MANY_ELEMENTS, "too many FFT lines for one launch");
    *blocks = (int)nb;
    return LUCEM_OK;
}

namespace {

// exp(-2 pi i m / n), m < n, computed in FP64 and cached per client: the twiddles between the
// two passes of a four-step transform of length n.
lucem_status long_twiddles(lucem_client* c, int64_t n, const float2** out) {
    auto it = c->long_twiddles.find(n);
    if (it == c->long_twiddles.end()) {
        const double pi = 3.14159265358979323846264338327950288;
        std::vector<float2> host((size_t)n);
        for (int64_t m = 0; m < n; m++) {
            const std::complex<double> w = std::polar(1.0, -2.0 * pi * (double)m / (double)n);
            host[(size_t)m] = make_float2((float)w.real(), (float)w.imag());
        }
        float2* dev = nullptr;
        hipError_t e = hipMalloc(&dev, sizeof(float2) * host.size());
        if (e != hipSuccess) return hip_fail(e, "hipMalloc(four-step twiddles)");
        e = upload_sync(dev, host.data(), sizeof(float2) * host.size(), c->stream);
        if (e != hipSuccess) {
            (void)hipFree(dev);
            return hip_fail(e, "hipMemcpy(four-step twiddles)");
        }
        it = c->long_twiddles.emplace(n, dev).first;
    }
    *out = it->second;
    return LUCEM_OK;
}

// One launch of the single-pass kernel over [outer, n, inner].
lucem_status one_pass(lucem_client* c, const float* ri, const float* ii, float* ro, float* io, int64_t outer, int n,
                      int64_t inner, bool inverse, bool unit_scale, const AxisFft* extra) {
    AxisFft d;
    int blocks = 0;
    lucem_status st = axis_fft_desc(c, n, outer, inner, inverse, &d, &blocks);
    if (st) return st;
    if (extra) {
        d.post_tw = extra->post_tw;
        d.post_n = extra->post_n;
        d.post_inner = extra->post_inner;
        d.tr_n1 = extra->tr_n1;
        d.tr_n = extra->tr_n;
    }
    if (unit_scale) d.scale = 1.0f;
    d.re_in = ri;
    d.im_in = ii;
    d.re_out = ro;
    d.im_out = io;
    const int err = launch_axis_fft(d, blocks, c->stream);
    return err == hipSuccess ? LUCEM_OK : hip_fail(err, "fft axis pass");
}

// The 7-smooth factor of n (n 7-smooth) closest to sqrt(n) with both factors at most the
// single-pass limit; 0 when there is none.
int64_t split_smooth(int64_t n) {
    int64_t best = 0;
    for (int64_t a = 2; a <= LUCEM_MAX_FFT_LEN && a < n; a++) {
        if (n % a || n / a > LUCEM_MAX_FFT_LEN || !smooth7((int)a)) continue;
        if (!best || std::llabs(a * a - n) < std::llabs(best * best - n)) best = a;
    }
    return best;
}

lucem_status fft_lines(lucem_client* c, const float* ri, const float* ii, float* ro, float* io, int64_t outer,
                       int64_t n, int64_t inner, bool inverse, bool unit_scale);

// Four-step transform of the 7-smooth length n = n1 * n2 (Bailey): with j = j1 * n2 + j2 and
// k = k1 + n1 * k2, X[k] = sum_j2 w_n2^(j2 k2) w_n^(j2 k1) sum_j1 x[j] w_n1^(j1 k1). Pass 1: the
// length-n1 transforms over j1 (a strided view), times w_n^(j2 k1), into a scratch buffer; pass 2:
// the length-n2 transforms over j2, stored in the transposed order k1 + n1 * k2.
lucem_status four_step(lucem_client* c, const float* ri, const float* ii, float* ro, float* io, int64_t outer,
                       int64_t n, int64_t inner, bool inverse, bool unit_scale) {
    const int64_t n1 = split_smooth(n);
    if (!n1)
        return fail(LUCEM_UNSUPPORTED_SHAPE, "FFT length " + std::to_string(n) + ": no split into two lengths of at "
                                                                                   "most 16384");
    const int64_t n2 = n / n1;
    const int64_t total = outer * n * inner;
    float *tr, *ti;
    lucem_status st;
    lucem_workspace* ws = c->internal_ws.get();
    if ((st = ws_slot(ws, "fft4_re", sizeof(float) * total, (void**)&tr))) return st;
    if ((st = ws_slot(ws, "fft4_im", sizeof(float) * total, (void**)&ti))) return st;
    AxisFft x;
    std::memset(&x, 0, sizeof x);
    if ((st = long_twiddles(c, n, &x.post_tw))) return st;
    x.post_n = n;
    x.post_inner = inner;
    if ((st = one_pass(c, ri, ii, tr, ti, outer, (int)n1, n2 * inner, inverse, unit_scale, &x))) return st;
    std::memset(&x, 0, sizeof x);
    x.tr_n1 = n1;
    x.tr_n = n;
    return one_pass(c, tr, ti, ro, io, outer * n1, (int)n2, inner, inverse, unit_scale, &x);
}

// Bluestein's convolution in several passes, for lengths whose power-of-two convolution M is
// longer than one workgroup holds: the chirped input zero-padded to M, its transform (four-step),
// the product with the chirp filter's transform, the unnormalised inverse, the output chirp.
lucem_status bluestein_long(lucem_client* c, const float* ri, const float* ii, float* ro, float* io, int64_t outer,
                            int64_t n, int64_t inner, bool inverse, bool unit_scale) {
    int64_t M = 1;
    while (M < 2 * n - 1) M <<= 1;
    if (M > (int64_t)LUCEM_MAX_FFT_LEN * LUCEM_MAX_FFT_LEN)
        return fail(LUCEM_UNSUPPORTED_SHAPE, "FFT length " + std::to_string(n) + " is too long");
    const float2 *chirp, *hat;
    lucem_status st = chirp_tables(c, (int)n, (int)M, &chirp, &hat);
    if (st) return st;
    float *ur, *ui;
    lucem_workspace* ws = c->internal_ws.get();
    const int64_t total = outer * M * inner;
    if ((st = ws_slot(ws, "fftb_re", sizeof(float) * total, (void**)&ur))) return st;
    if ((st = ws_slot(ws, "fftb_im", sizeof(float) * total, (void**)&ui))) return st;
    auto check = [](int err, const char* what) { return err == hipSuccess ? LUCEM_OK : hip_fail(err, what); };
    if ((st = check(launch_bluestein_pre(ri, ii, ur, ui, chirp, outer, n, M, inner, inverse ? 1 : 0, c->stream),
                    "bluestein pre")))
        return st;
    if ((st = fft_lines(c, ur, ui, ur, ui, outer, M, inner, false, true))) return st;
    if ((st = check(launch_bluestein_mid(ur, ui, hat, outer, M, inner, c->stream), "bluestein mid"))) return st;
    if ((st = fft_lines(c, ur, ui, ur, ui, outer, M, inner, true, true))) return st;
    const float scale = inverse && !unit_scale ? (float)(1.0 / (double)n) : 1.0f;
    return check(launch_bluestein_post(ur, ui, ro, io, chirp, outer, n, M, inner, inverse ? 1 : 0, scale, c->stream),
                 "bluestein post");
}

// The transform of every line of [outer, n, inner] (forward, or inverse with 1/n unless
// unit_scale): one pass when a workgroup holds the line, else four-step or Bluestein in passes.
// ro/io may alias ri/ii.
lucem_status fft_lines(lucem_client* c, const float* ri, const float* ii, float* ro, float* io, int64_t outer,
                       int64_t n, int64_t inner, bool inverse, bool unit_scale) {
    const bool smooth = n <= 0x7FFFFFFF && smooth7((int)n);
    if ((smooth && n <= LUCEM_MAX_FFT_LEN) || (!smooth && n <= LUCEM_MAX_FFT_LEN / 2))
        return one_pass(c, ri, ii, ro, io, outer, (int)n, inner, inverse, unit_scale, nullptr);
    if (smooth && split_smooth(n)) return four_step(c, ri, ii, ro, io, outer, n, inner, inverse, unit_scale);
    return bluestein_long(c, ri, ii, ro, io, outer, n, inner, inverse, unit_scale);
}

}  // namespace

static lucem_status run_axis(lucem_client* c, const float* ri, const float* ii, float* ro, float* io, const uint64_t* dims,
                             uint32_t ndim, uint32_t axis, bool inverse) {
    int64_t outer = 1, inner = 1;
    for (uint32_t a = 0; a < axis; a++) outer *= (int64_t)dims[a];
    for (uint32_t a = axis + 1; a < ndim; a++) inner *= (int64_t)dims[a];
    return fft_lines(c, ri, ii, ro, io, outer, (int64_t)dims[axis], inner, inverse, false);
}

extern "C" lucem_status lucem_fft_1d_into_f32(lucem_client* client, const float* re_in, const uint64_t* re_in_dims,
                                              uint32_t re_in_ndim, const float* im_in, const uint64_t* im_in_dims,
                                              uint32_t im_in_ndim, float* re_out, const uint64_t* re_out_dims,
                                              uint32_t re_out_ndim, float* im_out, const uint64_t* im_out_dims,
                                              uint32_t im_out_ndim, int64_t axis, int32_t inverse) {
    lucem_status st;
    if ((st = shape_check("re_in", re_in_dims, re_in_ndim, "im_in", im_in_dims, im_in_ndim)) ||
        (st = shape_check("re_in", re_in_dims, re_in_ndim, "re_out", re_out_dims, re_out_ndim)) ||
        (st = shape_check("re_in", re_in_dims, re_in_ndim, "im_out", im_out_dims, im_out_ndim)))
        return st;
    bool overflow = false;
    const uint64_t n = elems(re_in_dims, re_in_ndim, &overflow);
    if (n > 0 && !overflow) {
        if (overlap(re_in, re_out, n * sizeof(float)) || overlap(im_in, im_out, n * sizeof(float)))
            return fail(LUCEM_ALIASED_BUFFERS, "input and output arrays must not share the same device buffer");
    }
    const int64_t nd = (int64_t)re_in_ndim;
    const int64_t ax = axis < 0 ? axis + nd : axis;
    if (ax < 0 || ax >= nd)
        return fail(LUCEM_AXIS_OUT_OF_BOUNDS,
                    "axis " + std::to_string(axis) + " is out of bounds for " + std::to_string(nd) + "-D array");
    if (n == 0) return LUCEM_OK;
    if ((st = element_limit(n, overflow, re_in_dims, re_in_ndim))) return st;
    if (!client) return fail(LUCEM_INVALID_HANDLE, "client is NULL");
    if (!re_in || !im_in || !re_out || !im_out) return fail(LUCEM_INVALID_HANDLE, "array pointer is NULL");
    hipError_t e = hipSetDevice(client->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    return run_axis(client, re_in, im_in, re_out, im_out, re_in_dims, re_in_ndim, (uint32_t)ax, inverse != 0);
}

extern "C" lucem_status lucem_fft_nd_into_f32(lucem_client* client, const float* re_in, const uint64_t* re_in_dims,
                                              uint32_t re_in_ndim, const float* im_in, const uint64_t* im_in_dims,
                                              uint32_t im_in_ndim, float* re_out, const uint64_t* re_out_dims,
                                              uint32_t re_out_ndim, float* im_out, const uint64_t* im_out_dims,
                                              uint32_t im_out_ndim, int32_t inverse, lucem_workspace* ws) {
    lucem_status st;
    if ((st = shape_check("re_in", re_in_dims, re_in_ndim, "im_in", im_in_dims, im_in_ndim)) ||
        (st = shape_check("re_in", re_in_dims, re_in_ndim, "re_out", re_out_dims, re_out_ndim)) ||
        (st = shape_check("re_in", re_in_dims, re_in_ndim, "im_out", im_out_dims, im_out_ndim)))
        return st;
    const uint32_t ndim = re_in_ndim;
    if (ndim == 1)  // fft_1d_into along axis 0, with its aliasing check
        return lucem_fft_1d_into_f32(client, re_in, re_in_dims, ndim, im_in, im_in_dims, ndim, re_out, re_out_dims,
                                     ndim, im_out, im_out_dims, ndim, 0, inverse);
    bool overflow = false;
    const uint64_t n = elems(re_in_dims, ndim, &overflow);
    if (n == 0) return LUCEM_OK;
    if ((st = element_limit(n, overflow, re_in_dims, ndim))) return st;
    if (!client) return fail(LUCEM_INVALID_HANDLE, "client is NULL");
    if (!re_in || !im_in || !re_out || !im_out) return fail(LUCEM_INVALID_HANDLE, "array pointer is NULL");
    hipError_t e = hipSetDevice(client->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    if (ndim == 0) {  // a scalar: copied (core/fft.rs:69-72)
        int err = launch_copy_f32(re_in, re_out, 1, client->stream);
        if (err == hipSuccess) err = launch_copy_f32(im_in, im_out, 1, client->stream);
        return err == hipSuccess ? LUCEM_OK : hip_fail(err, "fft copy");
    }
    // axis 0 from the input into the ping slots, the middle axes in place there, the last axis
    // into the output: outputs may alias the inputs in any way
    if (!ws) ws = client->internal_ws.get();
    if (ws->client != client) return fail(LUCEM_INVALID_HANDLE, "workspace belongs to another client");
    float *pr, *pi;
    if ((st = ws_slot(ws, "fft_ping_re", sizeof(float) * n, (void**)&pr))) return st;
    if ((st = ws_slot(ws, "fft_ping_im", sizeof(float) * n, (void**)&pi))) return st;
    const bool inv = inverse != 0;
    if ((st = run_axis(client, re_in, im_in, pr, pi, re_in_dims, ndim, 0, inv))) return st;
    for (uint32_t a = 1; a + 1 < ndim; a++)
        if ((st = run_axis(client, pr, pi, pr, pi, re_in_dims, ndim, a, inv))) return st;
    return run_axis(client, pr, pi, re_out, im_out, re_in_dims, ndim, ndim - 1, inv);
}

extern "C" lucem_status lucem_complex_mul_into_f32(lucem_client* client, const float* a_re, const uint64_t* a_dims,
                                                   uint32_t a_ndim, const float* a_im, const float* b_re,
                                                   const uint64_t* b_dims, uint32_t b_ndim, const float* b_im,
                                                   float* out_re, const uint64_t* out_dims, uint32_t out_ndim,
                                                   float* out_im, int32_t conj_b) {
    lucem_status st;
    if ((st = shape_check("a_re", a_dims, a_ndim, "b_re", b_dims, b_ndim)) ||
        (st = shape_check("a_re", a_dims, a_ndim, "out_re", out_dims, out_ndim)))
        return st;
    bool overflow = false;
    const uint64_t n = elems(a_dims, a_ndim, &overflow);
    if (n == 0) return LUCEM_OK;
    if ((st = element_limit(n, overflow, a_dims, a_ndim))) return st;
    if (!client) return fail(LUCEM_INVALID_HANDLE, "client is NULL");
    if (!a_re || !a_im || !b_re || !b_im || !out_re || !out_im)
        return fail(LUCEM_INVALID_HANDLE, "array pointer is NULL");
    hipError_t e = hipSetDevice(client->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    const int err = launch_complex_mul(a_re, a_im, b_re, b_im, out_re, out_im, (int64_t)n, conj_b ? 1 : 0,
                                       client->stream);
    return err == hipSuccess ? LUCEM_OK : hip_fail(err, "complex_mul");
}
