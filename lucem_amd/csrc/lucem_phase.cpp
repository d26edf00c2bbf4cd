// C ABI of phase_correlation (crates/lucem-ops/src/register/phase.rs:52-250): subpixel image
// registration by Fourier phase cross-correlation (Guizar-Sicairos, Thurman and Fienup, 2008) on
// the general FFT of lucem_fft.cpp, which replaces the reference's O(N^2) DFT there.
//
// Device: both spectra, the cross-power spectrum, its inverse transform, the first index of the
// largest |cc|^2. Host: the integer shift, then for upsample_factor > 1 the reference's two
// coordinate sweeps, each candidate's local inverse DFT summed on the device in FP64
// (phase.rs:160-231); the phase at the peak; the RMS error from the images' powers and the
// unnormalised correlation at the peak index (phase.rs:233-257), that one value summed directly
// rather than through a whole inverse transform. The call returns when the result is on the host,
// as the reference's does.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/lucem_hip.h"
#include "lucem_internal_host.hpp"
#include "rl_internal.hpp"

using namespace lucem;

extern "C" void lucem_phase_correlation_settings_default(lucem_phase_correlation_settings* out) {
    if (!out) return;
    out->upsample_factor = 1;
    out->normalize_phase = 1;
    out->reserved = 0;
}

namespace {

constexpr int kMaxCandidates = 64;  // positions per plane-wave launch

struct PhaseRun {
    lucem_client* c;
    int64_t dims3[3];
    int64_t n;
    int nb;
    double* pos;       // device, 3 * kMaxCandidates
    double2* part;     // device, kMaxCandidates * nb
    std::vector<double2> host_part;
};

lucem_status check(int err, const char* what) { return err == hipSuccess ? LUCEM_OK : hip_fail(err, what); }

// sum_idx z[idx] exp(+2 pi i f(idx) . s) for each position s (3 doubles, z y x).
lucem_status plane_wave_sums(PhaseRun& r, const float* re, const float* im, const std::vector<double>& pos,
                             std::vector<double2>* out) {
    const int C = (int)(pos.size() / 3);
    out->assign(C, make_double2(0.0, 0.0));
    for (int c0 = 0; c0 < C; c0 += kMaxCandidates) {
        const int cn = std::min(kMaxCandidates, C - c0);
        lucem_status st;
        hipError_t e = hipMemcpyAsync(r.pos, pos.data() + 3 * (size_t)c0, sizeof(double) * 3 * cn,
                                      hipMemcpyHostToDevice, r.c->stream);
        if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync(positions)");
        if ((st = check(launch_plane_wave_sums(re, im, r.dims3, r.pos, cn, r.part, r.c->stream), "plane-wave sums")))
            return st;
        r.host_part.resize((size_t)cn * r.nb);
        e = hipMemcpyAsync(r.host_part.data(), r.part, sizeof(double2) * r.host_part.size(), hipMemcpyDeviceToHost,
                           r.c->stream);
        if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync(partial sums)");
        if ((e = hipStreamSynchronize(r.c->stream)) != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
        for (int c = 0; c < cn; c++) {
            double sr = 0.0, si = 0.0;
            for (int b = 0; b < r.nb; b++) {
                sr += r.host_part[(size_t)c * r.nb + b].x;
                si += r.host_part[(size_t)c * r.nb + b].y;
            }
            (*out)[c0 + c] = make_double2(sr, si);
        }
    }
    return LUCEM_OK;
}

lucem_status sum_squares(PhaseRun& r, const float* x, double* out) {
    lucem_status st = check(launch_sum_squares(x, r.n, reinterpret_cast<double*>(r.part), r.c->stream), "sum squares");
    if (st) return st;
    std::vector<double> h(r.nb);
    hipError_t e = hipMemcpyAsync(h.data(), r.part, sizeof(double) * r.nb, hipMemcpyDeviceToHost, r.c->stream);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync(sum squares)");
    if ((e = hipStreamSynchronize(r.c->stream)) != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
    double s = 0.0;
    for (double v : h) s += v;
    *out = s;
    return LUCEM_OK;
}

}  // namespace

extern "C" lucem_status lucem_phase_correlation_f32(lucem_client* client, const float* reference,
                                                    const uint64_t* ref_dims, uint32_t ref_ndim, const float* moving,
                                                    const uint64_t* mov_dims, uint32_t mov_ndim,
                                                    const lucem_phase_correlation_settings* settings,
                                                    lucem_phase_correlation_result* out, lucem_workspace* ws) {
    if (!settings || !out) return fail(LUCEM_INVALID_HANDLE, "settings or result pointer is NULL");
    if ((ref_ndim && !ref_dims) || (mov_ndim && !mov_dims)) return fail(LUCEM_INVALID_HANDLE, "dims pointer is NULL");
    // phase.rs:58-68: same shape, non-empty, upsample_factor >= 1
    bool same = ref_ndim == mov_ndim;
    for (uint32_t i = 0; same && i < ref_ndim; i++) same = ref_dims[i] == mov_dims[i];
    if (!same)
        return fail(LUCEM_SHAPE_MISMATCH, "reference shape " + dims_str(ref_dims, ref_ndim) +
                                              " does not match moving shape " + dims_str(mov_dims, mov_ndim));
    uint64_t n = 1;
    bool overflow = false;
    for (uint32_t i = 0; i < ref_ndim; i++) {
        if (ref_dims[i] && n > UINT64_MAX / ref_dims[i]) overflow = true;
        n *= ref_dims[i];
    }
    if (n == 0) return fail(LUCEM_EMPTY_INPUT, "phase_correlation cannot be computed on an empty array");
    if (settings->upsample_factor == 0)
        return fail(LUCEM_INVALID_PARAMETER, "invalid parameter `upsample_factor`: upsample_factor must be >= 1");
    if (ref_ndim > 3)
        return fail(LUCEM_UNSUPPORTED_SHAPE, "phase_correlation takes 1-D to 3-D arrays, got " +
                                                 std::to_string(ref_ndim) + "-D shape " + dims_str(ref_dims, ref_ndim));
    if (overflow || n > 0xFFFFFFFFull)
        return fail(LUCEM_TOO_MANY_ELEMENTS,
                    "dims " + dims_str(ref_dims, ref_ndim) + " exceed the device element limit of 4294967295");
    if (!client) return fail(LUCEM_INVALID_HANDLE, "client is NULL");
    if (!reference || !moving) return fail(LUCEM_INVALID_HANDLE, "reference or moving pointer is NULL");
    if (!ws) ws = client->internal_ws.get();
    if (ws->client != client) return fail(LUCEM_INVALID_HANDLE, "workspace belongs to another client");
    hipError_t e = hipSetDevice(client->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");

    PhaseRun r{};
    r.c = client;
    r.n = (int64_t)n;
    r.dims3[0] = r.dims3[1] = r.dims3[2] = 1;
    for (uint32_t i = 0; i < ref_ndim; i++) r.dims3[3 - ref_ndim + i] = (int64_t)ref_dims[i];
    r.nb = phase_blocks(r.n);
    const size_t fb = sizeof(float) * n;
    float *zero, *ar, *ai, *br, *bi, *pr, *pi, *cr, *ci, *part_v;
    int64_t* part_i;
    lucem_status st;
    if ((st = ws_slot(ws, "pc_zero", fb, (void**)&zero)) || (st = ws_slot(ws, "pc_ar", fb, (void**)&ar)) ||
        (st = ws_slot(ws, "pc_ai", fb, (void**)&ai)) || (st = ws_slot(ws, "pc_br", fb, (void**)&br)) ||
        (st = ws_slot(ws, "pc_bi", fb, (void**)&bi)) || (st = ws_slot(ws, "pc_pr", fb, (void**)&pr)) ||
        (st = ws_slot(ws, "pc_pi", fb, (void**)&pi)) || (st = ws_slot(ws, "pc_cr", fb, (void**)&cr)) ||
        (st = ws_slot(ws, "pc_ci", fb, (void**)&ci)) ||
        (st = ws_slot(ws, "pc_pos", sizeof(double) * 3 * kMaxCandidates, (void**)&r.pos)) ||
        (st = ws_slot(ws, "pc_part", sizeof(double2) * kMaxCandidates * r.nb, (void**)&r.part)) ||
        (st = ws_slot(ws, "pc_best_v", sizeof(float) * r.nb, (void**)&part_v)) ||
        (st = ws_slot(ws, "pc_best_i", sizeof(int64_t) * r.nb, (void**)&part_i)))
        return st;
    hipStream_t sm = client->stream;
    if ((st = check(launch_zero_f32(zero, r.n, sm), "zero"))) return st;
    const uint64_t* d = ref_dims;
    const uint32_t nd = ref_ndim;
    // phase.rs:83-84: both spectra; :88-103: the cross-power spectrum; :105-107: its inverse
    if ((st = lucem_fft_nd_into_f32(client, reference, d, nd, zero, d, nd, ar, d, nd, ai, d, nd, 0, ws))) return st;
    if ((st = lucem_fft_nd_into_f32(client, moving, d, nd, zero, d, nd, br, d, nd, bi, d, nd, 0, ws))) return st;
    if ((st = check(launch_cross_power(ar, ai, br, bi, pr, pi, r.n, settings->normalize_phase ? 1 : 0, sm),
                    "cross power")))
        return st;
    if ((st = lucem_fft_nd_into_f32(client, pr, d, nd, pi, d, nd, cr, d, nd, ci, d, nd, 1, ws))) return st;
    // phase.rs:109-136: the first index of the largest |cc|^2, and cc there
    if ((st = check(launch_argmax_mag2(cr, ci, r.n, part_v, part_i, sm), "argmax"))) return st;
    std::vector<float> hv(r.nb);
    std::vector<int64_t> hi(r.nb);
    if ((e = hipMemcpyAsync(hv.data(), part_v, sizeof(float) * r.nb, hipMemcpyDeviceToHost, sm)) != hipSuccess ||
        (e = hipMemcpyAsync(hi.data(), part_i, sizeof(int64_t) * r.nb, hipMemcpyDeviceToHost, sm)) != hipSuccess ||
        (e = hipStreamSynchronize(sm)) != hipSuccess)
        return hip_fail(e, "argmax read-back");
    int64_t max_idx = hi[0];
    float max_v = hv[0];
    for (int b = 1; b < r.nb; b++)
        if (hv[b] > max_v || (hv[b] == max_v && hi[b] < max_idx)) {
            max_v = hv[b];
            max_idx = hi[b];
        }
    if (max_idx < 0 || max_idx >= r.n) max_idx = 0;  // every |cc|^2 NaN: the reference's kernel starts at 0
    float peak_f[2];
    if ((e = hipMemcpyAsync(&peak_f[0], cr + max_idx, sizeof(float), hipMemcpyDeviceToHost, sm)) != hipSuccess ||
        (e = hipMemcpyAsync(&peak_f[1], ci + max_idx, sizeof(float), hipMemcpyDeviceToHost, sm)) != hipSuccess ||
        (e = hipStreamSynchronize(sm)) != hipSuccess)
        return hip_fail(e, "peak read-back");
    double peak_re = peak_f[0], peak_im = peak_f[1];
    double shift[3] = {0.0, 0.0, 0.0};  // right-aligned (z y x)
    int64_t coords[3];
    {
        int64_t rest = max_idx;
        for (int a = 2; a >= 0; a--) {
            coords[a] = rest % r.dims3[a];
            rest /= r.dims3[a];
        }
        for (int a = 0; a < 3; a++) {
            const int64_t k = coords[a], na = r.dims3[a];
            shift[a] = k <= na / 2 ? (double)k : (double)(k - na);
        }
    }
    // phase.rs:139-231: two coordinate sweeps of the local inverse DFT of the cross-power spectrum
    if (settings->upsample_factor > 1) {
        const double u = (double)settings->upsample_factor;
        const int half_span = (int)std::ceil(1.5 * u);
        std::vector<double> pos;
        std::vector<double2> sums;
        for (int pass = 0; pass < 2; pass++) {
            const int radius = pass == 0 ? half_span : 2;
            const double step = pass == 0 ? 1.0 / u : 0.25 / u;
            for (uint32_t axi = 0; axi < nd; axi++) {
                const int ax = 3 - (int)nd + (int)axi;
                const double center = shift[ax];
                pos.clear();
                for (int m = -radius; m <= radius; m++) {
                    double s3[3] = {shift[0], shift[1], shift[2]};
                    s3[ax] = center + (double)m * step;
                    pos.insert(pos.end(), s3, s3 + 3);
                }
                if ((st = plane_wave_sums(r, pr, pi, pos, &sums))) return st;
                double best = -1.0, best_s = center, best_re = peak_re, best_im = peak_im;
                for (int m = -radius; m <= radius; m++) {
                    const double2 v = sums[m + radius];
                    const double sre = v.x / (double)r.n, sim = v.y / (double)r.n;
                    const double mag2 = sre * sre + sim * sim;
                    if (mag2 > best) {
                        best = mag2;
                        best_s = center + (double)m * step;
                        best_re = sre;
                        best_im = sim;
                    }
                }
                shift[ax] = best_s;
                peak_re = best_re;
                peak_im = best_im;
            }
        }
        for (int a = 0; a < 3; a++) shift[a] = std::round(shift[a] * u) / u;
    }
    const double phasediff = std::atan2(peak_im, peak_re);
    // phase.rs:233-257: the normalised RMS error
    double ref_power = 0.0, mov_power = 0.0;
    if ((st = sum_squares(r, reference, &ref_power)) || (st = sum_squares(r, moving, &mov_power))) return st;
    const double denom = ref_power * mov_power;
    double error = 0.0;
    if (denom > DBL_EPSILON) {
        double ur = peak_re, ui = peak_im;
        if (settings->normalize_phase) {
            // the unnormalised correlation at max_idx: (1/n) sum_k A_k conj(B_k) exp(+2 pi i f(k) . idx)
            if ((st = check(launch_complex_mul(ar, ai, br, bi, pr, pi, r.n, 1, sm), "complex mul"))) return st;
            std::vector<double> p3 = {(double)coords[0], (double)coords[1], (double)coords[2]};
            std::vector<double2> v;
            if ((st = plane_wave_sums(r, pr, pi, p3, &v))) return st;
            ur = v[0].x / (double)r.n;
            ui = v[0].y / (double)r.n;
        }
        const double amp2 = (ur * ur + ui * ui) * (double)r.n * (double)r.n;
        const double q = std::min(std::max(amp2 / denom, 0.0), 1.0);
        error = std::sqrt(1.0 - q);
    }
    for (int i = 0; i < 3; i++) out->shift[i] = 0.0;
    for (uint32_t i = 0; i < nd; i++) out->shift[i] = shift[3 - nd + i];
    out->ndim = nd;
    out->error = error;
    out->phasediff = phasediff;
    return LUCEM_OK;
}
