// CPU oracle for lucem-ops Richardson-Lucy deconvolution.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker the HIP library is compared against; only
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it. Nothing in the
// product path (lucem_amd/, include/) calls it.
//
// It restates, in C++, the reference algorithm literally (direct spatial correlation, no FFT):
//   - pad_index                   crates/lucem-ops/src/core/kernels/pad.rs:11-34
//   - correlate_2d_kernel         crates/lucem-ops/src/filter/kernels/convolve.rs:61-118
//                                 (left = k/2 + origin: filter/convolve.rs:144-145; sequential
//                                 accumulation ky-major then kx: convolve.rs:88-115)
//   - rl_div_clamp_kernel         crates/lucem-ops/src/restore/kernels/deconvolve.rs:7-22
//   - rl_mul_clamp_kernel         crates/lucem-ops/src/restore/kernels/deconvolve.rs:24-43
//   - richardson_lucy_into        crates/lucem-ops/src/restore/deconvolve.rs:45-149
//                                 (check order :53-71, f64 PSF prep :73-90, 0.5 start :93,
//                                 eps rule :104, ping-pong :108-147)
// in FP32 and FP64. The 3-D form (ndim 3, accumulation kz, ky, kx) is this project's own N-D
// extension; the reference RL is 2-D only (deconvolve.rs:55-56).
//
// It also holds the FP64 adjoint (reverse-mode gradient) of unrolled RL followed by the Poisson
// negative log-likelihood of the re-blurred estimate. The reference has no gradient; this is the
// exact derivative of the FP64 forward above and is pinned by central finite differences.
//
// Floating-point contraction is off (-ffp-contract=off in the Makefile) so that every product
// and sum is a separately rounded IEEE operation, in the order the reference kernel writes them.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

enum : int { PAD_REFLECT = 0, PAD_CONSTANT = 1, PAD_NEAREST = 2, PAD_MIRROR = 3, PAD_WRAP = 4 };

// Status codes in the reference's check order (deconvolve.rs:53-71, error.rs:8-59).
enum : int {
    ST_OK = 0,
    ST_SHAPE_MISMATCH = 1,
    ST_ALIASED_BUFFERS = 2,
    ST_INVALID_DIMENSION = 3,
    ST_INVALID_WEIGHTS = 4,
    ST_INVALID_PARAMETER = 5,
    ST_TOO_MANY_ELEMENTS = 6,
    ST_UNKNOWN_PAD_MODE = 8,
};

// core/kernels/pad.rs:11-34. Returns -1 for Constant outside [0, len).
inline int64_t pad_index(int64_t p, int64_t len, int mode) {
    if (p >= 0 && p < len) return p;
    switch (mode) {
        case PAD_CONSTANT: return -1;
        case PAD_NEAREST: return std::max<int64_t>(std::min<int64_t>(p, len - 1), 0);
        case PAD_WRAP: return ((p % len) + len) % len;
        case PAD_REFLECT: {
            int64_t period = 2 * len;
            int64_t r = ((p % period) + period) % period;
            return r < len ? r : period - 1 - r;
        }
        default: {  // Mirror; a length-1 axis has period clamped to 1 (pad.rs:110-114)
            int64_t period = std::max<int64_t>(2 * len - 2, 1);
            int64_t r = ((p % period) + period) % period;
            return r < len ? r : period - r;
        }
    }
}

// Direct correlation of a [n0, n1, n2] array with [k0, k1, k2] weights centred at k/2 + origin
// (filter/kernels/convolve.rs:61-118 extended to three axes). Accumulates per pixel in F in the
// order kz, ky, kx, as the reference's ky-major, kx-minor loop does for 2-D.
template <class F>
void correlate3(const F* in, const int64_t n[3], const F* w, const int64_t k[3], const int64_t org[3],
                int mode, F cval, F* out) {
    const int64_t l0 = k[0] / 2 + org[0], l1 = k[1] / 2 + org[1], l2 = k[2] / 2 + org[2];
    const int64_t rows = n[0] * n[1];
    const int64_t pw = n[2] + k[2] - 1;
    std::vector<int64_t> xmap(pw);
    for (int64_t j = 0; j < pw; j++) xmap[j] = pad_index(j - l2, n[2], mode);
#pragma omp parallel
    {
        std::vector<F> prow(pw), acc(n[2]);
#pragma omp for schedule(static)
        for (int64_t r = 0; r < rows; r++) {
            const int64_t z = r / n[1], y = r % n[1];
            std::fill(acc.begin(), acc.end(), F(0));
            for (int64_t kz = 0; kz < k[0]; kz++) {
                const int64_t sz = pad_index(z - l0 + kz, n[0], mode);
                for (int64_t ky = 0; ky < k[1]; ky++) {
                    const int64_t sy = pad_index(y - l1 + ky, n[1], mode);
                    if (sz < 0 || sy < 0) {
                        std::fill(prow.begin(), prow.end(), cval);
                    } else {
                        const F* src = in + (sz * n[1] + sy) * n[2];
                        for (int64_t j = 0; j < pw; j++) prow[j] = xmap[j] < 0 ? cval : src[xmap[j]];
                    }
                    const F* wr = w + (kz * k[1] + ky) * k[2];
                    F* a = acc.data();
                    const F* pr = prow.data();
                    for (int64_t kx = 0; kx < k[2]; kx++) {
                        const F wv = wr[kx];
                        const F* s = pr + kx;
                        for (int64_t x = 0; x < n[2]; x++) a[x] += wv * s[x];
                    }
                }
            }
            std::memcpy(out + r * n[2], acc.data(), sizeof(F) * n[2]);
        }
    }
}

// Transpose of the padded correlation above (the adjoint used by the gradient oracle):
// out[src] += sum over every padded position p that reads src of w[...] * g[...].
template <class F>
void correlate3_adjoint(const F* g, const int64_t n[3], const F* w, const int64_t k[3], int mode, F* out) {
    const int64_t l0 = k[0] / 2, l1 = k[1] / 2, l2 = k[2] / 2;
    const int64_t total = n[0] * n[1] * n[2];
    std::fill(out, out + total, F(0));
    for (int64_t z = 0; z < n[0]; z++)
        for (int64_t y = 0; y < n[1]; y++)
            for (int64_t x = 0; x < n[2]; x++) {
                const F gv = g[(z * n[1] + y) * n[2] + x];
                for (int64_t kz = 0; kz < k[0]; kz++) {
                    const int64_t sz = pad_index(z - l0 + kz, n[0], mode);
                    if (sz < 0) continue;
                    for (int64_t ky = 0; ky < k[1]; ky++) {
                        const int64_t sy = pad_index(y - l1 + ky, n[1], mode);
                        if (sy < 0) continue;
                        for (int64_t kx = 0; kx < k[2]; kx++) {
                            const int64_t sx = pad_index(x - l2 + kx, n[2], mode);
                            if (sx < 0) continue;
                            out[(sz * n[1] + sy) * n[2] + sx] += w[(kz * k[1] + ky) * k[2] + kx] * gv;
                        }
                    }
                }
            }
}

// d/dw of <g, correlate(x, w)>: dw[kz,ky,kx] = sum_p g[p] * P(x)[p - l + k].
template <class F>
void correlate3_weight_grad(const F* x, const F* g, const int64_t n[3], const int64_t k[3], int mode, F* dw) {
    const int64_t l0 = k[0] / 2, l1 = k[1] / 2, l2 = k[2] / 2;
    const int64_t nk = k[0] * k[1] * k[2];
    for (int64_t q = 0; q < nk; q++) {
        const int64_t kz = q / (k[1] * k[2]), ky = (q / k[2]) % k[1], kx = q % k[2];
        F s = 0;
        for (int64_t z = 0; z < n[0]; z++) {
            const int64_t sz = pad_index(z - l0 + kz, n[0], mode);
            if (sz < 0) continue;
            for (int64_t y = 0; y < n[1]; y++) {
                const int64_t sy = pad_index(y - l1 + ky, n[1], mode);
                if (sy < 0) continue;
                for (int64_t xx = 0; xx < n[2]; xx++) {
                    const int64_t sx = pad_index(xx - l2 + kx, n[2], mode);
                    if (sx < 0) continue;
                    s += g[(z * n[1] + y) * n[2] + xx] * x[(sz * n[1] + sy) * n[2] + sx];
                }
            }
        }
        dw[q] = s;
    }
}

bool shape3(const int64_t* dims, int ndim, int64_t out[3]) {
    if (ndim < 1 || ndim > 3) return false;
    out[0] = out[1] = out[2] = 1;
    for (int i = 0; i < ndim; i++) out[3 - ndim + i] = dims[i];
    return true;
}

bool ranges_overlap(const void* a, size_t abytes, const void* b, size_t bbytes) {
    const char* pa = static_cast<const char*>(a);
    const char* pb = static_cast<const char*>(b);
    return pa < pb + bbytes && pb < pa + abytes;
}

// PSF prep exactly as deconvolve.rs:73-90: f64 sequential sum, norm = 1/sum if |sum| > 1e-12,
// norm_psf = psf * norm, rev_psf = 180-degree flip, both rounded once to F.
template <class F>
void prep_psf(const F* psf, int64_t nk, std::vector<F>& norm_psf, std::vector<F>& rev_psf) {
    double sum = 0.0;
    for (int64_t i = 0; i < nk; i++) sum += static_cast<double>(psf[i]);
    const double norm = std::fabs(sum) > 1e-12 ? 1.0 / sum : 1.0;
    norm_psf.assign(nk, F(0));
    rev_psf.assign(nk, F(0));
    for (int64_t i = 0; i < nk; i++) {
        const double v = static_cast<double>(psf[i]) * norm;
        norm_psf[i] = static_cast<F>(v);
        rev_psf[nk - 1 - i] = static_cast<F>(v);  // flipping every axis reverses the flat index
    }
}

template <class F>
int rl(const F* input, const int64_t* dims, int ndim, const F* psf, const int64_t* kdims, int kndim, F* output,
       const int64_t* out_dims, int out_ndim, uint64_t num_iter, int clip, int has_eps, double eps_in, int mode,
       int max_ndim) {
    // check::same_shape(input, output)
    if (ndim != out_ndim) return ST_SHAPE_MISMATCH;
    for (int i = 0; i < ndim; i++)
        if (dims[i] != out_dims[i]) return ST_SHAPE_MISMATCH;
    int64_t n = 1;
    for (int i = 0; i < ndim; i++) n *= dims[i];
    // check::not_aliased(input, output): buffer overlap
    if (n > 0 && ranges_overlap(input, n * sizeof(F), output, n * sizeof(F))) return ST_ALIASED_BUFFERS;
    // check::ndim_in(input, [2]) then ndim_in(psf, [2]) (N-D form: 2 or 3, same rank as input)
    if (ndim < 2 || ndim > max_ndim) return ST_INVALID_DIMENSION;
    if (kndim < 2 || kndim > max_ndim || (max_ndim > 2 && kndim != ndim)) return ST_INVALID_DIMENSION;
    int64_t nk = 1;
    for (int i = 0; i < kndim; i++) nk *= kdims[i];
    if (nk == 0) return ST_INVALID_WEIGHTS;
    if (num_iter == 0) return ST_INVALID_PARAMETER;
    if (mode < 0 || mode > 4) return ST_UNKNOWN_PAD_MODE;
    if (n == 0) return ST_OK;

    int64_t n3[3], k3[3];
    shape3(dims, ndim, n3);
    shape3(kdims, kndim, k3);
    const int64_t org[3] = {0, 0, 0};
    std::vector<F> norm_psf, rev_psf;
    prep_psf(psf, nk, norm_psf, rev_psf);

    const double e = has_eps ? eps_in : 1.0e-12;
    const F eps = static_cast<F>(std::max(e, 1.0e-15));
    std::vector<F> est_a(n, static_cast<F>(0.5)), est_b(n), est_conv(n), rel_blur(n), err_est(n);
    bool ping = true;
    for (uint64_t iter = 0; iter < num_iter; iter++) {
        const bool is_last = iter + 1 == num_iter;
        const F* src = ping ? est_a.data() : est_b.data();
        F* dst = is_last ? output : (ping ? est_b.data() : est_a.data());
        correlate3<F>(src, n3, rev_psf.data(), k3, org, mode, F(0), est_conv.data());
        for (int64_t i = 0; i < n; i++) {  // rl_div_clamp_kernel
            const F d = est_conv[i];
            const F safe = std::fabs(d) < eps ? eps : d;
            rel_blur[i] = input[i] / safe;
        }
        correlate3<F>(rel_blur.data(), n3, norm_psf.data(), k3, org, mode, F(0), err_est.data());
        const bool do_clip = is_last && clip;
        for (int64_t i = 0; i < n; i++) {  // rl_mul_clamp_kernel
            const F v = src[i] * err_est[i];
            dst[i] = do_clip ? std::min(std::max(v, F(-1)), F(1)) : v;
        }
        ping = !ping;
    }
    return ST_OK;
}

}  // namespace

extern "C" {

int64_t lo_pad_index(int64_t p, int64_t len, int mode) { return pad_index(p, len, mode); }

int lo_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void lo_set_num_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

// Correlation with explicit origin (filter/convolve.rs:125-166), ndim 1..3.
void lo_correlate_f32(const float* in, const int64_t* dims, int ndim, const float* w, const int64_t* kdims,
                      const int64_t* origin, int mode, float cval, float* out) {
    int64_t n3[3], k3[3], o3[3] = {0, 0, 0};
    shape3(dims, ndim, n3);
    shape3(kdims, ndim, k3);
    if (origin)
        for (int i = 0; i < ndim; i++) o3[3 - ndim + i] = origin[i];
    correlate3<float>(in, n3, w, k3, o3, mode, cval, out);
}

void lo_correlate_f64(const double* in, const int64_t* dims, int ndim, const double* w, const int64_t* kdims,
                      const int64_t* origin, int mode, double cval, double* out) {
    int64_t n3[3], k3[3], o3[3] = {0, 0, 0};
    shape3(dims, ndim, n3);
    shape3(kdims, ndim, k3);
    if (origin)
        for (int i = 0; i < ndim; i++) o3[3 - ndim + i] = origin[i];
    correlate3<double>(in, n3, w, k3, o3, mode, cval, out);
}

// richardson_lucy_into, FP32. max_ndim = 2 reproduces the reference's 2-D-only contract;
// max_ndim = 3 is the N-D extension.
int lo_rl_f32(const float* input, const int64_t* dims, int ndim, const float* psf, const int64_t* kdims, int kndim,
              float* output, const int64_t* out_dims, int out_ndim, uint64_t num_iter, int clip, int has_eps,
              double eps, int mode, int max_ndim) {
    return rl<float>(input, dims, ndim, psf, kdims, kndim, output, out_dims, out_ndim, num_iter, clip, has_eps, eps,
                     mode, max_ndim);
}

int lo_rl_f64(const double* input, const int64_t* dims, int ndim, const double* psf, const int64_t* kdims,
              int kndim, double* output, const int64_t* out_dims, int out_ndim, uint64_t num_iter, int clip,
              int has_eps, double eps, int mode, int max_ndim) {
    return rl<double>(input, dims, ndim, psf, kdims, kndim, output, out_dims, out_ndim, num_iter, clip, has_eps,
                      eps, mode, max_ndim);
}

// FP64 forward + reverse-mode gradient of
//   x_0 given;  x_{t+1} = x_t * C_h( y / safe(A x_t) ),  A = correlate with flip(h_n), C_h = correlate with h_n
//   loss = sum( u - y * log(u) ),  u = A x_T
// with h_n = psf / sum(psf) (the reference's normalisation). Returns the loss and writes
// d loss / d x_0 and d loss / d psf (through the normalisation). clip is not applied.
double lo_rl_nll_grad_f64(const double* y, const double* x0, const int64_t* dims, int ndim, const double* psf,
                          const int64_t* kdims, uint64_t num_iter, double eps_in, int mode, double* grad_x0,
                          double* grad_psf) {
    int64_t n3[3], k3[3];
    shape3(dims, ndim, n3);
    shape3(kdims, ndim, k3);
    const int64_t org[3] = {0, 0, 0};
    const int64_t n = n3[0] * n3[1] * n3[2];
    const int64_t nk = k3[0] * k3[1] * k3[2];
    double sum = 0.0;
    for (int64_t i = 0; i < nk; i++) sum += psf[i];
    const double norm = std::fabs(sum) > 1e-12 ? 1.0 / sum : 1.0;
    std::vector<double> h(nk), rev(nk);
    for (int64_t i = 0; i < nk; i++) {
        h[i] = psf[i] * norm;
        rev[nk - 1 - i] = h[i];
    }
    const double eps = std::max(eps_in, 1.0e-15);
    const uint64_t T = num_iter;
    std::vector<std::vector<double>> xs(T + 1, std::vector<double>(n)), us(T + 1, std::vector<double>(n)),
        rs(T, std::vector<double>(n)), cs(T, std::vector<double>(n));
    std::copy(x0, x0 + n, xs[0].begin());
    for (uint64_t t = 0; t < T; t++) {
        correlate3<double>(xs[t].data(), n3, rev.data(), k3, org, mode, 0.0, us[t].data());
        for (int64_t i = 0; i < n; i++) {
            const double d = us[t][i];
            rs[t][i] = y[i] / (std::fabs(d) < eps ? eps : d);
        }
        correlate3<double>(rs[t].data(), n3, h.data(), k3, org, mode, 0.0, cs[t].data());
        for (int64_t i = 0; i < n; i++) xs[t + 1][i] = xs[t][i] * cs[t][i];
    }
    correlate3<double>(xs[T].data(), n3, rev.data(), k3, org, mode, 0.0, us[T].data());
    double loss = 0.0;
    std::vector<double> ubar(n), xbar(n), tmp(n), cbar(n), rbar(n), dw(nk), drev(nk, 0.0), dh(nk, 0.0);
    for (int64_t i = 0; i < n; i++) {
        loss += us[T][i] - y[i] * std::log(us[T][i]);
        ubar[i] = 1.0 - y[i] / us[T][i];
    }
    correlate3_adjoint<double>(ubar.data(), n3, rev.data(), k3, mode, xbar.data());
    correlate3_weight_grad<double>(xs[T].data(), ubar.data(), n3, k3, mode, dw.data());
    for (int64_t q = 0; q < nk; q++) drev[q] += dw[q];
    for (uint64_t tt = T; tt-- > 0;) {
        // x_{t+1} = x_t * c_t
        for (int64_t i = 0; i < n; i++) cbar[i] = xbar[i] * xs[tt][i];
        for (int64_t i = 0; i < n; i++) xbar[i] = xbar[i] * cs[tt][i];
        // c_t = C_h r_t
        correlate3_adjoint<double>(cbar.data(), n3, h.data(), k3, mode, rbar.data());
        correlate3_weight_grad<double>(rs[tt].data(), cbar.data(), n3, k3, mode, dw.data());
        for (int64_t q = 0; q < nk; q++) dh[q] += dw[q];
        // r_t = y / safe(u_t): d r / d u = -y / u^2 where |u| >= eps, else 0
        for (int64_t i = 0; i < n; i++) {
            const double u = us[tt][i];
            ubar[i] = std::fabs(u) < eps ? 0.0 : -rbar[i] * y[i] / (u * u);
        }
        // u_t = A x_t
        correlate3_adjoint<double>(ubar.data(), n3, rev.data(), k3, mode, tmp.data());
        for (int64_t i = 0; i < n; i++) xbar[i] += tmp[i];
        correlate3_weight_grad<double>(xs[tt].data(), ubar.data(), n3, k3, mode, dw.data());
        for (int64_t q = 0; q < nk; q++) drev[q] += dw[q];
    }
    std::copy(xbar.begin(), xbar.end(), grad_x0);
    // rev = flip(h): gradient flows back flipped; h = psf * norm.
    for (int64_t q = 0; q < nk; q++) dh[q] += drev[nk - 1 - q];
    if (std::fabs(sum) > 1e-12) {
        double dot = 0.0;
        for (int64_t q = 0; q < nk; q++) dot += dh[q] * h[q];
        for (int64_t q = 0; q < nk; q++) grad_psf[q] = (dh[q] - dot) * norm;
    } else {
        for (int64_t q = 0; q < nk; q++) grad_psf[q] = dh[q];
    }
    return loss;
}

// Loss only (same definition as above), for finite differences.
double lo_rl_nll_f64(const double* y, const double* x0, const int64_t* dims, int ndim, const double* psf,
                     const int64_t* kdims, uint64_t num_iter, double eps_in, int mode) {
    int64_t n3[3], k3[3];
    shape3(dims, ndim, n3);
    shape3(kdims, ndim, k3);
    const int64_t org[3] = {0, 0, 0};
    const int64_t n = n3[0] * n3[1] * n3[2];
    const int64_t nk = k3[0] * k3[1] * k3[2];
    double sum = 0.0;
    for (int64_t i = 0; i < nk; i++) sum += psf[i];
    const double norm = std::fabs(sum) > 1e-12 ? 1.0 / sum : 1.0;
    std::vector<double> h(nk), rev(nk);
    for (int64_t i = 0; i < nk; i++) {
        h[i] = psf[i] * norm;
        rev[nk - 1 - i] = h[i];
    }
    const double eps = std::max(eps_in, 1.0e-15);
    std::vector<double> x(x0, x0 + n), u(n), r(n), c(n);
    for (uint64_t t = 0; t < num_iter; t++) {
        correlate3<double>(x.data(), n3, rev.data(), k3, org, mode, 0.0, u.data());
        for (int64_t i = 0; i < n; i++) r[i] = y[i] / (std::fabs(u[i]) < eps ? eps : u[i]);
        correlate3<double>(r.data(), n3, h.data(), k3, org, mode, 0.0, c.data());
        for (int64_t i = 0; i < n; i++) x[i] *= c[i];
    }
    correlate3<double>(x.data(), n3, rev.data(), k3, org, mode, 0.0, u.data());
    double loss = 0.0;
    for (int64_t i = 0; i < n; i++) loss += u[i] - y[i] * std::log(u[i]);
    return loss;
}

}  // extern "C"
