// Internal interface between the C ABI (lucem_hip.cpp) and the HIP kernels (rl_kernels.hip).
//
// The deconvolution runs every correlation of the reference (filter/kernels/convolve.rs:61-118)
// as an FFT linear correlation on a padded domain L >= n + k - 1 per axis:
//   Q = P_mode(x) gathered on [0, L) with Q[i] = x[pad_index(i - k/2)],
//   corr = IFFT( FFT(Q) * conj(FFT(w zero-padded to L)) ), cropped to [0, n).
// Axis order is (z, y, x), x contiguous. The x axis is transformed two rows at a time
// (z = row_a + i*row_b, one complex FFT of length L_x, then split into two half spectra of
// L_x/2 + 1 bins); the other axes are complex line FFTs over those half spectra.
#pragma once

#include <cstdint>

#include "../../include/lucem_hip.h"

namespace lucem {

constexpr int kMaxPasses = 16;
constexpr int kThreads = 256;         // threads per workgroup of every FFT kernel
constexpr int kElemsPerThread = 16;   // complex values per thread: 4096 per workgroup, 32 KiB LDS
constexpr int kLdsComplex = kThreads * kElemsPerThread;
constexpr int kThreadsWide = 512;     // long lines: 8192 complex per workgroup
constexpr int kLdsComplexWide = kThreadsWide * kElemsPerThread;
constexpr int kThreadsXL = 1024;      // lines up to 16384: 128 KiB of LDS, one workgroup per CU
constexpr int kLdsComplexXL = kThreadsXL * kElemsPerThread;
// Thread-count tier of an FFT kernel (the `wide` field of the descriptors): 0, 1 or 2.
constexpr int tier_threads(int tier) { return tier == 2 ? kThreadsXL : tier == 1 ? kThreadsWide : kThreads; }
constexpr int tier_for_length(int L) { return L <= kLdsComplex ? 0 : L <= kLdsComplexWide ? 1 : 2; }


// One FFT length and its Stockham radix sequence: pass q has radix code byte q of the packed
// words (code = R1 | R2 << 4, R2 = 1 for a simple pass).
struct FftPlan {
    int L;
    int npass;
    int tw;                            // offset (in float2) of this length's twiddle table
    int reserved;
    unsigned long long codes_lo;       // passes 0..7
    unsigned long long codes_hi;       // passes 8..15
};

inline int plan_code(const FftPlan& p, int q) {
    const unsigned long long w = q < 8 ? p.codes_lo : p.codes_hi;
    return (int)((w >> (8 * (q & 7))) & 0xffull);
}

enum RowOp : int {
    ROW_LOAD = 0,        // gather a real array, forward R2C
    ROW_LOAD_CONST = 1,  // a constant array (the 0.5 start), forward R2C
    ROW_RATIO = 2,       // C2R, ratio y / safe(c), forward R2C
    ROW_UPDATE = 3,      // C2R, est' = est * c (clip on the last), forward R2C unless last
};

enum ColOp : int {
    COL_FWD = 0,    // gather along the axis, forward FFT, store all L bins
    COL_FUSED = 1,  // gather, forward FFT, multiply by the transfer function, inverse FFT, crop
    COL_INV = 2,    // load L bins, inverse FFT, crop
    COL_FWD_OTF = 3,  // as COL_FWD, storing conj(v) * scale (transfer function build)
    COL_FUSED_T = 4,  // transposed correlation: gather, FFT, (accumulate), * conj(H), IFFT, fold back
    COL_INV_FOLD = 5, // load L bins, inverse FFT, fold the padded positions back onto [0, n)
};

// Row (x axis) pass over `rows` real rows of length nx, paired two by two.
struct RowPass {
    int op;
    int nx, rows;
    int left, mode;      // pad gather along x
    int Lc, pitch;       // half-spectrum bins per row, row pitch (complex) of `spec`
    int P;               // row pairs per workgroup (any count, P * L <= threads * 16)
    int wide;            // thread tier: 0 = 256, 1 = 512, 2 = 1024 threads
    int first;           // ROW_UPDATE: the previous estimate is the constant `init`
    int last;            // ROW_UPDATE: write `out`, no forward transform
    int clip;            // ROW_UPDATE && last: clamp to [-1, 1]
    float eps;           // ROW_RATIO: |d| < eps => d = eps
    float init;          // ROW_LOAD_CONST / first ROW_UPDATE: the constant start value
    FftPlan f;
    const float* src;    // ROW_LOAD: the real rows
    const float* obs;    // ROW_RATIO: observed image
    float* est;          // ROW_UPDATE: estimate read (unless first) and written (unless last)
    float* out;          // ROW_UPDATE && last: output
    float2* spec;        // half spectra, read (C2R) and written (R2C) in place
    const float2* tw;    // twiddle tables
    // Strided views of the batched form's tile layouts (all zero: the tile's own contiguous
    // arrays). Row r of the tile is (z, y) = (r / ny, r % ny).
    int ny;              // rows per plane
    int strided_obs;     // ROW_RATIO reads obs[z * obs_ps + y * obs_rs + x]
    int window;          // the last ROW_UPDATE stores only own0 <= (z, y, x) < own1, at
                         // out[(z - own0[0]) * out_ps + (y - own0[1]) * out_rs + (x - own0[2])]
    int own0[3], own1[3];
    int64_t obs_rs, obs_ps, out_rs, out_ps;
    int stat;            // ROW_RATIO / ROW_UPDATE: run the compile-time build of (f.L, P)
};

// Line pass along the y or z axis of the half spectra.
struct ColPass {
    int op;
    int n_src;            // rows gathered along the axis (FWD/FUSED) or L (INV)
    int left, mode;       // gather along the axis (FWD/FUSED)
    int n_dst;            // rows stored along the axis
    int Lc;               // valid columns
    int B;                // columns per workgroup (any count, B * L <= threads * 16)
    int nkb;              // column blocks per outer index
    int wide;             // thread tier: 0 = 256, 1 = 512, 2 = 1024 threads
    float scale;          // COL_FWD_OTF
    int64_t src_outer, src_axis, dst_outer, dst_axis, otf_outer, otf_axis;
    FftPlan f;
    const float2* src;
    float2* dst;
    const float2* otf;
    const float2* tw;
    // differentiable form (laid out like otf; slot k at base + k * slot_stride):
    float2* save;           // COL_FUSED: the padded input's spectrum, stored before the multiply
    float2* gacc;           // COL_FUSED_T: gacc += save[slot] * conj(spectrum of the input)
    int64_t slot_stride;
    int fold_left, fold_mode, fold_n;  // COL_FUSED_T / COL_INV_FOLD: the forward's pad (transposed)
    int fold_end;                      // n + k - 1: end of the padded support
    int stat;                          // COL_FUSED: run the compile-time build of (f.L, B)
};

// PSF preparation (deconvolve.rs:73-90 on the device): f64 sum, norm = 1/sum when |sum| > 1e-12
// else 1, h = psf*norm rounded to f32, rev = 180-degree flip of h.
struct PsfPrep {
    int64_t nk;
    const float* psf;
    float* h;
    float* rev;
};

// Batched launches: one kernel over the work items of many tiles. Each workgroup of the
// persistent grid takes items (tile, block) from an atomic counter until none is left.
struct RowBatch {
    const RowPass* tiles;   // per-tile descriptors (geometry, buffers)
    const int2* items;      // (tile, row-pair block)
    int nitems;
    int op, first, last, clip;  // overrides of the per-tile fields for this launch
    int wide;
    int* counter;           // work queue (persistent grid), or null
    int stride;             // counter null: 1 = persistent grid striding over the items, 0 = one item per block
};

struct ColBatch {
    const ColPass* tiles;
    const int2* items;      // (tile, workgroup index within the tile's pass)
    int nitems;
    int op;
    int wide;
    int* counter;           // work queue (persistent grid), or null
    int slot;               // differentiable form: storage slot of save / gacc operands
    int stride;             // counter null: 1 = persistent grid striding over the items, 0 = one item per block
};

struct PsfBatch {
    const PsfPrep* psfs;
    int count;
};

// Launchers (rl_kernels.hip). Each returns the hipError_t of the launch as int.
int launch_row_pass(const RowPass& p, int blocks, void* stream);
int launch_col_pass(const ColPass& p, int blocks, void* stream);
int launch_psf_prep(const PsfPrep& p, void* stream);
int launch_row_batch(const RowBatch& b, int grid, size_t lds, void* stream);
int launch_col_batch(const ColBatch& b, int grid, size_t lds, void* stream);
int launch_psf_batch(const PsfBatch& b, void* stream);
int launch_zero_ints(int* p, int n, void* stream);
// Workgroups of a batched kernel that fit on one CU with `lds` bytes of dynamic LDS.
int batch_occupancy(bool row, int tier, size_t lds);
// Phantoms: background + sum of the objects' values inside which a pixel lies.
int launch_render_ellipses(float* out, int H, int W, const lucem_ellipse* obj, int64_t count, double background,
                           void* stream);
int launch_render_spheres(float* out, int D, int H, int W, const lucem_sphere* obj, int64_t count, double background,
                          void* stream);

// Spectral filters (spectral.hip): per-axis transfer of the signed frequency on [outer, len,
// inner] (kind 0 Gaussian, 1 box sinc, 2 phase shift; parameter p), and the N-D Butterworth
// weight on a right-aligned 4-D shape.
int launch_spectral_axis(float* re, float* im, int64_t total, int64_t len, int64_t inner, int kind, double p,
                         void* stream);
int launch_butterworth_weight(float* re, float* im, const int64_t dims4[4], double cutoff, double order, int high_pass,
                              void* stream);

// Phase correlation (phase.hip): workgroups of the reductions; A * conj(B) (normalised when
// asked); per-workgroup largest re^2 + im^2 and its first index; per-workgroup FP64 sums of a
// spectrum against plane waves exp(+2 pi i f(k) . s) at C positions (3 doubles each, z y x);
// per-workgroup FP64 sums of squares.
int phase_blocks(int64_t n);
int launch_cross_power(const float* ar, const float* ai, const float* br, const float* bi, float* pr, float* pi,
                       int64_t n, int normalize, void* stream);
int launch_argmax_mag2(const float* re, const float* im, int64_t n, float* part_v, int64_t* part_i, void* stream);
int launch_plane_wave_sums(const float* re, const float* im, const int64_t dims3[3], const double* pos, int C,
                           double2* part, void* stream);
int launch_sum_squares(const float* x, int64_t n, double* part, void* stream);

// correlate_1d along an axis viewed as [outer, len, inner] (total elements), window of `size`
// taps starting `left` before each output, `cval` outside the array in Constant mode.
int launch_correlate1d(const float* in, float* out, const float* w, int64_t total, int64_t len, int64_t inner,
                       int64_t left, float cval, int size, int mode, void* stream);
int launch_correlate1d(const double* in, double* out, const double* w, int64_t total, int64_t len, int64_t inner,
                       int64_t left, double cval, int size, int mode, void* stream);

// Constant-mode cval of the f32 correlation: out = in - cval; *sum = sum(w) in f64;
// out += cval * *sum.
int launch_cval_shift(const float* in, float* out, float cval, int64_t n, void* stream);
int launch_weight_sum(const float* w, int64_t nk, double* sum, void* stream);
int launch_add_weight_sum(float* out, const double* sum, double cval, int64_t n, void* stream);

// Direct form (richardson_lucy_into::<f64>, and f32 arrays beyond the FFT length limit):
// out = correlate(in, w) with left offsets (ly, lx) and `cval` outside the array in Constant mode,
// the ratio, the update, a fill, and the f64 PSF normalisation.
int launch_correlate2d_direct(const double* in, const double* w, double* out, int H, int W, int kh, int kw, int ly,
                              int lx, int mode, double cval, void* stream);
int launch_correlate2d_direct(const float* in, const float* w, float* out, int H, int W, int kh, int kw, int ly,
                              int lx, int mode, float cval, void* stream);
int launch_ratio_direct(const double* obs, const double* c, double* rel, double eps, int64_t n, void* stream);
int launch_ratio_direct(const float* obs, const float* c, float* rel, float eps, int64_t n, void* stream);
int launch_update_direct(const double* src, const double* err, double* dst, int clip, int64_t n, void* stream);
int launch_update_direct(const float* src, const float* err, float* dst, int clip, int64_t n, void* stream);
int launch_fill_direct(double* p, double v, int64_t n, void* stream);
int launch_fill_direct(float* p, float v, int64_t n, void* stream);
int launch_psf_prep_f64(const double* psf, double* h, double* rev, int64_t nk, void* stream);
// Complex FFT along one axis of split re/im arrays viewed as [outer, n, inner] (fft_axis.hip).
struct AxisFft {
    const float* re_in;
    const float* im_in;
    float* re_out;
    float* im_out;
    int n;                   // axis length
    int B;                   // lines per workgroup (B * f.L <= threads * 16)
    int64_t inner, outer;    // elements after / slabs before the axis
    int64_t nkb;             // line blocks per slab (inner > 1)
    int inverse;             // conj(DFT(conj(x))) * scale
    int wide;                // thread tier: 0 = 256, 1 = 512, 2 = 1024 threads
    int bluestein;           // n not 7-smooth: circular convolution of length f.L = M
    float scale;             // 1 forward, 1/n inverse
    FftPlan f;
    const float2* tw;
    const float2* chirp;      // Bluestein: exp(-i pi m^2 / n), m < n
    const float2* chirp_hat;  // Bluestein: DFT_M of the conjugate chirp sequence, times 1/M
    // Four-step passes of a longer transform (all zero otherwise). First pass: output k of
    // column j is multiplied by post_tw[((j / post_inner) * k) mod post_n]. Second pass: line
    // o' = o * tr_n1 + k1 stores its output k at ((o * tr_n + k1 + tr_n1 * k) * inner + j), the
    // transposed order of the long axis.
    const float2* post_tw;
    int64_t post_n, post_inner;
    int64_t tr_n1, tr_n;
};
int launch_axis_fft(const AxisFft& d, int blocks, void* stream);
// Bluestein in several passes (lines longer than one workgroup holds), over [outer, M, inner]:
// pre:  u[o][m][j] = m < n ? (x or conj(x))[o][m][j] * chirp[m] : 0
// mid:  u *= chirp_hat[m]
// post: y[o][k][j] = (u[o][k][j] * chirp[k]) (conjugated for the inverse) * scale, k < n
int launch_bluestein_pre(const float* re, const float* im, float* ure, float* uim, const float2* chirp, int64_t outer,
                         int64_t n, int64_t M, int64_t inner, int inverse, void* stream);
int launch_bluestein_mid(float* ure, float* uim, const float2* hat, int64_t outer, int64_t M, int64_t inner,
                         void* stream);
int launch_bluestein_post(const float* ure, const float* uim, float* re, float* im, const float2* chirp, int64_t outer,
                          int64_t n, int64_t M, int64_t inner, int inverse, float scale, void* stream);
size_t axis_fft_lds_bytes(int L, int B);
int launch_complex_mul(const float* ar, const float* ai, const float* br, const float* bi, float* orr, float* oi,
                       int64_t n, int conj_b, void* stream);
int launch_copy_f32(const float* src, float* dst, int64_t n, void* stream);
int allow_big_lds_axis_kernels();

// Wiener forms (wiener.hip). The transfer function of a centred PSF on the circular Hn x Wn
// domain: tables exp(-2 pi i m / N) into tw_h / tw_w, the PSF rows' DFT into A [kh][Lc], then
// either the Wiener filter (filter != 0, float2 at G[u * pitch + v], times scale) or H itself
// (FP64 at Hout[u * Lc + v]).
int launch_wiener_otf(const float* psf, int kh, int kw, int Hn, int Wn, int Lc, int pitch, double2* tw_h,
                      double2* tw_w, double2* A, int filter, double balance, double scale, float2* G, double2* Hout,
                      void* stream);
// Unsupervised Wiener: workgroups of the reductions (at most kUwMaxBlocks), the image moments and
// initial state, one sweep + precision update per iteration, and the posterior mean as the half
// spectrum of the inverse passes.
constexpr int kUwMaxBlocks = 1024;
int uw_blocks(int64_t n);
int launch_uw_init(const float* x, int64_t n, double* part, int nb, double2* mean, int64_t nmean, double* state,
                   void* stream);
// `full`: the spectrum is the whole H x W one (weight 1 everywhere), not the half spectrum.
int launch_uw_iteration(const float2* spec, int pitch, const double2* Hc, const double2* tw_h, const double2* tw_w,
                        int Hn, int Wn, int Lc, double n, double* state, int post, int full, double2* mean, double* part,
                        int nb, double jitter, double* chain, int iter, void* stream);
// General-size route: (re + i im) *= G in place; split planes <-> interleaved (times scale);
// out = re, clamped to [-1, 1] when clip.
int launch_apply_filter(float* re, float* im, const float2* G, int64_t n, void* stream);
int launch_pack_spectrum(const float* re, const float* im, float scale, float2* spec, int64_t n, void* stream);
int launch_unpack_spectrum(const float2* spec, float* re, float* im, int64_t n, void* stream);
int launch_real_part(const float* re, float* out, int64_t n, int clip, void* stream);
int launch_zero_f32(float* p, int64_t n, void* stream);
int launch_uw_finish(const double2* mean, double scale, int Hn, int Lc, int pitch, float2* spec, void* stream);
// add_noise (noise.hip): kind is a LUCEM_NOISE_* id; a, b its parameters as the kernel takes them.
int launch_add_noise(const float* in, float* out, int kind, uint32_t seed, float a, float b, int clip, uint32_t total,
                     void* stream);
// Allows the 1024-thread kernels their 128 KiB of dynamic LDS (once per process).
int allow_big_lds_kernels();

// One whole RL iteration of a 2-D array in one launch (rl_tiled.hip). Workgroup (ty, tx) owns the
// output tile [ty*T0, ty*T0 + T0) x [tx*T1, tx*T1 + T1). Both correlations run as overlap-save FFT
// correlations on an L0 x L1 domain held in LDS, L_a >= T_a + 2 (k_a - 1) per axis:
//   Q1[i] = P(x)(t0 - 2 left + i), i < L        the estimate with the pad rule, 2 halos deep
//   u(t0 - left + p) = corr(Q1, rev)[p],  p < T + k - 1
//   Q2[i] = P(rel)(t0 - left + i) for i < T + k - 1, 0 beyond; rel = y / safe(u)
//   c(t0 + p)        = corr(Q2, h)[p],    p < T
//   x'(t0 + p)       = x(t0 + p) * c(t0 + p)
// so the iteration reads x with a halo of 2 (k - 1), y with one of k - 1, and writes x' once. The
// host admits a geometry only when every pad target of rel lies inside the tile's own u window.
struct TileIter {
    int n0, n1;          // image rows, columns
    int T0, T1;          // output tile extent per axis
    int left0, left1;    // window offsets k / 2
    int k0, k1;          // PSF extent
    int L0, L1;          // FFT lengths (L0 even: rows are transformed in pairs)
    int tiles1;          // tiles along x (blockIdx.x = ty * tiles1 + tx)
    int mode;            // pad mode of both correlations
    int pitch;           // row pitch (complex) of the transfer functions, [L0][pitch]
    int first, last, clip;
    int generic;         // run the kernel compiled for run-time lengths even where a constant one exists
    float eps, init;
    FftPlan f0, f1;      // column (L0) and row (L1) transforms
    const float* est_in;  // estimate (unused when first: the constant init)
    float* est_out;       // next estimate (unless last)
    const float* obs;     // observed image, rows obs_rs elements apart
    float* out;           // output (last): pixels own0 <= (y, x) < own1 at out[(y - own0y) * out_rs + x - own0x]
    int64_t obs_rs, out_rs;
    int own0y, own0x, own1y, own1x;
    const float2* otf1;   // transfer function of rev (est -> u)
    const float2* otf2;   // transfer function of h (rel -> c)
    const float2* tw0;    // twiddle tables of L0 and L1
    const float2* tw1;
};
int launch_tile_iter(const TileIter& d, int blocks, void* stream);
// One iteration of many arrays in one launch: workgroup i runs tile items[i].y of array
// jobs[items[i].x]; the launch sets first / last / clip, and `swap` exchanges each job's est_in and
// est_out (the ping-pong of the iterations). Every job shares the lengths L0 x L1.
struct TileBatch {
    const TileIter* jobs;
    const int2* items;
    int nitems;
    int first, last, clip, swap;
    int L0, L1;
};
int launch_tile_batch(const TileBatch& b, void* stream);
// Whether the tile kernel has a build with these lengths as compile-time constants.
bool tile_len_static(int L0, int L1);
// Allows the tiled kernel its dynamic LDS (once per process, at client creation).
int allow_tile_lds();
size_t tile_iter_lds_bytes(int L0, int L1);
// Threads of the tiled kernel and the most complex values one of its transforms holds (L * lines).
constexpr int kTileThreads = 1024;
constexpr int kTileElems = 9;
constexpr int kTileMaxLen = 128;
// Interleaved lines of the tile's LDS layouts: n data lines, plus one when n is even.
__host__ __device__ constexpr int tile_lines(int n) { return n | 1; }
// Complex values of one of the tile's two LDS buffers: either layout of an L0 x L1 tile.
__host__ __device__ constexpr int tile_buffer_elems(int L0, int L1) {
    return L1 * tile_lines(L0 / 2) > L0 * tile_lines(L1 / 2 + 1) ? L1 * tile_lines(L0 / 2) : L0 * tile_lines(L1 / 2 + 1);
}
size_t row_pass_lds_bytes(int L, int P);
size_t col_pass_lds_bytes(int L, int B);
// A compile-time COL_FUSED build exists for lines of length L in blocks of B columns (256 threads).
bool col_static_available(int L, int B);
// A compile-time COL_FWD / COL_INV build exists for lines of length L in blocks of B columns (512 threads).
bool line_static_available(int L, int B);
// A compile-time ROW_RATIO / ROW_UPDATE build exists for rows of length L, P pairs per workgroup.
bool row_static_available(int L, int P);

}  // namespace lucem
