/*
 * lucem_hip.h — C ABI of the MI355X (gfx950) Richardson–Lucy deconvolution library.
 *
 * Every entry point takes device pointers owned by the caller, plain sizes and POD settings,
 * enqueues work on the client's HIP stream and returns without blocking. Status codes are
 * returned in the same order as the reference's checks.
 *
 * Reference interfaces replaced (paths in neuro-kitchen/lucem):
 *   lucem_client_*        lucem_core::compute::Client (crates/lucem-core/src/compute/target.rs:95-115)
 *                         and its sync (crates/lucem-core/src/compute/bench.rs:10-13)
 *   lucem_workspace_*     lucem_core::memory::Workspace (crates/lucem-core/src/memory/workspace.rs:17-37)
 *   lucem_status          lucem_ops::error::OpsError (crates/lucem-ops/src/error.rs:8-59) and
 *                         lucem_core::CoreError::{TooManyElements, OutOfDeviceMemory}
 *                         (crates/lucem-core/src/compute/limits.rs:11-22, memory/budget.rs:232-244)
 *   lucem_pad_mode_parse  lucem_ops::core::PadMode::parse (crates/lucem-ops/src/core/pad.rs:48-57)
 *   lucem_rl_settings     lucem_ops::restore::RichardsonLucy (crates/lucem-ops/src/restore/deconvolve.rs:19-41)
 *   lucem_richardson_lucy_into_f32
 *                         lucem_ops::restore::richardson_lucy_into::<f32>
 *                         (crates/lucem-ops/src/restore/deconvolve.rs:45-149)
 *   lucem_richardson_lucy_f32
 *                         lucem_ops::restore::richardson_lucy::<f32> (deconvolve.rs:152-162): the
 *                         output is caller-allocated here, the workspace is the client's own.
 *   lucem_richardson_lucy_into_f64 / lucem_richardson_lucy_f64
 *                         the same two for F = f64 (deconvolve.rs:45-162)
 *   lucem_richardson_lucy_nd_into_f32
 *                         the same operator for 2-D or 3-D arrays (no reference counterpart: the
 *                         reference accepts 2-D only, deconvolve.rs:55-56)
 *   lucem_rl_batch_*      many tiles of uneven size in one graph-captured launch sequence; each
 *                         tile is what a lucem-pipeline custom::transform closure receives
 *                         (engine/lucem-pipeline/src/custom.rs:233-251, Block geometry
 *                         crates/lucem-core/src/tiling.rs:13-23); no reference counterpart
 *   lucem_rl_grad_*       differentiable unrolled RL (forward with checkpoints, adjoint); no
 *                         reference counterpart
 *   lucem_wiener_settings lucem_ops::restore::WienerDeconvolve (deconvolve.rs:164-181)
 *   lucem_wiener_deconvolve_into_f32 / lucem_wiener_deconvolve_f32
 *                         wiener_deconvolve_into / wiener_deconvolve (deconvolve.rs:213-327)
 *   lucem_unsupervised_wiener_settings
 *                         lucem_ops::restore::UnsupervisedWiener (deconvolve.rs:330-353)
 *   lucem_unsupervised_wiener_f32
 *                         unsupervised_wiener (deconvolve.rs:357-509): output caller-allocated,
 *                         the balance chain written to a caller host buffer
 *   lucem_add_noise_f32   lucem_ops::intensity::add_noise (intensity/noise.rs:35-143), NoiseModel
 *                         as lucem_noise_model; output caller-allocated
 *   lucem_fft_1d_into_f32 / lucem_fft_nd_into_f32 / lucem_complex_mul_into_f32
 *                         lucem_ops::core::fft::{fft_1d_into, fft_nd_into, complex_mul_into}
 *                         (crates/lucem-ops/src/core/fft.rs:13-135) on split re/im arrays
 *   lucem_correlate_settings
 *                         lucem_ops::filter::Correlate2d (crates/lucem-ops/src/filter/convolve.rs:113-123)
 *   lucem_correlate_2d_into_f32 / lucem_correlate_2d_into_f64
 *                         lucem_ops::filter::correlate_2d_into (filter/convolve.rs:126-166), the
 *                         correlation A5/A7 of the RL iteration call; f32 through the FFT passes,
 *                         f64 direct
 *   lucem_correlate_nd_into_f32
 *                         the same operator on 2-D or 3-D arrays (no reference counterpart)
 *   lucem_gabor_2d_f32    lucem_ops::filter::gabor_2d (filter/spectral.rs:216-268), on the correlation
 *   lucem_correlate1d_settings / lucem_correlate_1d_into_f32 / lucem_correlate_1d_into_f64
 *                         lucem_ops::filter::{Correlate1d, correlate_1d_into} (filter/convolve.rs:15-75)
 *   lucem_fourier_gaussian_into_f32 / lucem_fourier_box_into_f32 / lucem_fourier_shift_into_f32 /
 *   lucem_butterworth_f32
 *                         lucem_ops::filter::{fourier_gaussian_into, fourier_box_into,
 *                         fourier_shift_into, butterworth} (crates/lucem-ops/src/filter/spectral.rs:
 *                         20-212)
 *   lucem_phase_correlation_f32
 *                         lucem_ops::register::phase_correlation (crates/lucem-ops/src/register/phase.rs:
 *                         18-250), PhaseCorrelation as lucem_phase_correlation_settings; the result
 *                         written to a caller host struct
 *   lucem_render_ellipses_f32 / lucem_render_spheres_f32
 *                         the synthetic workloads' ground truth rendered on the device (no
 *                         reference counterpart; with lucem_add_noise_f32, device-side phantoms
 *                         and Poisson noise)
 */
#ifndef LUCEM_HIP_H
#define LUCEM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. Order and meaning follow OpsError / CoreError. */
typedef enum lucem_status {
    LUCEM_OK = 0,
    LUCEM_SHAPE_MISMATCH = 1,      /* OpsError::ShapeMismatch       (core/check.rs:9-26)  */
    LUCEM_ALIASED_BUFFERS = 2,     /* OpsError::AliasedBuffers      (core/check.rs:47-57) */
    LUCEM_INVALID_DIMENSION = 3,   /* OpsError::InvalidDimension    (core/check.rs:84-91) */
    LUCEM_INVALID_WEIGHTS = 4,     /* OpsError::InvalidWeights      (deconvolve.rs:57-62) */
    LUCEM_INVALID_PARAMETER = 5,   /* OpsError::InvalidParameter    (deconvolve.rs:63-68) */
    LUCEM_TOO_MANY_ELEMENTS = 6,   /* CoreError::TooManyElements    (compute/limits.rs:16-22) */
    LUCEM_OUT_OF_DEVICE_MEMORY = 7,/* CoreError::OutOfDeviceMemory  (memory/budget.rs:232-244) */
    LUCEM_UNKNOWN_PAD_MODE = 8,    /* OpsError::UnknownPadMode      (core/pad.rs:55)      */
    LUCEM_UNSUPPORTED_SHAPE = 9,   /* an axis longer than this library's FFT or launch limits */
    LUCEM_HIP_ERROR = 10,          /* a HIP runtime call failed; see lucem_last_error()   */
    LUCEM_INVALID_HANDLE = 11,     /* NULL client / workspace / settings pointer          */
    LUCEM_AXIS_OUT_OF_BOUNDS = 12, /* CoreError::AxisOutOfBounds    (lucem-core shape.rs:50-58) */
    LUCEM_EMPTY_INPUT = 13,        /* OpsError::EmptyInput          (error.rs:54)         */
    LUCEM_INVALID_ORIGIN = 14      /* OpsError::InvalidOrigin       (core/check.rs:59-72) */
} lucem_status;

/* PadMode ids (core/pad.rs:29-33). RL always uses cval = 0. */
enum {
    LUCEM_PAD_REFLECT = 0,
    LUCEM_PAD_CONSTANT = 1,
    LUCEM_PAD_NEAREST = 2,
    LUCEM_PAD_MIRROR = 3,
    LUCEM_PAD_WRAP = 4
};

/* Largest padded FFT length per axis handled in one LDS-resident pass. */
#define LUCEM_MAX_FFT_LEN 16384

/* RichardsonLucy settings (deconvolve.rs:19-41). Defaults: num_iter 30, clip 1,
 * has_filter_epsilon 0 (=> 1e-12), mode LUCEM_PAD_REFLECT. The epsilon used is
 * max(filter_epsilon or 1e-12, 1e-15) rounded to f32 (deconvolve.rs:104). */
typedef struct lucem_rl_settings {
    uint64_t num_iter;
    int32_t clip;
    int32_t has_filter_epsilon;
    double filter_epsilon;
    uint32_t mode;
    uint32_t reserved;
} lucem_rl_settings;

typedef struct lucem_client lucem_client;
typedef struct lucem_workspace lucem_workspace;

/* The message of the last failing call on this thread ("" if none). */
const char* lucem_last_error(void);
/* Library version string. */
const char* lucem_version(void);

/* Default settings, as RichardsonLucy::default(). */
void lucem_rl_settings_default(lucem_rl_settings* out);
/* PadMode::parse: names and aliases of core/pad.rs:48-57; LUCEM_UNKNOWN_PAD_MODE otherwise. */
lucem_status lucem_pad_mode_parse(const char* name, uint32_t* mode_out);

/* A client binds one device and one HIP stream. own_stream != 0 creates a stream owned by the
 * client (default priority, `stream` ignored); otherwise the client enqueues on `stream` (a
 * hipStream_t; NULL is the device's null stream). Not thread-safe: one client per thread of work. */
lucem_status lucem_client_create(int device, void* stream, int32_t own_stream, lucem_client** out);
lucem_status lucem_client_destroy(lucem_client* client);
/* Waits for every operation queued on the client's stream. */
lucem_status lucem_client_sync(lucem_client* client);
/* The client's hipStream_t. */
void* lucem_client_stream(lucem_client* client);
/* Number of compute units of the client's device. */
int lucem_client_num_cus(lucem_client* client);

/* Per-kernel timing of the client's launches (HIP events on the client stream around each
 * launch) with each launch's algorithmic HBM bytes (the minimum traffic of that pass: its
 * distinct inputs and outputs). Enabling clears earlier records. No reference counterpart
 * (the reference times whole calls, crates/lucem-core/src/compute/bench.rs:17-30). */
typedef struct lucem_kernel_stat {
    char name[48];
    uint64_t launches;
    double total_ms;
    double bytes;
} lucem_kernel_stat;
lucem_status lucem_client_profile(lucem_client* client, int32_t enable);
/* Waits for the stream, aggregates the records by kernel name into out[0..max), sets *count
 * to the number of distinct names, and clears the records. */
lucem_status lucem_client_profile_read(lucem_client* client, lucem_kernel_stat* out, uint32_t max, uint32_t* count);

/* Grow-only device buffers reused across calls (Workspace semantics: one per thread of work,
 * never shared between concurrent calls). */
lucem_status lucem_workspace_create(lucem_client* client, lucem_workspace** out);
lucem_status lucem_workspace_destroy(lucem_workspace* ws);
uint64_t lucem_workspace_capacity_bytes(const lucem_workspace* ws);

/* Device memory budget (crates/lucem-core/src/memory/budget.rs). A limit comes from
 * LUCEM_DEVICE_MEMORY, else LUCEM_DEVICE_MEMORY_LIMIT (read once, budget.rs:246-261; values as
 * parse_bytes takes them, e.g. "7GiB", "6500MB", "7516192768"), or from
 * lucem_set_device_memory_limit (set_device_memory_limit, budget.rs:175-179; has_limit 0 clears it).
 * With a limit, a workspace slot's growth or a batched / gradient plan of 64 MiB or more fails
 * with LUCEM_OUT_OF_DEVICE_MEMORY when it exceeds the limit minus what this library already holds
 * on that device (check_allocation, budget.rs:232-244; message as CoreError::OutOfDeviceMemory).
 * The caller's own arrays are not counted. None of these calls touch the device. */
int32_t lucem_parse_bytes(const char* text, uint64_t* out); /* parse_bytes, budget.rs:263-279; 1 = parsed */
void lucem_set_device_memory_limit(int32_t has_limit, uint64_t bytes);
int32_t lucem_device_memory_limit(uint64_t* out);           /* 1 and *out = the limit, or 0 (no limit) */
uint64_t lucem_device_bytes_in_use(int device);             /* bytes this library holds on `device` */
lucem_status lucem_check_allocation(int device, const char* what, uint64_t bytes);

/* richardson_lucy_into::<f32>. input, psf and output are device pointers, C-contiguous.
 * Checks, in order: same shape (input, output); no overlap of input and output; input ndim == 2;
 * psf ndim == 2; psf non-empty; num_iter >= 1; empty input => LUCEM_OK with nothing done.
 * Correlations run as FFTs on the padded domain n + k - 1 per axis; when that is longer than
 * LUCEM_MAX_FFT_LEN, the call runs the direct form instead (the reference's own correlations and
 * order in FP32), so every 2-D shape the reference accepts is accepted. */
lucem_status lucem_richardson_lucy_into_f32(lucem_client* client, const float* input, const uint64_t* in_dims,
                                            uint32_t in_ndim, const float* psf, const uint64_t* psf_dims,
                                            uint32_t psf_ndim, float* output, const uint64_t* out_dims,
                                            uint32_t out_ndim, const lucem_rl_settings* settings,
                                            lucem_workspace* ws);

/* Allocation-free variant of lucem_richardson_lucy_into_f32. No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). Same checks in the same order and the same launches,
 * but every buffer beyond input, psf and output lives in one fixed device array per GPU of
 * LUCEM_NOALLOC_SCRATCH_BYTES bytes, held by the library until the process ends (set up when a
 * client is first created on the device). The call creates no buffer, event or stream and takes
 * no workspace; launch shapes are the heuristics or the ones this client already measured for the
 * geometry; profiling records nothing. A problem whose buffers exceed the scratch returns
 * LUCEM_UNSUPPORTED_SHAPE (a 1024 x 1024 image with a 25 x 25 PSF needs about 18 MB, 2048 x 2048
 * about 70 MB). All calls on one GPU share its scratch: calls that overlap on one GPU race. */
#define LUCEM_NOALLOC_SCRATCH_BYTES (96ull * 1024ull * 1024ull)
lucem_status lucem_richardson_lucy_into_noalloc_f32(lucem_client* client, const float* input, const uint64_t* in_dims,
                                                    uint32_t in_ndim, const float* psf, const uint64_t* psf_dims,
                                                    uint32_t psf_ndim, float* output, const uint64_t* out_dims,
                                                    uint32_t out_ndim, const lucem_rl_settings* settings);
/* Size of that scratch in bytes (LUCEM_NOALLOC_SCRATCH_BYTES). */
uint64_t lucem_noalloc_scratch_bytes(void);

/* richardson_lucy::<f32> with a caller-allocated output and the client's internal workspace. */
lucem_status lucem_richardson_lucy_f32(lucem_client* client, const float* input, const uint64_t* dims,
                                       uint32_t ndim, const float* psf, const uint64_t* psf_dims, uint32_t psf_ndim,
                                       float* output, const lucem_rl_settings* settings);

/* richardson_lucy_into::<f64> (deconvolve.rs:45-149 with F = f64): the same checks in the same
 * order, FP64 arithmetic throughout, direct correlations (filter/kernels/convolve.rs:61-118) rather
 * than the FFT form of the f32 path. 2-D only, as the reference. */
lucem_status lucem_richardson_lucy_into_f64(lucem_client* client, const double* input, const uint64_t* in_dims,
                                            uint32_t in_ndim, const double* psf, const uint64_t* psf_dims,
                                            uint32_t psf_ndim, double* output, const uint64_t* out_dims,
                                            uint32_t out_ndim, const lucem_rl_settings* settings, lucem_workspace* ws);
/* richardson_lucy::<f64> with a caller-allocated output and the client's internal workspace. */
lucem_status lucem_richardson_lucy_f64(lucem_client* client, const double* input, const uint64_t* dims, uint32_t ndim,
                                       const double* psf, const uint64_t* psf_dims, uint32_t psf_ndim, double* output,
                                       const lucem_rl_settings* settings);
/* Settings of the regularised Wiener filter, WienerDeconvolve (deconvolve.rs:164-181).
 * Defaults (lucem_wiener_settings_default): balance 0.1, clip 1. */
typedef struct lucem_wiener_settings {
    double balance;
    int32_t clip;
    uint32_t reserved;
} lucem_wiener_settings;
void lucem_wiener_settings_default(lucem_wiener_settings* out);

/* wiener_deconvolve_into::<f32> (deconvolve.rs:213-315): X = conj(H) Y / max(|H|^2 + balance |L|^2,
 * 1e-12) on the circular H x W domain of the image, L the centred 3x3 Laplacian, then the
 * inverse transform and the optional clip to [-1, 1]. Checks in the reference's order:
 * same_shape, not_aliased, input 2-D, psf 2-D, psf non-empty, balance finite and >= 0, empty
 * input (OK), psf no larger than the image (LUCEM_INVALID_WEIGHTS). Image sizes: 7-smooth up to
 * 16384 take the R2C passes, others the general FFT (any length up to 2^27 per axis;
 * LUCEM_UNSUPPORTED_SHAPE beyond). Asynchronous on the client's stream. */
lucem_status lucem_wiener_deconvolve_into_f32(lucem_client* client, const float* input, const uint64_t* in_dims,
                                              uint32_t in_ndim, const float* psf, const uint64_t* psf_dims,
                                              uint32_t psf_ndim, float* output, const uint64_t* out_dims,
                                              uint32_t out_ndim, const lucem_wiener_settings* settings,
                                              lucem_workspace* ws);
/* wiener_deconvolve::<f32> (deconvolve.rs:318-327) with a caller-allocated output. */
lucem_status lucem_wiener_deconvolve_f32(lucem_client* client, const float* input, const uint64_t* dims, uint32_t ndim,
                                         const float* psf, const uint64_t* psf_dims, uint32_t psf_ndim, float* output,
                                         const lucem_wiener_settings* settings);

/* Settings of the unsupervised Wiener filter, UnsupervisedWiener (deconvolve.rs:330-353).
 * Defaults (lucem_unsupervised_wiener_settings_default): num_iter 30, burnin 10, clip 1, seed 0. */
typedef struct lucem_unsupervised_wiener_settings {
    uint64_t num_iter;
    uint64_t burnin;
    int32_t clip;
    uint32_t reserved;
    uint64_t seed;
} lucem_unsupervised_wiener_settings;
void lucem_unsupervised_wiener_settings_default(lucem_unsupervised_wiener_settings* out);

/* unsupervised_wiener::<f32> (deconvolve.rs:357-509): the empirical-Bayes loop over num_iter
 * iterations runs on the device; `output` (the input's dims) receives the posterior-mean image
 * and `balance_chain` (host memory, room for num_iter values) the balance of every iteration;
 * *chain_len is set to num_iter (0 for an empty input). Checks in the reference's order: input
 * 2-D, psf 2-D, psf non-empty, num_iter >= 1 and > burnin, empty input (OK); then
 * not_aliased, chain capacity (LUCEM_INVALID_PARAMETER), psf no larger than the image, sizes as
 * for the Wiener filter. Returns after the chain is on the host (synchronises the stream). */
lucem_status lucem_unsupervised_wiener_f32(lucem_client* client, const float* input, const uint64_t* dims,
                                           uint32_t ndim, const float* psf, const uint64_t* psf_dims, uint32_t psf_ndim,
                                           float* output, const lucem_unsupervised_wiener_settings* settings,
                                           double* balance_chain, uint64_t chain_capacity, uint64_t* chain_len,
                                           lucem_workspace* ws);

/* fft_1d_into (core/fft.rs:13-51): the DFT along `axis` (negative counts from the end) of the
 * split complex array (re_in, im_in) into (re_out, im_out); forward exp(-2 pi i k m / n), inverse
 * exp(+...) / n. Checks in the reference's order: the three shapes equal to re_in's, re_in/re_out
 * and im_in/im_out not overlapping, the axis in range (LUCEM_AXIS_OUT_OF_BOUNDS), empty (OK). Any
 * length up to 2^27: one pass for 7-smooth lengths up to 16384 and others up to 8192, four-step
 * passes beyond (7-smooth lengths as two single-pass lengths, others through Bluestein). */
lucem_status lucem_fft_1d_into_f32(lucem_client* client, const float* re_in, const uint64_t* re_in_dims,
                                   uint32_t re_in_ndim, const float* im_in, const uint64_t* im_in_dims,
                                   uint32_t im_in_ndim, float* re_out, const uint64_t* re_out_dims,
                                   uint32_t re_out_ndim, float* im_out, const uint64_t* im_out_dims,
                                   uint32_t im_out_ndim, int64_t axis, int32_t inverse);
/* fft_nd_into (core/fft.rs:53-100): the DFT over every axis, in order. A 0-D array is copied, a
 * 1-D array goes through lucem_fft_1d_into_f32 (and its aliasing check); otherwise the outputs may
 * alias the inputs (two workspace slots hold the intermediate axes). */
lucem_status lucem_fft_nd_into_f32(lucem_client* client, const float* re_in, const uint64_t* re_in_dims,
                                   uint32_t re_in_ndim, const float* im_in, const uint64_t* im_in_dims,
                                   uint32_t im_in_ndim, float* re_out, const uint64_t* re_out_dims,
                                   uint32_t re_out_ndim, float* im_out, const uint64_t* im_out_dims,
                                   uint32_t im_out_ndim, int32_t inverse, lucem_workspace* ws);
/* complex_mul_into (core/fft.rs:102-135): out = a * b, or a * conj(b). Shapes checked as the
 * reference does (b_re and out_re against a_re); the imaginary parts have the same length. */
lucem_status lucem_complex_mul_into_f32(lucem_client* client, const float* a_re, const uint64_t* a_dims,
                                        uint32_t a_ndim, const float* a_im, const float* b_re, const uint64_t* b_dims,
                                        uint32_t b_ndim, const float* b_im, float* out_re, const uint64_t* out_dims,
                                        uint32_t out_ndim, float* out_im, int32_t conj_b);

/* NoiseModel (intensity/noise.rs:17-32): kind and its two parameters. */
enum {
    LUCEM_NOISE_GAUSSIAN = 0,        /* p0 mean, p1 std: x + mean + std * N(0, 1) */
    LUCEM_NOISE_SPECKLE = 1,         /* p0 mean, p1 std: x + x * (mean + std * N(0, 1)) */
    LUCEM_NOISE_SALT = 2,            /* p0 amount: that fraction of pixels set to 1 */
    LUCEM_NOISE_PEPPER = 3,          /* p0 amount: that fraction of pixels set to 0 */
    LUCEM_NOISE_SALT_AND_PEPPER = 4, /* p0 amount, p1 salt_vs_pepper */
    LUCEM_NOISE_POISSON = 5          /* p0 scale: photons per unit intensity (at least 1) */
};
typedef struct lucem_noise_model {
    uint32_t kind;
    uint32_t reserved;
    double p0;
    double p1;
} lucem_noise_model;

/* add_noise::<f32> (intensity/noise.rs:35-143): `output` (the input's dims; it may be the input)
 * receives the noisy image. The per-pixel counter generator and the seed folding are the
 * reference's, so uniform draws and the decisions on them match it bit for bit. An empty input
 * returns LUCEM_EMPTY_INPUT, as the reference's EmptyInput. Asynchronous on the client's stream. */
lucem_status lucem_add_noise_f32(lucem_client* client, const float* input, const uint64_t* dims, uint32_t ndim,
                                 float* output, const lucem_noise_model* model, uint64_t seed, int32_t clip);

/* The same operator for 2-D or 3-D arrays (psf of the same rank as the input). */
lucem_status lucem_richardson_lucy_nd_into_f32(lucem_client* client, const float* input, const uint64_t* in_dims,
                                               uint32_t in_ndim, const float* psf, const uint64_t* psf_dims,
                                               uint32_t psf_ndim, float* output, const uint64_t* out_dims,
                                               uint32_t out_ndim, const lucem_rl_settings* settings,
                                               lucem_workspace* ws);

/* ---- Batched form: many tiles of uneven size in one graph-captured launch sequence ----------
 * Each tile is an independent RL problem (what a custom::transform closure receives,
 * engine/lucem-pipeline/src/custom.rs:233-251; Block read geometry crates/lucem-core/src/tiling.rs:13-23)
 * with its own PSF from a PSF table. All tiles have the PSF table's rank (2 or 3); a 2-D tile uses
 * dims[0..2) = [h, w]. Tiles sharing an FFT shape and a PSF share transfer functions. Every pass
 * of every iteration is one persistent kernel over the work items of all tiles (an atomic work
 * queue, largest tiles first), and the whole sequence is captured once in a hipGraph that
 * lucem_rl_batch_execute replays. Results equal lucem_richardson_lucy_nd_into_f32 per tile. */
typedef struct lucem_rl_tile {
    const float* input;
    float* output;
    uint64_t dims[3];
    uint32_t ndim;
    uint32_t psf_index;
} lucem_rl_tile;

typedef struct lucem_rl_batch lucem_rl_batch;

/* Plans a batch for these tile shapes, PSF shapes (num_psfs x psf_ndim) and settings. The
 * checks of lucem_richardson_lucy_nd_into_f32 apply to every tile (index order), then
 * psf_index < num_psfs (LUCEM_INVALID_PARAMETER). Pointers in `tiles` are not used here. */
lucem_status lucem_rl_batch_create(lucem_client* client, const lucem_rl_tile* tiles, uint64_t num_tiles,
                                   const uint64_t* psf_dims, uint32_t psf_ndim, uint64_t num_psfs,
                                   const lucem_rl_settings* settings, lucem_rl_batch** out);
/* Runs the planned batch on these buffers (same shapes and PSF indices as at creation; device
 * pointers may change from call to call) and `psfs[j]` (device pointers). Enqueues and returns. */
lucem_status lucem_rl_batch_execute(lucem_rl_batch* batch, const lucem_rl_tile* tiles, const float* const* psfs);
/* Where each tile reads its input and stores its result, for tiles that are views into larger
 * arrays (custom::transform's block walk, engine/lucem-pipeline/src/custom.rs:177-196, with the
 * owned-part crop fused into the last store). Strides count elements; 0 means the tile's own
 * contiguous layout. The stored window (origin and size per axis in tile coordinates, z y x; a
 * 2-D tile uses the last two entries; size 0: to the tile's end) is written at `output`, which
 * points at the window's first element. No reference counterpart. */
typedef struct lucem_rl_tile_layout {
    uint64_t in_row_stride, in_plane_stride;
    uint64_t out_row_stride, out_plane_stride;
    uint64_t own_origin[3];
    uint64_t own_size[3];
} lucem_rl_tile_layout;
/* lucem_rl_batch_execute with one layout per tile. A window outside its tile, or a stride shorter
 * than the rows it spans, is LUCEM_INVALID_PARAMETER; a stored range overlapping any tile's input
 * range is LUCEM_ALIASED_BUFFERS (every tile reads its input in every iteration). Stored windows
 * of different tiles must not overlap one another (they are written concurrently). */
lucem_status lucem_rl_batch_execute_layout(lucem_rl_batch* batch, const lucem_rl_tile* tiles,
                                           const lucem_rl_tile_layout* layouts, const float* const* psfs);
lucem_status lucem_rl_batch_destroy(lucem_rl_batch* batch);
/* Device memory the plan holds (workspace, transfer functions, descriptors). */
uint64_t lucem_rl_batch_device_bytes(const lucem_rl_batch* batch);
/* Number of kernel launches one execute replays (the captured graph's nodes). */
uint64_t lucem_rl_batch_launches(const lucem_rl_batch* batch);
/* One call: create, execute, destroy (the plan is released after the stream drains). */
lucem_status lucem_richardson_lucy_batched_f32(lucem_client* client, const lucem_rl_tile* tiles, uint64_t num_tiles,
                                               const float* const* psfs, const uint64_t* psf_dims,
                                               uint32_t psf_ndim, uint64_t num_psfs,
                                               const lucem_rl_settings* settings);

/* ---- Differentiable form --------------------------------------------------------------------
 * loss_t = sum(u_T - y_t log u_T), u_T = A x_T, after T = num_iter RL iterations (no clip) from a
 * given initial estimate x0_t, for num_tiles tiles of rank ndim (2 or 3; shapes may differ,
 * tile_dims is num_tiles x ndim) sharing one PSF. The backward returns d(sum_t w_t loss_t)/d x0_t
 * and d(...)/d psf (through the reference's normalisation psf / sum(psf), deconvolve.rs:77-79).
 * The forward keeps a checkpoint of x every `checkpoint_every` iterations (0: only the start);
 * the backward recomputes the segments before the last from them. x0, y and psf must not change
 * between a forward and its backward. No reference counterpart. */
typedef struct lucem_rl_grad lucem_rl_grad;
lucem_status lucem_rl_grad_create(lucem_client* client, const uint64_t* tile_dims, uint32_t ndim,
                                  uint64_t num_tiles, const uint64_t* psf_dims, const lucem_rl_settings* settings,
                                  uint32_t checkpoint_every, lucem_rl_grad** out);
/* x0[t], y[t]: device pointers; psf: device; x_out: NULL or per-tile device pointers (entries may
 * be NULL) for x_T; loss: NULL or device double[num_tiles]. */
lucem_status lucem_rl_grad_forward(lucem_rl_grad* plan, const float* const* x0, const float* const* y,
                                   const float* psf, float* const* x_out, double* loss);
/* The same with the loss taken against target[t] instead of y[t]: loss_t = sum(u_T - target_t log u_T)
 * (target NULL, or an entry NULL: y). RL still runs on y. With y and target two halves of the
 * same counts (binomial thinning), this held-out likelihood is minimised near the true PSF, where
 * the likelihood of y itself is minimised by a delta PSF. No reference counterpart. */
lucem_status lucem_rl_grad_forward_target(lucem_rl_grad* plan, const float* const* x0, const float* const* y,
                                          const float* const* target, const float* psf, float* const* x_out,
                                          double* loss);
/* loss_weight: NULL (all 1) or device float[num_tiles]; grad_x0[t]: device; grad_psf: NULL or device. */
lucem_status lucem_rl_grad_backward(lucem_rl_grad* plan, const float* loss_weight, float* const* grad_x0,
                                    float* grad_psf);
lucem_status lucem_rl_grad_destroy(lucem_rl_grad* plan);
uint64_t lucem_rl_grad_device_bytes(const lucem_rl_grad* plan);

/* ---- Correlation (filter/convolve.rs:113-166) ------------------------------------------------
 * Correlate2d settings: pad mode (PadMode::id), cval (read outside the array in Constant mode)
 * and the window origin per axis of the input (a 2-D input uses origin[0..2)). Defaults
 * (lucem_correlate_settings_default): Reflect, cval 0, origin 0. */
typedef struct lucem_correlate_settings {
    uint32_t mode;
    int32_t origin[3];
    double cval;
} lucem_correlate_settings;
void lucem_correlate_settings_default(lucem_correlate_settings* out);

/* correlate_2d_into::<f32> (filter/convolve.rs:126-166): out[p] = sum_t w[t] P(in)[p - (k/2 + origin) + t]
 * over the window t, P the boundary mode with cval. Checks in the reference's order: input ndim
 * == 2, weights ndim == 2, same shape (input, output), no overlap of input and output, then per
 * axis the origin (core/check.rs:59-72: a weights axis of size 0 is LUCEM_INVALID_WEIGHTS, an
 * origin outside -(k/2)..=(k-1)/2 LUCEM_INVALID_ORIGIN); an empty input returns LUCEM_OK with
 * nothing done. Windows of up to 169 taps (13 x 13) run the direct sum in the reference's order,
 * the faster form there; larger windows run as an FFT linear correlation (the passes of the RL
 * path) on the padded domain n + k - 1 per axis: FP32 with a rounding error of about 1e-7 of the
 * output's largest value. A domain longer than LUCEM_MAX_FFT_LEN runs the direct sum too (2-D;
 * 3-D returns LUCEM_UNSUPPORTED_SHAPE). Asynchronous on the client's stream; ws may be NULL (the
 * client's own). */
lucem_status lucem_correlate_2d_into_f32(lucem_client* client, const float* input, const uint64_t* in_dims,
                                         uint32_t in_ndim, const float* weights, const uint64_t* w_dims,
                                         uint32_t w_ndim, float* output, const uint64_t* out_dims, uint32_t out_ndim,
                                         const lucem_correlate_settings* settings, lucem_workspace* ws);
/* The same operator on 2-D or 3-D arrays (weights of the input's rank). */
lucem_status lucem_correlate_nd_into_f32(lucem_client* client, const float* input, const uint64_t* in_dims,
                                         uint32_t in_ndim, const float* weights, const uint64_t* w_dims,
                                         uint32_t w_ndim, float* output, const uint64_t* out_dims, uint32_t out_ndim,
                                         const lucem_correlate_settings* settings, lucem_workspace* ws);
/* correlate_2d_into::<f64>: the same checks, the direct sum in FP64 accumulated in the order of
 * filter/kernels/convolve.rs:61-118 (row of the window outer, column inner). */
lucem_status lucem_correlate_2d_into_f64(lucem_client* client, const double* input, const uint64_t* in_dims,
                                         uint32_t in_ndim, const double* weights, const uint64_t* w_dims,
                                         uint32_t w_ndim, double* output, const uint64_t* out_dims,
                                         uint32_t out_ndim, const lucem_correlate_settings* settings);

/* ---- Frequency-domain filters (filter/spectral.rs:20-212) -----------------------------------
 * In place on a split spectrum (re, im), one axis at a time: fourier_gaussian_into multiplies by
 * exp(-2 pi^2 sigma_d^2 f_d^2), fourier_box_into by sinc(size_d f_d) (1 where |pi size f| < 1e-12),
 * fourier_shift_into by exp(-2 pi i shift_d f_d), with f = k / n for k <= n / 2, else (k - n) / n.
 * The transfer values are computed in FP64 and rounded to FP32 as the reference's tables are.
 * Checks in the reference's order: same shape (re, im), then 1 or ndim values
 * (LUCEM_INVALID_PARAMETER), empty (OK). */
lucem_status lucem_fourier_gaussian_into_f32(lucem_client* client, float* re, const uint64_t* re_dims,
                                             uint32_t re_ndim, float* im, const uint64_t* im_dims, uint32_t im_ndim,
                                             const double* sigmas, uint32_t num_sigmas);
lucem_status lucem_fourier_box_into_f32(lucem_client* client, float* re, const uint64_t* re_dims, uint32_t re_ndim,
                                        float* im, const uint64_t* im_dims, uint32_t im_ndim, const double* sizes,
                                        uint32_t num_sizes);
lucem_status lucem_fourier_shift_into_f32(lucem_client* client, float* re, const uint64_t* re_dims, uint32_t re_ndim,
                                          float* im, const uint64_t* im_dims, uint32_t im_ndim, const double* shifts,
                                          uint32_t num_shifts);
/* butterworth::<f32> (spectral.rs:166-212) with a caller-allocated output of the input's shape:
 * the forward FFT, the radial weight 1 / (1 + |f / cutoff|^(2 order)) (or 1 minus it for the
 * high-pass), the inverse FFT's real part. cutoff_ratio outside (0, 0.5] is
 * LUCEM_INVALID_PARAMETER (checked first, as the reference); 1-D to 4-D arrays. */
lucem_status lucem_butterworth_f32(lucem_client* client, const float* input, const uint64_t* in_dims, uint32_t in_ndim,
                                   float* output, const uint64_t* out_dims, uint32_t out_ndim, double cutoff_ratio,
                                   int32_t high_pass, double order, lucem_workspace* ws);

/* ---- Phase correlation (register/phase.rs:18-250) --------------------------------------------
 * Settings as PhaseCorrelation (defaults: upsample_factor 1, normalize_phase 1). */
typedef struct lucem_phase_correlation_settings {
    uint64_t upsample_factor;
    int32_t normalize_phase;
    uint32_t reserved;
} lucem_phase_correlation_settings;
void lucem_phase_correlation_settings_default(lucem_phase_correlation_settings* out);
/* PhaseCorrelationResult: shift[0..ndim) (pixels per axis that register moving onto reference),
 * error (normalised RMS, [0, 1]) and phasediff (radians). */
typedef struct lucem_phase_correlation_result {
    double shift[3];
    uint32_t ndim;
    uint32_t reserved;
    double error;
    double phasediff;
} lucem_phase_correlation_result;
/* phase_correlation::<f32> (phase.rs:52-250) for 1-D to 3-D arrays. Checks in the reference's
 * order: same shape (LUCEM_SHAPE_MISMATCH), non-empty (LUCEM_EMPTY_INPUT), upsample_factor >= 1
 * (LUCEM_INVALID_PARAMETER); then ndim <= 3 (LUCEM_UNSUPPORTED_SHAPE) and the element limit. The
 * transforms run on the device (the general FFT), the subpixel sweeps' local inverse DFTs are
 * FP64 sums on the device. Blocks until the result is written to `out` (host memory), as the
 * reference returns host values. */
lucem_status lucem_phase_correlation_f32(lucem_client* client, const float* reference, const uint64_t* ref_dims,
                                         uint32_t ref_ndim, const float* moving, const uint64_t* mov_dims,
                                         uint32_t mov_ndim, const lucem_phase_correlation_settings* settings,
                                         lucem_phase_correlation_result* out, lucem_workspace* ws);

/* ---- Synthetic phantoms on the device (no reference counterpart) ----------------------------
 * out = background + the sum, in list order and in float64, of the value of every object the
 * pixel lies inside, rounded to float32 once. Inside tests, in float64 with no contraction:
 *   ellipse: u = (dx cos + dy sin) / a, v = (-dx sin + dy cos) / b, u*u + v*v <= 1
 *   sphere:  dz*dz + dy*dy + dx*dx <= radius_sq
 * with d = (pixel index) - (centre). cos_theta, sin_theta and radius_sq are the caller's (so that a
 * host generator and the device agree bit for bit); radius bounds the sphere for culling.
 * `objects` is a DEVICE array of `count` entries; out is [h, w] (ellipses) or [d, h, w] (spheres).
 * Checks, in order: ndim (LUCEM_INVALID_DIMENSION), an empty output (LUCEM_OK, nothing done),
 * element limit (LUCEM_TOO_MANY_ELEMENTS), grid limits (LUCEM_UNSUPPORTED_SHAPE: h <= 1048560, or
 * d <= 262140 and h <= 524280), NULL handles. Asynchronous on the client's stream. */
typedef struct lucem_ellipse {
    double cy, cx, a, b, cos_theta, sin_theta, value;
} lucem_ellipse;
typedef struct lucem_sphere {
    double cz, cy, cx, radius, radius_sq, value;
} lucem_sphere;
lucem_status lucem_render_ellipses_f32(lucem_client* client, float* out, const uint64_t* dims, uint32_t ndim,
                                       const lucem_ellipse* objects, uint64_t count, double background);
lucem_status lucem_render_spheres_f32(lucem_client* client, float* out, const uint64_t* dims, uint32_t ndim,
                                      const lucem_sphere* objects, uint64_t count, double background);

/* Correlate1d settings (filter/convolve.rs:15-29): axis (negative counts from the end), pad mode,
 * window origin, cval. Defaults (lucem_correlate1d_settings_default): axis -1, Reflect, 0, 0. */
typedef struct lucem_correlate1d_settings {
    int64_t axis;
    uint32_t mode;
    int32_t origin;
    double cval;
} lucem_correlate1d_settings;
void lucem_correlate1d_settings_default(lucem_correlate1d_settings* out);
/* correlate_1d_into (filter/convolve.rs:40-75): out[o, i, k] = sum_t w[t] P(in)[o, i + t - left, k]
 * along `axis`, left = size / 2 + origin, the direct sum in the reference's order with its
 * interior fast path (filter/kernels/convolve.rs:8-55). `weights` is a device array of
 * num_weights taps. Checks in the reference's order: same shape, no overlap of input and output,
 * the origin (a window of 0 taps is LUCEM_INVALID_WEIGHTS), the axis (LUCEM_AXIS_OUT_OF_BOUNDS),
 * empty (OK). */
lucem_status lucem_correlate_1d_into_f32(lucem_client* client, const float* input, const uint64_t* dims, uint32_t ndim,
                                         const float* weights, uint64_t num_weights, float* output,
                                         const uint64_t* out_dims, uint32_t out_ndim,
                                         const lucem_correlate1d_settings* settings);
lucem_status lucem_correlate_1d_into_f64(lucem_client* client, const double* input, const uint64_t* dims,
                                         uint32_t ndim, const double* weights, uint64_t num_weights, double* output,
                                         const uint64_t* out_dims, uint32_t out_ndim,
                                         const lucem_correlate1d_settings* settings);

/* gabor_2d::<f32> (filter/spectral.rs:216-268): the complex Gabor kernel of gabor_kernel_2d
 * (FP64 on the host, rounded to f32) correlated with the 2-D input by lucem_correlate_2d_into_f32,
 * real part into out_re and imaginary part into out_im (both caller-allocated, the input's
 * shape). Defaults (lucem_gabor_settings_default): frequency 0.1, theta 0, sigma 1 and 1, n_stds 3,
 * offset 0, Reflect, cval 0. The correlation's checks apply. Blocks until the kernel is uploaded. */
typedef struct lucem_gabor_settings {
    double frequency, theta, sigma_x, sigma_y, n_stds, offset;
    uint32_t mode;
    uint32_t reserved;
    double cval;
} lucem_gabor_settings;
void lucem_gabor_settings_default(lucem_gabor_settings* out);
lucem_status lucem_gabor_2d_f32(lucem_client* client, const float* input, const uint64_t* dims, uint32_t ndim,
                                float* out_re, float* out_im, const lucem_gabor_settings* settings,
                                lucem_workspace* ws);

/* Batched complex-to-complex FFTs of `nlines` contiguous lines of length L (interleaved re/im
 * float pairs, device pointers; src == dst allowed). Forward is exp(-2 pi i j k / L), inverse is
 * unnormalised. L must be 7-smooth and <= LUCEM_MAX_FFT_LEN. The engine the RL passes use; no
 * reference counterpart (the reference's naive DFT is crates/lucem-ops/src/core/fft.rs:13-51). */
lucem_status lucem_fft_c2c_lines_f32(lucem_client* client, const float* src, float* dst, uint32_t L, uint64_t nlines,
                                     int32_t inverse);

#ifdef __cplusplus
}
#endif

#endif /* LUCEM_HIP_H */
