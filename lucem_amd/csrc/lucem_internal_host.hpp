// Host-side internals shared by the C ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/lucem_hip.h"
#include "rl_internal.hpp"

struct lucem_workspace;

namespace lucem {

// Host-to-device upload ordered on `s` and waited for. Unlike hipMemcpy it does not touch the
// legacy stream, so it may run while another thread captures a graph.
inline hipError_t upload_sync(void* dst, const void* src, size_t bytes, hipStream_t s) {
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s);
    return e == hipSuccess ? hipStreamSynchronize(s) : e;
}

// One profiled launch: events on the client stream around it and its algorithmic HBM bytes.
struct ProfRecord {
    const char* name;
    hipEvent_t start, stop;
    double bytes;
};

// Launch shape of one pass kind at one FFT length: lines per workgroup, the 512-thread line
// kernel, radix-sequence variant.
struct ShapeChoice {
    int lines = 0;
    int wide = 0;
    int variant = 0;
};
enum ShapeKind : int { SHAPE_ROW = 0, SHAPE_COL_FWD = 1, SHAPE_COL_FUSED = 2, SHAPE_COL_INV = 3 };
using ShapeTable = std::map<std::pair<int, int>, ShapeChoice>;  // (kind, L) -> choice
}  // namespace lucem

struct lucem_client {
    // One reference for the caller's handle and one for every workspace and plan made on the
    // client: lucem_client_destroy drops the caller's, and whichever release comes last frees it.
    std::atomic<int> refs{1};
    int device = 0;
    int num_cus = 256;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::map<int, float2*> twiddles;               // FFT length -> device twiddle table
    std::map<int, float2*> chirps;                 // Bluestein length n -> chirp [n] + its transform [M]
    std::map<int64_t, float2*> long_twiddles;      // four-step length n -> exp(-2 pi i m / n), m < n
    std::unique_ptr<lucem_workspace> internal_ws;  // used by the allocating-style entry points
    bool profiling = false;
    std::vector<lucem::ProfRecord> records;        // launches since profiling was enabled
    std::vector<hipEvent_t> event_pool;            // events of read records, reused
    std::map<std::vector<int>, lucem::ShapeTable> tuned;  // plain path: tuned shapes per geometry
    ~lucem_client();
};

namespace lucem {
struct Slot {
    void* ptr = nullptr;
    size_t bytes = 0;
};
}  // namespace lucem

struct lucem_workspace {
    lucem_client* client = nullptr;
    bool internal = false;  // the client's own workspace: holds no reference to it
    std::map<std::string, lucem::Slot> slots;
    ~lucem_workspace();
};

namespace lucem {

// Client references held by workspaces and plans (see lucem_client::refs).
void client_retain(lucem_client* c);
void client_release(lucem_client* c);
// True once the process has begun to exit and the HIP runtime may be torn down: the mark is set
// by the destructor of a static made right after the runtime's first use, so it runs before the
// runtime's own teardown. Destroy calls made after it free host memory only and call no HIP API.
bool exiting();

// Geometry of one RL array: sizes per axis (z, y, x), FFT lengths and half-spectrum layout.
struct Geometry {
    int ndim;
    int n[3];
    int k[3];
    int L[3];
    int Lc;
    int pitch;
    int mode;
    int origin[3];  // shift of the window per axis: the gather starts at k / 2 + origin before each output
};

lucem_status fail(lucem_status st, const std::string& msg);
lucem_status hip_fail(int err, const char* what);
std::string dims_str(const uint64_t* dims, uint32_t ndim);
lucem_status twiddles_for(lucem_client* c, int L, const float2** out);
lucem_status ws_slot(lucem_workspace* ws, const char* name, size_t bytes, void** out);

// Device memory budget (lucem_budget.cpp). budget_check: LUCEM_OUT_OF_DEVICE_MEMORY when a limit is
// set, `bytes` is at least 64 MiB and more than the limit minus what the library holds on `device`.
// dev_malloc checks, allocates and counts; dev_free frees and uncounts; dev_forget uncounts memory
// that is left to the process (freed without a HIP call at exit).
lucem_status budget_check(int device, const char* what, uint64_t bytes);
lucem_status dev_malloc(int device, void** p, size_t bytes, const char* what);
void dev_free(int device, void* p, size_t bytes);
void dev_forget(int device, size_t bytes);

int choose_fft_len(int target);
// Radix sequence of an FFT of length L. Variant 0: greedy, largest composite radix first;
// 1: the same radices in reverse order; 2: single radices only (8, 7, 5, 4, 3, 2).
FftPlan make_plan(int L, int variant = 0);
// Shape-choice variant of a COL_FUSED pass that runs the compile-time build (plan as variant 0).
constexpr int kVariantStatic = 3;
int col_static_choice(bool tuned, int L, int B, int wide);
int row_static_choice(bool tuned, int L, int P, int wide);
int line_static_choice(bool tuned, int L, int B, int wide);

// How the descriptor builders pick launch shapes: the device size, the parallelism of the launch
// a descriptor joins (row pairs and lines of all arrays launched together; 0 = the array's own)
// and measured choices that override the heuristics.
struct ShapePolicy {
    int num_cus = 0;
    int64_t launch_pairs = 0;
    int64_t launch_lines = 0;
    const ShapeTable* table = nullptr;
};
int choose_P(int L, int rows, const ShapePolicy& sp, int* tier, int* variant);
int choose_B(int L, int Lc, int64_t outers, int kind, const ShapePolicy& sp, int* wide, int* variant);
bool make_geometry(const int64_t n[3], const int64_t k[3], int ndim, Geometry* g);

// Launch with optional profiling (events around the launch, algorithmic bytes recorded).
lucem_status run_row(lucem_client* c, const RowPass& r, int blocks, const char* name);
lucem_status run_col(lucem_client* c, const ColPass& p, int blocks, const char* name);
lucem_status run_tile(lucem_client* c, const TileIter& d, int blocks, double bytes, const char* name);
double row_pass_bytes(const RowPass& r);
double col_pass_bytes(const ColPass& p, int blocks);

// Descriptor builders shared by the plain and batched drivers.
int make_otf_descs(const Geometry& g, const ShapePolicy& sp, const float* w, float2* pa, float2* pb, float2* otf,
                   const float2* const tws[3], RowPass* row, int* row_blocks, ColPass cols[2], int col_blocks[2]);
int make_corr_descs(const Geometry& g, const ShapePolicy& sp, float2* sa, float2* sb, const float2* otf,
                    const float2* const tws[3], ColPass out[3], int blocks[3]);
RowPass make_row_desc(const Geometry& g, const ShapePolicy& sp, const float2* tw_x, float eps, const float* obs, float* est,
                      float* out, float2* spec, int* blocks);
float rl_eps(const lucem_rl_settings& s);
// One linear correlation of the row spectra in spec_a (and spec_b in 3-D) with transfer function
// `otf`, and the build of a transfer function from real weights `w` (dims g.k) through the
// scratch spectra pa, pb.
lucem_status run_corr(lucem_client* c, const Geometry& g, const ShapePolicy& sp, float2* sa, float2* sb,
                      const float2* otf, const float2* const tws[3]);
lucem_status run_otf(lucem_client* c, const Geometry& g, const float* w, float2* pa, float2* pb, float2* otf,
                     const float2* const tws[3]);

// Autotuner: times candidate shapes of these descriptors' passes (the one-array kernels, or the
// batch kernels scheduled as the batched form runs them) and records the fastest per (kind, L).
enum TuneMode : int { TUNE_PLAIN = 0, TUNE_BATCH_DIRECT = 1, TUNE_BATCH_GRID = 2 };
struct TuneCol {
    ColPass d;
    int outers;  // outer lines of the pass (work items = outers * column blocks)
};
lucem_status tune_shapes(lucem_client* c, const std::vector<RowPass>& rows, const std::vector<TuneCol>& cols,
                         int mode, ShapeTable* out);
lucem_status tuned_shapes(lucem_client* c, const Geometry& g, float* est, float2* sa, float2* sb, const float2* otf,
                          const float2* const tws[3], const ShapeTable** out);

// One axis transform of the general FFT (lucem_fft.cpp): descriptor and grid for an array
// viewed as [outer, n, inner]; the caller fills in the buffers.
lucem_status axis_fft_desc(lucem_client* c, int n, int64_t outer, int64_t inner, bool inverse, AxisFft* d,
                           int* blocks);

lucem_status validate_rl(const void* input, const uint64_t* in_dims, uint32_t in_ndim, const void* psf,
                         const uint64_t* psf_dims, uint32_t psf_ndim, const void* output, const uint64_t* out_dims,
                         uint32_t out_ndim, const lucem_rl_settings* s, uint32_t max_ndim, bool* empty, Geometry* g,
                         size_t elem_size = sizeof(float), bool fft = true);
lucem_status rl_run(lucem_client* c, lucem_workspace* ws, const float* input, const float* psf, float* output,
                    const Geometry& g, const lucem_rl_settings& s);

// The one-launch iteration of 2-D arrays (lucem_tiled.cpp, rl_tiled.hip): output tiles of T0 x T1
// on transform domains of L0 x L1, tiles0 x tiles1 of them.
struct TileGrid {
    int L0, L1, T0, T1, tiles0, tiles1;
};
// The cheapest admissible tile geometry of a 2-D array (false: none; the whole-array passes run).
bool tile_plan(const Geometry& g, int num_cus, TileGrid* out);
// Whether the plain 2-D path runs tiled (LUCEM_RL_TILED=0 / 1 force it; otherwise small arrays).
bool use_tiled(const Geometry& g, int num_cus, TileGrid* plan);
lucem_status rl_run_tiled(lucem_client* c, lucem_workspace* ws, const float* input, const float* psf, float* output,
                          const Geometry& g, const lucem_rl_settings& s, const TileGrid& tp);
// Device buffers of the tiled call (workspace slots, or the allocation-free scratch): two
// estimates (ping-pong), the transfer functions [L0][pitch], the normalised PSF and its flip, and
// the PSF row spectra of the transfer-function build. Sizes in bytes of one of each.
struct TileBufs {
    float *est_a, *est_b, *h, *rev;
    float2 *otf1, *otf2, *pa;
};
struct TileSizes {
    size_t est, otf, psf, pa;
};
TileSizes tile_sizes(const Geometry& g, const TileGrid& tp);
// The tile domain's geometry (1, L0, L1), on which the transfer functions are built.
Geometry tile_geometry(const Geometry& g, const TileGrid& tp);
// The descriptor of one array's tiled iteration, without its array pointers (obs, out, est_in,
// est_out) and with the whole array as its output window.
TileIter tile_desc(const Geometry& g, const TileGrid& tp, const lucem_rl_settings& s, const float2* otf1,
                   const float2* otf2, const float2* tw0, const float2* tw1);
// One compile-time length for a batch of 2-D arrays: the grids (same L0 = L1 for all) that admit
// every array at the least modelled time (false when no such length exists), that time and the
// modelled time of the per-pass batch of the same arrays, in us per iteration.
bool tile_batch_plan(const std::vector<Geometry>& gs, int num_cus, std::vector<TileGrid>* grids, double* tiled_us,
                     double* whole_us);
// LUCEM_RL_TILED: -1 unset, 0 off, 1 on.
int tiled_mode_env();
// The launches of one tiled call on buffers b with the twiddle tables of L0 and L1. Allocates nothing.
lucem_status rl_tiled_launch(lucem_client* c, const Geometry& g, const lucem_rl_settings& s, const float* input,
                             const float* psf, float* output, const TileBufs& b, const float2* tw0, const float2* tw1,
                             const TileGrid& tp);

// Device buffers of one plain RL call: the workspace slots of rl_run, or carved from the
// allocation-free scratch (lucem_noalloc.hip).
struct RlBuffers {
    float* est;     // current estimate, n floats
    float2* sa;     // row half spectra, rows * pitch
    float2* sb;     // 3-D: y-transformed planes, n0 * L1 * pitch
    float2* otf1;   // transfer function of the flipped PSF (A5), L0 * L1 * pitch
    float2* otf2;   // transfer function of the PSF (A7)
    float* h;       // normalised PSF
    float* rev;     // its flip
    float2* pa;     // PSF row spectra, k0 * k1 * pitch
    float2* pb;     // 3-D: PSF y-transformed planes, k0 * L1 * pitch
};
// Bytes of each RlBuffers member for geometry g (sb and pb are 0 in 2-D; otf is per transfer
// function, psf per PSF copy).
struct RlSizes {
    size_t est, sa, sb, otf, psf, pa, pb;
};
RlSizes rl_sizes(const Geometry& g);
// The launches of one plain RL call on buffers b: PSF preparation, both transfer functions, the
// transformed constant start and num_iter iterations. `table`: measured launch shapes, or null
// for the heuristics. Allocates nothing.
lucem_status rl_launch(lucem_client* c, const Geometry& g, const lucem_rl_settings& s, const float* input,
                       const float* psf, float* output, const RlBuffers& b, const float2* const tws[3],
                       const ShapeTable* table);
// The direct form's launches (2-D, f32) on bufs = {est_a, est_b, conv, rel, err} (n floats each)
// and {h, rev} (the PSF's size each). Allocates nothing.
lucem_status rl_direct_launch_f32(lucem_client* c, const float* input, const float* psf, float* output,
                                  const Geometry& g, const lucem_rl_settings& s, float* const bufs[7]);
// The launch shapes measured earlier on this client for geometry g, without measuring (null when
// none are cached or tuning is off).
const ShapeTable* tuned_lookup(lucem_client* c, const Geometry& g);
// While one of these lives on a thread, the client's profiling records nothing (no events are
// created or taken).
struct NoProfileScope {
    NoProfileScope();
    ~NoProfileScope();
    NoProfileScope(const NoProfileScope&) = delete;
    NoProfileScope& operator=(const NoProfileScope&) = delete;
};
// Resolves the allocation-free scratch of `device` (that device current); called once per device
// by lucem_client_create. The status is kept and reported by the allocation-free entry point.
void noalloc_setup(int device);

}  // namespace lucem
