// Batched Richardson-Lucy: many tiles of uneven size (slide-edge tiles, volumes of different
// depth), one plan, one hipGraph.
//
// The plan gives every tile its own FFT geometry and workspace and every (FFT shape, PSF) pair
// its own transfer functions. Each pass of the RL iteration (restore/deconvolve.rs:109-147:
// blur, ratio, adjoint, update) becomes one launch of a persistent-style kernel over the work
// items of all tiles, largest tiles first: by default one workgroup per (tile, block) item;
// LUCEM_BATCH_SCHEDULE=stride runs a grid of occupancy x CUs workgroups striding over the items,
// =queue one taking them from an atomic counter. The PSF
// preparation, the transfer-function builds and every iteration are captured once into a
// hipGraph (the counters are cleared by a kernel at its start, not by a memset node); execute
// uploads the caller's buffer pointers into the descriptors and replays the graph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <numeric>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/lucem_hip.h"
#include "lucem_internal_host.hpp"
#include "rl_internal.hpp"

using namespace lucem;

namespace {

struct TilePlan {
    Geometry g;
    uint32_t psf;
    int group;
    int64_t n;
    int row_blocks;
    int col_blocks[3];
};

struct GroupPlan {
    int L[3];
    uint32_t psf;
    int tile;  // a tile of this group (its geometry builds the transfer functions)
};

// One persistent launch of the captured sequence.
struct Launch {
    bool row;
    int op;
    int first, last, clip;
    const void* descs;
    const int2* items;
    int nitems;
    size_t lds;
    int wide;
    int grid;
};

}  // namespace

struct lucem_rl_batch {
    lucem_client* client = nullptr;
    lucem_rl_settings settings{};
    uint32_t ndim = 0;
    std::vector<TilePlan> tiles;
    std::vector<GroupPlan> groups;
    std::vector<std::vector<uint64_t>> psf_dims;
    void* mem = nullptr;
    size_t mem_bytes = 0;
    RowPass* d_rows = nullptr;          // per tile
    PsfPrep* d_psf = nullptr;           // per PSF
    std::vector<RowPass> h_rows;
    std::vector<PsfPrep> h_psf;
    RowPass* h_stage = nullptr;         // pinned upload staging
    PsfPrep* h_psf_stage = nullptr;
    hipEvent_t staged = nullptr;
    bool staged_pending = false;
    hipStream_t capture = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint64_t launches = 0;
    // Schedule of the batch launches: one workgroup per item, placed by the hardware dispatcher
    // (DIRECT, the default: measured 4-5% faster on the slide workload), or a persistent grid of
    // occupancy x CUs workgroups striding over the items (STRIDE) or taking them from an atomic
    // queue (QUEUE).
    enum Schedule { QUEUE, STRIDE, DIRECT } schedule = DIRECT;
    // Tiled mode (2-D tiles): one launch per iteration over the output tiles of every array
    // (rl_tiled.hip), one job descriptor per non-empty tile.
    bool tiled = false;
    std::vector<int> job_of;            // tile -> job (-1: empty tile)
    std::vector<TileIter> h_jobs;
    TileIter* d_jobs = nullptr;
    TileIter* h_job_stage = nullptr;
    ~lucem_rl_batch();
};

lucem_rl_batch::~lucem_rl_batch() {
    if (!lucem::exiting()) {
        if (client) {
            (void)hipSetDevice(client->device);
            (void)hipStreamSynchronize(client->stream);
        }
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (capture) (void)hipStreamDestroy(capture);
        if (staged) (void)hipEventDestroy(staged);
        if (h_stage) (void)hipHostFree(h_stage);
        if (h_psf_stage) (void)hipHostFree(h_psf_stage);
        if (h_job_stage) (void)hipHostFree(h_job_stage);
        if (mem) lucem::dev_free(client ? client->device : 0, mem, mem_bytes);
    } else if (mem) {
        lucem::dev_forget(client ? client->device : 0, mem_bytes);
    }
    lucem::client_release(client);
}

namespace {

// Carves aligned sub-buffers out of one allocation: first pass sizes, second pass assigns.
struct Carver {
    char* base = nullptr;
    size_t off = 0;
    template <class T>
    T* take(size_t count) {
        off = (off + 255) & ~size_t(255);
        T* p = base ? reinterpret_cast<T*>(base + off) : nullptr;
        off += sizeof(T) * std::max<size_t>(count, 1);
        return p;
    }
};

size_t pass_lds(bool row, const void* desc) {
    if (row) {
        const RowPass* r = static_cast<const RowPass*>(desc);
        return row_pass_lds_bytes(r->f.L, r->P);
    }
    const ColPass* c = static_cast<const ColPass*>(desc);
    return col_pass_lds_bytes(c->f.L, c->B);
}

}  // namespace

static lucem_status batch_validate(const lucem_rl_tile* tiles, uint64_t num_tiles, const uint64_t* psf_dims,
                                   uint32_t psf_ndim, uint64_t num_psfs, const lucem_rl_settings* settings,
                                   std::vector<Geometry>* geos, std::vector<bool>* empty) {
    if (!settings) return fail(LUCEM_INVALID_HANDLE, "settings is NULL");
    if (num_tiles && !tiles) return fail(LUCEM_INVALID_HANDLE, "tiles is NULL");
    if (num_psfs && !psf_dims) return fail(LUCEM_INVALID_HANDLE, "psf_dims is NULL");
    geos->resize(num_tiles);
    empty->assign(num_tiles, false);
    for (uint64_t t = 0; t < num_tiles; t++) {
        const lucem_rl_tile& tl = tiles[t];
        if (tl.psf_index >= num_psfs)
            return fail(LUCEM_INVALID_PARAMETER, "invalid parameter `psf_index`: tile " + std::to_string(t) +
                                                     " names PSF " + std::to_string(tl.psf_index) + " of " +
                                                     std::to_string(num_psfs));
        if (tl.ndim > 3) return fail(LUCEM_INVALID_DIMENSION, "tile " + std::to_string(t) + ": ndim > 3");
        const uint64_t* pd = psf_dims + (size_t)tl.psf_index * psf_ndim;
        bool e = false;
        Geometry g;
        // buffer pointers are checked at execute time
        lucem_status st = validate_rl(nullptr, tl.dims, tl.ndim, nullptr, pd, psf_ndim, nullptr, tl.dims, tl.ndim,
                                      settings, 3, &e, &g);
        if (st != LUCEM_OK) return fail(st, "tile " + std::to_string(t) + ": " + lucem_last_error());
        (*empty)[t] = e;
        (*geos)[t] = g;
    }
    return LUCEM_OK;
}

// The tiled plan: per tile its two estimates, per PSF its normalised copies and the two transfer
// functions on the common L x L domain; captured: PSF preparation, the transfer-function builds,
// then one tile launch per iteration. *done is false (and nothing is made) when the tiles do not
// all admit one compile-time length or the per-pass plan is modelled faster (batches with much
// work in all: the per-pass kernels then stream at HBM speed), and the per-pass plan is built.
static lucem_status create_tiled(lucem_rl_batch* b, const std::vector<Geometry>& geos, const std::vector<bool>& empty,
                                 const lucem_rl_settings& s, bool* done) {
    *done = false;
    if (b->ndim != 2 || tiled_mode_env() == 0) return LUCEM_OK;
    lucem_client* client = b->client;
    std::vector<Geometry> gs;
    for (size_t t = 0; t < geos.size(); t++)
        if (!empty[t]) gs.push_back(geos[t]);
    std::vector<TileGrid> grids;
    double tiled_us = 0.0, whole_us = 0.0;
    if (gs.empty() || !tile_batch_plan(gs, client->num_cus, &grids, &tiled_us, &whole_us)) return LUCEM_OK;
    // LUCEM_RL_TILED=1 takes the tiled form whenever it is admissible; otherwise when it is modelled faster
    if (tiled_mode_env() != 1 && tiled_us >= whole_us) return LUCEM_OK;
    const int L = grids[0].L0;
    const size_t ntiles = geos.size(), npsf = b->psf_dims.size();
    b->tiled = true;
    b->job_of.assign(ntiles, -1);
    int njobs = 0;
    for (size_t t = 0; t < ntiles; t++)
        if (!empty[t]) b->job_of[t] = njobs++;

    const float2* tw = nullptr;
    lucem_status st;
    if ((st = twiddles_for(client, L, &tw))) return st;
    // transfer functions per PSF on the (1, L, L) domain
    std::vector<Geometry> gpsf(npsf);
    for (size_t j = 0; j < npsf; j++) {
        Geometry gt{};
        gt.ndim = 2;
        gt.n[0] = 1;
        gt.n[1] = gt.n[2] = L;
        gt.k[0] = 1;
        gt.k[1] = (int)b->psf_dims[j][0];
        gt.k[2] = (int)b->psf_dims[j][1];
        gt.L[0] = 1;
        gt.L[1] = gt.L[2] = L;
        gt.Lc = L / 2 + 1;
        gt.pitch = (gt.Lc + 15) & ~15;
        gpsf[j] = gt;
    }
    // work items, largest tiles first
    std::vector<size_t> order;
    for (size_t t = 0; t < ntiles; t++)
        if (!empty[t]) order.push_back(t);
    std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) {
        return (int64_t)geos[x].n[1] * geos[x].n[2] > (int64_t)geos[y].n[1] * geos[y].n[2];
    });
    std::vector<int2> items;
    for (size_t t : order) {
        const TileGrid& tg = grids[b->job_of[t]];
        for (int q = 0; q < tg.tiles0 * tg.tiles1; q++) items.push_back(make_int2(b->job_of[t], q));
    }

    struct Ptrs {
        std::vector<float*> est_a, est_b, h, rev;
        std::vector<float2*> otf1, otf2, pa;
        TileIter* jobs;
        PsfPrep* psf;
        int2* items;
    } P;
    auto layout = [&](Carver& cv) {
        P.est_a.assign(njobs, nullptr);
        P.est_b.assign(njobs, nullptr);
        for (size_t t = 0; t < ntiles; t++) {
            if (b->job_of[t] < 0) continue;
            const size_t n = (size_t)geos[t].n[1] * geos[t].n[2];
            P.est_a[b->job_of[t]] = cv.take<float>(n);
            P.est_b[b->job_of[t]] = cv.take<float>(n);
        }
        P.h.assign(npsf, nullptr);
        P.rev = P.h;
        P.otf1.assign(npsf, nullptr);
        P.otf2 = P.pa = P.otf1;
        for (size_t j = 0; j < npsf; j++) {
            const Geometry& gt = gpsf[j];
            const size_t nk = (size_t)gt.k[1] * gt.k[2];
            P.h[j] = cv.take<float>(nk);
            P.rev[j] = cv.take<float>(nk);
            P.otf1[j] = cv.take<float2>((size_t)L * gt.pitch);
            P.otf2[j] = cv.take<float2>((size_t)L * gt.pitch);
            P.pa[j] = cv.take<float2>((size_t)gt.k[1] * gt.pitch);
        }
        P.jobs = cv.take<TileIter>(njobs);
        P.psf = cv.take<PsfPrep>(npsf);
        P.items = cv.take<int2>(items.size());
    };
    Carver sizing;
    layout(sizing);
    b->mem_bytes = sizing.off;
    {
        void* m = nullptr;
        if ((st = dev_malloc(client->device, &m, b->mem_bytes, "batch plan"))) return st;
        b->mem = m;
    }
    Carver real;
    real.base = static_cast<char*>(b->mem);
    layout(real);

    b->h_jobs.resize(njobs);
    for (size_t t = 0; t < ntiles; t++) {
        const int j = b->job_of[t];
        if (j < 0) continue;
        const uint32_t pi = b->tiles[t].psf;
        TileIter d = tile_desc(geos[t], grids[j], s, P.otf1[pi], P.otf2[pi], tw, tw);
        d.est_in = P.est_a[j];
        d.est_out = P.est_b[j];
        b->h_jobs[j] = d;
    }
    b->h_psf.resize(npsf);
    for (size_t j = 0; j < npsf; j++)
        b->h_psf[j] = PsfPrep{(int64_t)gpsf[j].k[1] * gpsf[j].k[2], nullptr, P.h[j], P.rev[j]};
    hipError_t e;
    if (!items.empty() && (e = upload_sync(P.items, items.data(), sizeof(int2) * items.size(), client->stream)) != hipSuccess)
        return hip_fail(e, "hipMemcpy(batch descriptors)");
    b->d_jobs = P.jobs;
    b->d_psf = P.psf;

    // capture: PSF preparation, transfer functions, iterations
    if ((e = hipStreamCreateWithFlags(&b->capture, hipStreamNonBlocking)) != hipSuccess)
        return hip_fail(e, "hipStreamCreate(capture)");
    if ((e = hipStreamBeginCapture(b->capture, hipStreamCaptureModeThreadLocal)) != hipSuccess)
        return hip_fail(e, "hipStreamBeginCapture");
    int err = npsf ? launch_psf_batch(PsfBatch{P.psf, (int)npsf}, b->capture) : 0;
    const ShapePolicy osp{client->num_cus, 0, 0, nullptr};
    const float2* tws[3] = {nullptr, tw, tw};
    uint64_t launches = npsf ? 1 : 0;
    for (size_t j = 0; j < npsf && !err; j++) {
        for (int k = 0; k < 2 && !err; k++) {
            RowPass r;
            int rb;
            ColPass cols[2];
            int cb[2];
            make_otf_descs(gpsf[j], osp, k == 0 ? P.rev[j] : P.h[j], P.pa[j], nullptr, k == 0 ? P.otf1[j] : P.otf2[j],
                           tws, &r, &rb, cols, cb);
            err = launch_row_pass(r, rb, b->capture);
            if (!err) err = launch_col_pass(cols[0], cb[0], b->capture);
            launches += 2;
        }
    }
    for (uint64_t it = 0; it < s.num_iter && !err; it++) {
        const TileBatch tb{P.jobs, P.items, (int)items.size(), it == 0 ? 1 : 0, it + 1 == s.num_iter ? 1 : 0,
                           s.clip ? 1 : 0, (int)(it & 1), L, L};
        err = launch_tile_batch(tb, b->capture);
        launches++;
    }
    hipGraph_t graph = nullptr;
    e = hipStreamEndCapture(b->capture, &graph);
    if (err) {
        if (graph) (void)hipGraphDestroy(graph);
        return hip_fail(err, "batched launch capture");
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    b->graph = graph;
    if ((e = hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0)) != hipSuccess)
        return hip_fail(e, "hipGraphInstantiate");
    b->launches = launches;
    if ((e = hipHostMalloc((void**)&b->h_job_stage, sizeof(TileIter) * std::max<int>(njobs, 1), hipHostMallocDefault)) !=
        hipSuccess)
        return hip_fail(e, "hipHostMalloc");
    if ((e = hipHostMalloc((void**)&b->h_psf_stage, sizeof(PsfPrep) * std::max<size_t>(npsf, 1), hipHostMallocDefault)) !=
        hipSuccess)
        return hip_fail(e, "hipHostMalloc");
    if ((e = hipEventCreateWithFlags(&b->staged, hipEventDisableTiming)) != hipSuccess)
        return hip_fail(e, "hipEventCreate");
    *done = true;
    return LUCEM_OK;
}

extern "C" lucem_status lucem_rl_batch_create(lucem_client* client, const lucem_rl_tile* tiles, uint64_t num_tiles,
                                              const uint64_t* psf_dims, uint32_t psf_ndim, uint64_t num_psfs,
                                              const lucem_rl_settings* settings, lucem_rl_batch** out) {
    if (!out) return fail(LUCEM_INVALID_HANDLE, "batch output pointer is NULL");
    *out = nullptr;
    std::vector<Geometry> geos;
    std::vector<bool> empty;
    lucem_status st = batch_validate(tiles, num_tiles, psf_dims, psf_ndim, num_psfs, settings, &geos, &empty);
    if (st) return st;
    if (!client) return fail(LUCEM_INVALID_HANDLE, "client is NULL");
    hipError_t e = hipSetDevice(client->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");

    auto b = std::make_unique<lucem_rl_batch>();
    b->client = client;
    client_retain(client);  // released by the plan's destructor, also when creation fails below
    if (const char* sched = std::getenv("LUCEM_BATCH_SCHEDULE")) {
        const std::string v(sched);
        b->schedule = v == "queue" ? lucem_rl_batch::QUEUE : v == "stride" ? lucem_rl_batch::STRIDE : lucem_rl_batch::DIRECT;
    }
    b->settings = *settings;
    b->ndim = psf_ndim;
    for (uint64_t j = 0; j < num_psfs; j++)
        b->psf_dims.emplace_back(psf_dims + j * psf_ndim, psf_dims + (j + 1) * psf_ndim);

    // tiles and (FFT shape, PSF) groups
    std::map<std::tuple<int, int, int, uint32_t>, int> gid;
    for (uint64_t t = 0; t < num_tiles; t++) {
        TilePlan tp{};
        tp.g = geos[t];
        tp.psf = tiles[t].psf_index;
        tp.n = empty[t] ? 0 : (int64_t)tp.g.n[0] * tp.g.n[1] * tp.g.n[2];
        tp.group = -1;
        if (!empty[t]) {
            auto key = std::make_tuple(tp.g.L[0], tp.g.L[1], tp.g.L[2], tp.psf);
            auto it = gid.find(key);
            if (it == gid.end()) {
                GroupPlan gp{};
                std::memcpy(gp.L, tp.g.L, sizeof gp.L);
                gp.psf = tp.psf;
                gp.tile = (int)t;
                it = gid.emplace(key, (int)b->groups.size()).first;
                b->groups.push_back(gp);
            }
            tp.group = it->second;
        }
        b->tiles.push_back(tp);
    }

    {
        bool done = false;
        if ((st = create_tiled(b.get(), geos, empty, *settings, &done))) return st;
        if (done) {
            *out = b.release();
            return LUCEM_OK;
        }
    }

    // twiddle tables for every length used
    std::map<int, const float2*> tw;
    for (const auto& tp : b->tiles) {
        if (tp.group < 0) continue;
        for (int a = (tp.g.ndim == 2 ? 1 : 0); a < 3; a++)
            if (!tw.count(tp.g.L[a])) {
                const float2* p;
                if ((st = twiddles_for(client, tp.g.L[a], &p))) return st;
                tw[tp.g.L[a]] = p;
            }
    }

    const size_t ntiles = b->tiles.size(), ngroups = b->groups.size(), npsf = num_psfs;
    const int ncorr = psf_ndim == 2 ? 1 : 3;
    const int notf = psf_ndim == 2 ? 1 : 2;

    // sizes: one allocation for everything the plan owns
    struct Ptrs {
        std::vector<float*> est;
        std::vector<float2*> sa, sb;
        std::vector<float2*> otf1, otf2, pa1, pa2, pb1, pb2;
        std::vector<float*> h, rev;
        RowPass* rows;
        PsfPrep* psf;
        RowPass* otf_rows;
        ColPass* otf_cols[2];
        ColPass* corr[2][3];
        int2* items_row;
        int2* items_col[3];
        int2* items_otf_row;
        int2* items_otf_col[2];
        int* counters;
    } P;
    std::vector<int2> it_row, it_col[3], it_otf_row, it_otf_col[2];
    // work items, largest tiles first
    std::vector<size_t> order(ntiles);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t bb) { return b->tiles[a].n > b->tiles[bb].n; });

    // Launch shapes: every tile joins one launch per pass, so the parallelism of a launch is the
    // whole batch's; the autotuner may replace the heuristic choices once the buffers exist.
    // Work-item arrays are sized for the smallest shapes (one line per workgroup) so that any
    // choice fits.
    ShapeTable table;
    ShapePolicy bsp{client->num_cus, 0, 0, nullptr};
    const ShapePolicy osp{client->num_cus, 0, 0, nullptr};
    size_t cap_row = 0, cap_col[3] = {0, 0, 0};
    for (const TilePlan& tp : b->tiles) {
        if (tp.group < 0) continue;
        const Geometry& g = tp.g;
        const int64_t pairs = ((int64_t)g.n[0] * g.n[1] + 1) / 2;
        bsp.launch_pairs += pairs;
        bsp.launch_lines += (int64_t)g.Lc * g.n[0];
        cap_row += (size_t)pairs;
        if (ncorr == 1) {
            cap_col[0] += (size_t)g.Lc;
        } else {
            cap_col[0] += (size_t)g.Lc * g.n[0];
            cap_col[1] += (size_t)g.Lc * g.L[1];
            cap_col[2] += (size_t)g.Lc * g.n[0];
        }
    }

    // descriptors need the final pointers, so size first with a null carver
    auto layout = [&](Carver& cv) {
        P.est.assign(ntiles, nullptr);
        P.sa.assign(ntiles, nullptr);
        P.sb.assign(ntiles, nullptr);
        for (size_t t = 0; t < ntiles; t++) {
            const TilePlan& tp = b->tiles[t];
            if (tp.group < 0) continue;
            const Geometry& g = tp.g;
            P.est[t] = cv.take<float>(tp.n);
            P.sa[t] = cv.take<float2>((size_t)g.n[0] * g.n[1] * g.pitch);
            if (g.ndim == 3) P.sb[t] = cv.take<float2>((size_t)g.n[0] * g.L[1] * g.pitch);
        }
        P.otf1.assign(ngroups, nullptr);
        P.otf2 = P.pa1 = P.pa2 = P.pb1 = P.pb2 = P.otf1;
        for (size_t gi = 0; gi < ngroups; gi++) {
            const Geometry& g = b->tiles[b->groups[gi].tile].g;
            const size_t otf = (size_t)g.L[0] * g.L[1] * g.pitch;
            P.otf1[gi] = cv.take<float2>(otf);
            P.otf2[gi] = cv.take<float2>(otf);
            P.pa1[gi] = cv.take<float2>((size_t)g.k[0] * g.k[1] * g.pitch);
            P.pa2[gi] = cv.take<float2>((size_t)g.k[0] * g.k[1] * g.pitch);
            if (g.ndim == 3) {
                P.pb1[gi] = cv.take<float2>((size_t)g.k[0] * g.L[1] * g.pitch);
                P.pb2[gi] = cv.take<float2>((size_t)g.k[0] * g.L[1] * g.pitch);
            }
        }
        P.h.assign(npsf, nullptr);
        P.rev.assign(npsf, nullptr);
        for (size_t j = 0; j < npsf; j++) {
            size_t nk = 1;
            for (auto d : b->psf_dims[j]) nk *= d;
            P.h[j] = cv.take<float>(nk);
            P.rev[j] = cv.take<float>(nk);
        }
        P.rows = cv.take<RowPass>(ntiles);
        P.psf = cv.take<PsfPrep>(npsf);
        P.otf_rows = cv.take<RowPass>(2 * ngroups);
        for (int i = 0; i < notf; i++) P.otf_cols[i] = cv.take<ColPass>(2 * ngroups);
        for (int k = 0; k < 2; k++)
            for (int i = 0; i < ncorr; i++) P.corr[k][i] = cv.take<ColPass>(ntiles);
        P.items_row = cv.take<int2>(cap_row);
        for (int i = 0; i < ncorr; i++) P.items_col[i] = cv.take<int2>(cap_col[i]);
        P.items_otf_row = cv.take<int2>(it_otf_row.size());
        for (int i = 0; i < notf; i++) P.items_otf_col[i] = cv.take<int2>(it_otf_col[i].size());
        P.counters = cv.take<int>(4096);
    };

    // host descriptors (built with the final pointers after the allocation)
    std::vector<RowPass> rows(ntiles), otf_rows(2 * ngroups);
    std::vector<ColPass> otf_cols[2], corr[2][3];
    for (int i = 0; i < 2; i++) otf_cols[i].resize(2 * ngroups);
    for (int k = 0; k < 2; k++)
        for (int i = 0; i < 3; i++) corr[k][i].resize(ntiles);
    std::vector<PsfPrep> psfp(npsf);
    const float eps = rl_eps(*settings);


    auto build_descs = [&]() {
        for (size_t t = 0; t < ntiles; t++) {
            TilePlan& tp = b->tiles[t];
            if (tp.group < 0) continue;
            const Geometry& g = tp.g;
            const float2* tws[3] = {g.ndim == 3 ? tw[g.L[0]] : nullptr, tw[g.L[1]], tw[g.L[2]]};
            rows[t] = make_row_desc(g, bsp, tws[2], eps, nullptr, P.est[t], nullptr, P.sa[t], &tp.row_blocks);
            ColPass d[3];
            int blk[3];
            make_corr_descs(g, bsp, P.sa[t], P.sb[t], P.otf1[tp.group], tws, d, blk);
            for (int i = 0; i < ncorr; i++) {
                corr[0][i][t] = d[i];
                tp.col_blocks[i] = blk[i];
            }
            make_corr_descs(g, bsp, P.sa[t], P.sb[t], P.otf2[tp.group], tws, d, blk);
            for (int i = 0; i < ncorr; i++) corr[1][i][t] = d[i];
        }
        for (size_t gi = 0; gi < ngroups; gi++) {
            const GroupPlan& gp = b->groups[gi];
            const Geometry& g = b->tiles[gp.tile].g;
            const float2* tws[3] = {g.ndim == 3 ? tw[g.L[0]] : nullptr, tw[g.L[1]], tw[g.L[2]]};
            for (int k = 0; k < 2; k++) {
                RowPass r;
                int rb;
                ColPass cols[2];
                int cb[2];
                make_otf_descs(g, osp, k == 0 ? P.rev[gp.psf] : P.h[gp.psf], k == 0 ? P.pa1[gi] : P.pa2[gi],
                               k == 0 ? P.pb1[gi] : P.pb2[gi], k == 0 ? P.otf1[gi] : P.otf2[gi], tws, &r, &rb, cols, cb);
                otf_rows[2 * gi + k] = r;
                for (int i = 0; i < notf; i++) otf_cols[i][2 * gi + k] = cols[i];
            }
        }
        for (size_t j = 0; j < npsf; j++) {
            int64_t nk = 1;
            for (auto d : b->psf_dims[j]) nk *= (int64_t)d;
            psfp[j] = PsfPrep{nk, nullptr, P.h[j], P.rev[j]};
        }
    };

    // work items of the current descriptors, largest tiles first
    auto build_items = [&]() {
        it_row.clear();
        for (int i = 0; i < 3; i++) it_col[i].clear();
        it_otf_row.clear();
        for (int i = 0; i < 2; i++) it_otf_col[i].clear();
        for (size_t oi : order) {
            const TilePlan& tp = b->tiles[oi];
            if (tp.group < 0) continue;
            for (int k = 0; k < tp.row_blocks; k++) it_row.push_back(make_int2((int)oi, k));
            for (int i = 0; i < ncorr; i++)
                for (int k = 0; k < tp.col_blocks[i]; k++) it_col[i].push_back(make_int2((int)oi, k));
        }
        for (size_t gi = 0; gi < ngroups; gi++)
            for (int k = 0; k < 2; k++) {
                const RowPass& r = otf_rows[2 * gi + k];
                const int rb = (((r.rows + 1) / 2) + r.P - 1) / r.P;
                for (int q = 0; q < rb; q++) it_otf_row.push_back(make_int2((int)(2 * gi + k), q));
                for (int i = 0; i < notf; i++) {
                    const ColPass& cp = otf_cols[i][2 * gi + k];
                    const Geometry& g = b->tiles[b->groups[gi].tile].g;
                    const int outers = i == 0 ? (notf == 1 ? 1 : g.k[0]) : g.L[1];
                    for (int q = 0; q < outers * cp.nkb; q++) it_otf_col[i].push_back(make_int2((int)(2 * gi + k), q));
                }
            }
    };
    // the transfer-function items do not depend on tuning: size them with a sizing pass
    {
        Carver sizing;
        layout(sizing);
        build_descs();
        build_items();
    }
    Carver sizing;
    layout(sizing);
    b->mem_bytes = sizing.off;
    {
        void* m = nullptr;
        if (lucem_status st = dev_malloc(client->device, &m, b->mem_bytes, "batch plan")) return st;
        b->mem = m;
    }
    Carver real;
    real.base = static_cast<char*>(b->mem);
    layout(real);
    build_descs();
    // autotune on the plan's own buffers, then rebuild the descriptors with the measured shapes
    {
        std::vector<RowPass> trows;
        std::vector<TuneCol> tcols;
        for (size_t t = 0; t < ntiles; t++) {
            const TilePlan& tp = b->tiles[t];
            if (tp.group < 0) continue;
            trows.push_back(rows[t]);
            for (int i = 0; i < ncorr; i++)
                tcols.push_back(TuneCol{corr[0][i][t], tp.col_blocks[i] / std::max(corr[0][i][t].nkb, 1)});
        }
        if ((st = tune_shapes(client, trows, tcols, b->schedule == lucem_rl_batch::DIRECT ? TUNE_BATCH_DIRECT : TUNE_BATCH_GRID,
                              &table)))
            return st;
        bsp.table = &table;
        build_descs();
    }
    build_items();
    b->h_rows = rows;
    b->h_psf = psfp;

    // upload the plan-owned descriptors (the pointer fields of rows/psf are filled per execute)
    hipStream_t s = client->stream;
    auto up = [&](void* dst, const void* src, size_t bytes) -> lucem_status {
        if (!bytes) return LUCEM_OK;
        hipError_t err = upload_sync(dst, src, bytes, s);
        return err == hipSuccess ? LUCEM_OK : hip_fail(err, "hipMemcpy(batch descriptors)");
    };
    if ((st = up(P.otf_rows, otf_rows.data(), sizeof(RowPass) * otf_rows.size()))) return st;
    for (int i = 0; i < notf; i++)
        if ((st = up(P.otf_cols[i], otf_cols[i].data(), sizeof(ColPass) * otf_cols[i].size()))) return st;
    for (int k = 0; k < 2; k++)
        for (int i = 0; i < ncorr; i++)
            if ((st = up(P.corr[k][i], corr[k][i].data(), sizeof(ColPass) * ntiles))) return st;
    if ((st = up(P.items_row, it_row.data(), sizeof(int2) * it_row.size()))) return st;
    for (int i = 0; i < ncorr; i++)
        if ((st = up(P.items_col[i], it_col[i].data(), sizeof(int2) * it_col[i].size()))) return st;
    if ((st = up(P.items_otf_row, it_otf_row.data(), sizeof(int2) * it_otf_row.size()))) return st;
    for (int i = 0; i < notf; i++)
        if ((st = up(P.items_otf_col[i], it_otf_col[i].data(), sizeof(int2) * it_otf_col[i].size()))) return st;
    b->d_rows = P.rows;
    b->d_psf = P.psf;

    // the launch sequence
    std::vector<Launch> seq;
    auto add = [&](bool row, int op, const void* descs, const int2* items, int nitems, size_t lds, int wide, int f = 0,
                   int l = 0, int cl = 0) {
        if (nitems <= 0) return;
        Launch L{};
        L.row = row;
        L.op = op;
        L.first = f;
        L.last = l;
        L.clip = cl;
        L.descs = descs;
        L.items = items;
        L.nitems = nitems;
        L.lds = lds;
        L.wide = wide;
        const int occ = batch_occupancy(row, wide, lds);
        L.grid = b->schedule != lucem_rl_batch::DIRECT ? std::max(1, std::min(nitems, occ * client->num_cus)) : nitems;
        seq.push_back(L);
    };
    auto max_lds = [&](bool row, const auto& descs, const std::vector<int2>& items) {
        size_t m = 0;
        for (const auto& it : items) m = std::max(m, pass_lds(row, &descs[it.x]));
        return m;
    };
    // one line-pass launch runs one kernel shape: the 512-thread one if any tile asks for it
    // (a tile planned for 256 threads runs correctly on 512: its items use fewer columns)
    auto any_wide = [&](const std::vector<ColPass>& descs, const std::vector<int2>& items) {
        int w = 0;
        for (const auto& it : items) w = std::max(w, descs[it.x].wide);
        return w;
    };
    auto row_tier = [&](const std::vector<RowPass>& descs, const std::vector<int2>& items) {
        int w = 0;
        for (const auto& it : items) w = std::max(w, descs[it.x].wide);
        return w;
    };
    add(true, ROW_LOAD, P.otf_rows, P.items_otf_row, (int)it_otf_row.size(), max_lds(true, otf_rows, it_otf_row),
        row_tier(otf_rows, it_otf_row));
    for (int i = 0; i < notf; i++)
        add(false, i + 1 == notf ? COL_FWD_OTF : COL_FWD, P.otf_cols[i], P.items_otf_col[i],
            (int)it_otf_col[i].size(), max_lds(false, otf_cols[i], it_otf_col[i]), any_wide(otf_cols[i], it_otf_col[i]));
    const size_t lds_row = max_lds(true, rows, it_row);
    const int tier_row = row_tier(rows, it_row);
    size_t lds_col[3];
    int wide_col[3];
    for (int i = 0; i < ncorr; i++) {
        lds_col[i] = max_lds(false, corr[0][i], it_col[i]);
        wide_col[i] = any_wide(corr[0][i], it_col[i]);
    }
    add(true, ROW_LOAD_CONST, P.rows, P.items_row, (int)it_row.size(), lds_row, tier_row);
    const int corr_ops2[1] = {COL_FUSED};
    const int corr_ops3[3] = {COL_FWD, COL_FUSED, COL_INV};
    for (uint64_t it = 0; it < settings->num_iter; it++) {
        const bool last = it + 1 == settings->num_iter;
        for (int k = 0; k < 2; k++) {
            for (int i = 0; i < ncorr; i++)
                add(false, ncorr == 1 ? corr_ops2[i] : corr_ops3[i], P.corr[k][i], P.items_col[i],
                    (int)it_col[i].size(), lds_col[i], wide_col[i]);
            if (k == 0)
                add(true, ROW_RATIO, P.rows, P.items_row, (int)it_row.size(), lds_row, tier_row);
            else
                add(true, ROW_UPDATE, P.rows, P.items_row, (int)it_row.size(), lds_row, tier_row, it == 0, last,
                    settings->clip ? 1 : 0);
        }
    }
    if (seq.size() + 1 > 4096) return fail(LUCEM_INVALID_PARAMETER, "invalid parameter `num_iter`: too many launches");

    // capture: counters cleared, PSF prep, transfer functions, iterations
    if ((e = hipStreamCreateWithFlags(&b->capture, hipStreamNonBlocking)) != hipSuccess)
        return hip_fail(e, "hipStreamCreate(capture)");
    if ((e = hipStreamBeginCapture(b->capture, hipStreamCaptureModeThreadLocal)) != hipSuccess)
        return hip_fail(e, "hipStreamBeginCapture");
    int err = launch_zero_ints(P.counters, (int)seq.size(), b->capture);
    if (!err && npsf) err = launch_psf_batch(PsfBatch{P.psf, (int)npsf}, b->capture);
    for (size_t i = 0; i < seq.size() && !err; i++) {
        const Launch& L = seq[i];
        if (L.row) {
            RowBatch rb{static_cast<const RowPass*>(L.descs), L.items, L.nitems, L.op, L.first, L.last, L.clip, L.wide,
                        b->schedule == lucem_rl_batch::QUEUE ? P.counters + i : nullptr,
                        b->schedule == lucem_rl_batch::STRIDE ? 1 : 0};
            err = launch_row_batch(rb, L.grid, L.lds, b->capture);
        } else {
            ColBatch cb{static_cast<const ColPass*>(L.descs), L.items, L.nitems, L.op, L.wide,
                        b->schedule == lucem_rl_batch::QUEUE ? P.counters + i : nullptr, 0,
                        b->schedule == lucem_rl_batch::STRIDE ? 1 : 0};
            err = launch_col_batch(cb, L.grid, L.lds, b->capture);
        }
    }
    hipGraph_t graph = nullptr;
    e = hipStreamEndCapture(b->capture, &graph);
    if (err) {
        if (graph) (void)hipGraphDestroy(graph);
        return hip_fail(err, "batched launch capture");
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    b->graph = graph;
    if ((e = hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0)) != hipSuccess)
        return hip_fail(e, "hipGraphInstantiate");
    b->launches = seq.size() + (npsf ? 2 : 1);

    if ((e = hipHostMalloc((void**)&b->h_stage, sizeof(RowPass) * std::max<size_t>(ntiles, 1), hipHostMallocDefault)) !=
        hipSuccess)
        return hip_fail(e, "hipHostMalloc");
    if ((e = hipHostMalloc((void**)&b->h_psf_stage, sizeof(PsfPrep) * std::max<size_t>(npsf, 1), hipHostMallocDefault)) !=
        hipSuccess)
        return hip_fail(e, "hipHostMalloc");
    if ((e = hipEventCreateWithFlags(&b->staged, hipEventDisableTiming)) != hipSuccess)
        return hip_fail(e, "hipEventCreate");
    *out = b.release();
    return LUCEM_OK;
}

namespace {

// Byte range [lo, hi) a strided view of (nz, ny, nx) elements spans.
struct Span {
    uintptr_t lo, hi;
    size_t tile;
};

Span view_span(const void* p, const int64_t n3[3], int64_t rs, int64_t ps, size_t t) {
    const uintptr_t lo = (uintptr_t)p;
    const uint64_t last = (uint64_t)(n3[0] - 1) * (uint64_t)ps + (uint64_t)(n3[1] - 1) * (uint64_t)rs + (uint64_t)n3[2];
    return Span{lo, lo + last * sizeof(float), t};
}

// The layout of tile t applied to its descriptor: strides default to the tile's own, the window
// to the whole tile; the window must lie inside the tile.
lucem_status apply_layout(const lucem_rl_tile_layout* lay, const TilePlan& tp, size_t t, RowPass& r, Span* in,
                          Span* out, const void* input, const void* output) {
    const int64_t n3[3] = {tp.g.n[0], tp.g.n[1], tp.g.n[2]};
    r.ny = tp.g.n[1];
    r.strided_obs = 0;
    r.window = 0;
    r.obs_rs = r.out_rs = n3[2];
    r.obs_ps = r.out_ps = n3[1] * n3[2];
    for (int a = 0; a < 3; a++) {
        r.own0[a] = 0;
        r.own1[a] = (int)n3[a];
    }
    int64_t win[3] = {n3[0], n3[1], n3[2]};
    if (lay) {
        const int off = 3 - tp.g.ndim;  // 2-D tiles: the layout's z entries are ignored
        if (lay->in_row_stride) r.obs_rs = (int64_t)lay->in_row_stride;
        if (lay->in_plane_stride) r.obs_ps = (int64_t)lay->in_plane_stride;
        if (lay->out_row_stride) r.out_rs = (int64_t)lay->out_row_stride;
        if (lay->out_plane_stride) r.out_ps = (int64_t)lay->out_plane_stride;
        for (int a = off; a < 3; a++) {
            const uint64_t o = lay->own_origin[a];
            const uint64_t sz = lay->own_size[a] ? lay->own_size[a] : (uint64_t)n3[a] - std::min<uint64_t>(o, n3[a]);
            if (o >= (uint64_t)n3[a] || sz == 0 || o + sz > (uint64_t)n3[a])
                return fail(LUCEM_INVALID_PARAMETER, "tile " + std::to_string(t) +
                                                         ": the stored window lies outside the tile");
            r.own0[a] = (int)o;
            r.own1[a] = (int)(o + sz);
            win[a] = (int64_t)sz;
        }
        if (r.obs_rs < n3[2] || (tp.g.ndim == 3 && r.obs_ps < n3[1] * r.obs_rs) || r.out_rs < win[2] ||
            (tp.g.ndim == 3 && r.out_ps < win[1] * r.out_rs))
            return fail(LUCEM_INVALID_PARAMETER, "tile " + std::to_string(t) + ": a stride is shorter than the rows it spans");
        r.strided_obs = r.obs_rs != n3[2] || r.obs_ps != n3[1] * n3[2];
        r.window = 1;
    }
    *in = view_span(input, n3, r.obs_rs, r.obs_ps, t);
    *out = view_span(output, win, r.out_rs, r.out_ps, t);
    return LUCEM_OK;
}

lucem_status batch_execute(lucem_rl_batch* b, const lucem_rl_tile* tiles, const lucem_rl_tile_layout* layouts,
                           const float* const* psfs) {
    if (!b) return fail(LUCEM_INVALID_HANDLE, "batch is NULL");
    if (b->tiles.size() && !tiles) return fail(LUCEM_INVALID_HANDLE, "tiles is NULL");
    if (b->psf_dims.size() && !psfs) return fail(LUCEM_INVALID_HANDLE, "psfs is NULL");
    lucem_client* c = b->client;
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    // the previous upload must have left the staging buffers before they are rewritten
    if (b->staged_pending) {
        if ((e = hipEventSynchronize(b->staged)) != hipSuccess) return hip_fail(e, "hipEventSynchronize");
        b->staged_pending = false;
    }
    std::vector<Span> ins, outs;
    ins.reserve(b->tiles.size());
    outs.reserve(b->tiles.size());
    for (size_t t = 0; t < b->tiles.size(); t++) {
        const lucem_rl_tile& tl = tiles[t];
        const TilePlan& tp = b->tiles[t];
        int64_t n3[3] = {1, 1, 1};
        if (tl.ndim > 3) return fail(LUCEM_SHAPE_MISMATCH, "tile " + std::to_string(t) + " differs from the plan");
        for (uint32_t i = 0; i < tl.ndim; i++) n3[3 - tl.ndim + i] = (int64_t)tl.dims[i];
        const bool same = tp.group < 0 ? (n3[0] * n3[1] * n3[2] == 0)
                                       : (n3[0] == tp.g.n[0] && n3[1] == tp.g.n[1] && n3[2] == tp.g.n[2] &&
                                          tl.psf_index == tp.psf);
        if (!same) return fail(LUCEM_SHAPE_MISMATCH, "tile " + std::to_string(t) + " differs from the plan");
        if (tp.group < 0) continue;
        if (!tl.input || !tl.output) return fail(LUCEM_INVALID_HANDLE, "tile pointer is NULL");
        Span in, out;
        RowPass view{};
        RowPass& r = b->tiled ? view : b->h_rows[t];
        lucem_status st = apply_layout(layouts ? &layouts[t] : nullptr, tp, t, r, &in, &out, tl.input, tl.output);
        if (st) return st;
        ins.push_back(in);
        outs.push_back(out);
        if (b->tiled) {
            TileIter& d = b->h_jobs[b->job_of[t]];
            d.obs = tl.input;
            d.out = tl.output;
            d.obs_rs = r.obs_rs;
            d.out_rs = r.out_rs;
            d.own0y = r.own0[1];
            d.own0x = r.own0[2];
            d.own1y = r.own1[1];
            d.own1x = r.own1[2];
        } else {
            r.obs = tl.input;
            r.out = tl.output;
        }
    }
    // Every tile reads its input in every iteration while the others store their results, so no
    // stored range may overlap any input range (strided views are compared by their extents).
    std::sort(ins.begin(), ins.end(), [](const Span& x, const Span& y) { return x.lo < y.lo; });
    std::vector<uintptr_t> reach(ins.size());
    for (size_t i = 0; i < ins.size(); i++) reach[i] = std::max(ins[i].hi, i ? reach[i - 1] : 0);
    for (const Span& o : outs) {
        // inputs starting before o.hi: the last of them reaching past o.lo overlaps
        const size_t k = std::lower_bound(ins.begin(), ins.end(), o.hi,
                                          [](const Span& x, uintptr_t v) { return x.lo < v; }) -
                         ins.begin();
        if (k > 0 && reach[k - 1] > o.lo) {
            if (!layouts)
                return fail(LUCEM_ALIASED_BUFFERS, "tile " + std::to_string(o.tile) +
                                                       ": input and output arrays must not share the same device buffer");
            return fail(LUCEM_ALIASED_BUFFERS, "tile " + std::to_string(o.tile) +
                                                   ": its stored window overlaps an input of the batch");
        }
    }
    for (size_t j = 0; j < b->h_psf.size(); j++) {
        if (!psfs[j]) return fail(LUCEM_INVALID_HANDLE, "psf pointer is NULL");
        b->h_psf[j].psf = psfs[j];
    }
    std::memcpy(b->h_psf_stage, b->h_psf.data(), sizeof(PsfPrep) * b->h_psf.size());
    if (b->tiled) {
        std::memcpy(b->h_job_stage, b->h_jobs.data(), sizeof(TileIter) * b->h_jobs.size());
        if (!b->h_jobs.empty() &&
            (e = hipMemcpyAsync(b->d_jobs, b->h_job_stage, sizeof(TileIter) * b->h_jobs.size(), hipMemcpyHostToDevice,
                                c->stream)) != hipSuccess)
            return hip_fail(e, "hipMemcpyAsync(descriptors)");
    } else {
        std::memcpy(b->h_stage, b->h_rows.data(), sizeof(RowPass) * b->h_rows.size());
        if (!b->h_rows.empty() &&
            (e = hipMemcpyAsync(b->d_rows, b->h_stage, sizeof(RowPass) * b->h_rows.size(), hipMemcpyHostToDevice,
                                c->stream)) != hipSuccess)
            return hip_fail(e, "hipMemcpyAsync(descriptors)");
    }
    if (!b->h_psf.empty() &&
        (e = hipMemcpyAsync(b->d_psf, b->h_psf_stage, sizeof(PsfPrep) * b->h_psf.size(), hipMemcpyHostToDevice,
                            c->stream)) != hipSuccess)
        return hip_fail(e, "hipMemcpyAsync(descriptors)");
    if ((e = hipEventRecord(b->staged, c->stream)) != hipSuccess) return hip_fail(e, "hipEventRecord");
    b->staged_pending = true;
    if ((e = hipGraphLaunch(b->exec, c->stream)) != hipSuccess) return hip_fail(e, "hipGraphLaunch");
    return LUCEM_OK;
}

}  // namespace

extern "C" lucem_status lucem_rl_batch_execute(lucem_rl_batch* b, const lucem_rl_tile* tiles, const float* const* psfs) {
    return batch_execute(b, tiles, nullptr, psfs);
}

extern "C" lucem_status lucem_rl_batch_execute_layout(lucem_rl_batch* b, const lucem_rl_tile* tiles,
                                                      const lucem_rl_tile_layout* layouts, const float* const* psfs) {
    if (b && b->tiles.size() && !layouts) return fail(LUCEM_INVALID_HANDLE, "layouts is NULL");
    return batch_execute(b, tiles, layouts, psfs);
}

extern "C" lucem_status lucem_rl_batch_destroy(lucem_rl_batch* b) {
    delete b;
    return LUCEM_OK;
}

extern "C" uint64_t lucem_rl_batch_device_bytes(const lucem_rl_batch* b) { return b ? b->mem_bytes : 0; }
extern "C" uint64_t lucem_rl_batch_launches(const lucem_rl_batch* b) { return b ? b->launches : 0; }

extern "C" lucem_status lucem_richardson_lucy_batched_f32(lucem_client* client, const lucem_rl_tile* tiles,
                                                          uint64_t num_tiles, const float* const* psfs,
                                                          const uint64_t* psf_dims, uint32_t psf_ndim,
                                                          uint64_t num_psfs, const lucem_rl_settings* settings) {
    lucem_rl_batch* b = nullptr;
    lucem_status st = lucem_rl_batch_create(client, tiles, num_tiles, psf_dims, psf_ndim, num_psfs, settings, &b);
    if (st) return st;
    st = lucem_rl_batch_execute(b, tiles, psfs);
    lucem_rl_batch_destroy(b);  // waits for the stream before releasing the plan's memory
    return st;
}
