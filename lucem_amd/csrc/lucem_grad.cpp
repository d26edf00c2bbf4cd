// Differentiable Richardson-Lucy: forward with checkpoints and a hand-written adjoint, for many
// tiles (2-D, or 3-D volumes of uneven depth) sharing one PSF.
//
// Forward: T RL iterations from the caller's x0 (the reference iteration of
// restore/deconvolve.rs:109-147, no clip), then u_T = A x_T and loss = sum(u_T - y log u_T).
// x is checkpointed every k iterations; the last segment keeps everything the backward needs
// (x_t, u_t, c_t and the spectra of the padded x_t and r_t). Backward: segments from last to
// first, recomputing each earlier segment from its checkpoint, then per iteration two
// transposed correlations with the pointwise adjoint steps fused into the row passes. The PSF
// gradient accumulates in the spectral domain and is transformed back once at the end, then
// taken through the normalisation h = psf / sum(psf) (deconvolve.rs:77-79).
//
// Both directions are captured once into hipGraphs; calls upload the caller's pointers into the
// tile descriptors and replay them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <numeric>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/lucem_hip.h"
#include "lucem_internal_host.hpp"
#include "rl_grad.hpp"
#include "rl_internal.hpp"

using namespace lucem;

namespace {

struct Carve {
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

// A batched launch of the captured sequences.
struct GLaunch {
    enum Kind { ROW, COL, GROW, PSF, ZERO, LOSS, PSFGRAD } kind;
    int op = 0;
    const void* descs = nullptr;
    const int2* items = nullptr;
    int nitems = 0;
    size_t lds = 0;
    int wide = 0;
    int slot = 0;
    GradRowBatch g{};
};

}  // namespace

struct lucem_rl_grad {
    lucem_client* client = nullptr;
    lucem_rl_settings settings{};
    int ndim = 0, k = 0, T = 0, S = 0;
    std::vector<Geometry> geos;
    std::vector<uint64_t> psf_dims;
    int64_t nk = 0;
    void* mem = nullptr;
    size_t mem_bytes = 0;
    // per-call descriptor uploads
    GradTile* d_tiles = nullptr;
    std::vector<GradTile> h_tiles;
    GradTile* h_stage = nullptr;
    hipEvent_t staged = nullptr;
    bool staged_pending = false;
    float* psf_copy = nullptr;    // the PSF of the last forward
    double* loss_buf = nullptr;   // per tile
    float* gpsf_buf = nullptr;    // d loss / d psf
    hipStream_t capture = nullptr;
    hipGraph_t g_fwd = nullptr, g_bwd = nullptr;
    hipGraphExec_t x_fwd = nullptr, x_bwd = nullptr;
    bool forward_done = false;
    uint64_t launches_fwd = 0, launches_bwd = 0;
    ~lucem_rl_grad();
};

lucem_rl_grad::~lucem_rl_grad() {
    if (!lucem::exiting()) {
        if (client) {
            (void)hipSetDevice(client->device);
            (void)hipStreamSynchronize(client->stream);
        }
        if (x_fwd) (void)hipGraphExecDestroy(x_fwd);
        if (x_bwd) (void)hipGraphExecDestroy(x_bwd);
        if (g_fwd) (void)hipGraphDestroy(g_fwd);
        if (g_bwd) (void)hipGraphDestroy(g_bwd);
        if (capture) (void)hipStreamDestroy(capture);
        if (staged) (void)hipEventDestroy(staged);
        if (h_stage) (void)hipHostFree(h_stage);
        if (mem) lucem::dev_free(client ? client->device : 0, mem, mem_bytes);
    } else if (mem) {
        lucem::dev_forget(client ? client->device : 0, mem_bytes);
    }
    lucem::client_release(client);
}

static lucem_status run_seq(hipStream_t s, const std::vector<GLaunch>& seq, int* counters_unused) {
    (void)counters_unused;
    for (const GLaunch& L : seq) {
        int err = 0;
        switch (L.kind) {
            case GLaunch::ROW: {
                RowBatch rb{static_cast<const RowPass*>(L.descs), L.items, L.nitems, L.op, 0, 0, 0, L.wide, nullptr};
                err = launch_row_batch(rb, L.nitems, L.lds, s);
                break;
            }
            case GLaunch::COL: {
                ColBatch cb{static_cast<const ColPass*>(L.descs), L.items, L.nitems, L.op, L.wide, nullptr, L.slot};
                err = launch_col_batch(cb, L.nitems, L.lds, s);
                break;
            }
            case GLaunch::GROW: err = launch_grad_row(L.g, L.lds, s); break;
            default: break;
        }
        if (err) return hip_fail(err, "gradient pass launch");
    }
    return LUCEM_OK;
}

extern "C" lucem_status lucem_rl_grad_create(lucem_client* client, const uint64_t* tile_dims, uint32_t ndim,
                                             uint64_t num_tiles, const uint64_t* psf_dims,
                                             const lucem_rl_settings* settings, uint32_t checkpoint_every,
                                             lucem_rl_grad** out) {
    if (!out) return fail(LUCEM_INVALID_HANDLE, "output pointer is NULL");
    *out = nullptr;
    if (!settings) return fail(LUCEM_INVALID_HANDLE, "settings is NULL");
    if (num_tiles == 0) return fail(LUCEM_INVALID_PARAMETER, "invalid parameter `num_tiles`: at least one tile");
    if (!tile_dims || !psf_dims) return fail(LUCEM_INVALID_HANDLE, "dims pointer is NULL");
    if (settings->clip) return fail(LUCEM_INVALID_PARAMETER, "invalid parameter `clip`: the differentiable form has no clip");
    if (settings->num_iter > 4096) return fail(LUCEM_INVALID_PARAMETER, "invalid parameter `num_iter`: at most 4096");
    auto P = std::make_unique<lucem_rl_grad>();
    P->geos.resize(num_tiles);
    for (uint64_t t = 0; t < num_tiles; t++) {
        bool empty = false;
        lucem_status st = validate_rl(nullptr, tile_dims + t * ndim, ndim, nullptr, psf_dims, ndim, nullptr,
                                      tile_dims + t * ndim, ndim, settings, 3, &empty, &P->geos[t]);
        if (st) return fail(st, "tile " + std::to_string(t) + ": " + lucem_last_error());
        if (empty) return fail(LUCEM_INVALID_PARAMETER, "tile " + std::to_string(t) + " is empty");
    }
    if (!client) return fail(LUCEM_INVALID_HANDLE, "client is NULL");
    hipError_t e = hipSetDevice(client->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    P->client = client;
    client_retain(client);  // released by the plan's destructor, also when creation fails below
    P->settings = *settings;
    P->ndim = (int)ndim;
    P->T = (int)settings->num_iter;
    P->k = (int)std::max<uint32_t>(1, std::min<uint32_t>(checkpoint_every ? checkpoint_every : P->T, P->T));
    P->S = (P->T + P->k - 1) / P->k;
    P->psf_dims.assign(psf_dims, psf_dims + ndim);
    P->nk = 1;
    for (auto d : P->psf_dims) P->nk *= (int64_t)d;
    const int ntiles = (int)num_tiles, K = P->k, T = P->T, S = P->S;
    const bool three = ndim == 3;
    const int ncorr = three ? 3 : 1;

    // groups of tiles sharing an FFT shape (transfer functions)
    std::map<std::tuple<int, int, int>, int> gid;
    std::vector<int> group_of(ntiles), group_tile;
    for (int t = 0; t < ntiles; t++) {
        const Geometry& g = P->geos[t];
        auto key = std::make_tuple(g.L[0], g.L[1], g.L[2]);
        auto it = gid.find(key);
        if (it == gid.end()) {
            it = gid.emplace(key, (int)group_tile.size()).first;
            group_tile.push_back(t);
        }
        group_of[t] = it->second;
    }
    const int ngroups = (int)group_tile.size();
    std::map<int, const float2*> tw;
    for (int t = 0; t < ntiles; t++)
        for (int a = three ? 0 : 1; a < 3; a++)
            if (!tw.count(P->geos[t].L[a])) {
                const float2* p;
                lucem_status st = twiddles_for(client, P->geos[t].L[a], &p);
                if (st) return st;
                tw[P->geos[t].L[a]] = p;
            }

    // ---- memory
    struct TileMem {
        float2 *sa, *sb, *sx, *sr, *gh, *grev, *ta_h, *tb_h, *ta_r, *tb_r;
        float *xs, *us, *cs, *ck, *xa, *xbar, *uT, *gout_h, *gout_r;
        double* partial;
        int64_t nstride, spec_slot;
    };
    std::vector<TileMem> tm(ntiles);
    std::vector<int> row_blocks(ntiles), first_block(ntiles);
    std::vector<float2*> otf1(ngroups), otf2(ngroups), pa1(ngroups), pa2(ngroups), pb1(ngroups), pb2(ngroups);
    float *h = nullptr, *rev = nullptr, *gh_all = nullptr, *gr_all = nullptr;
    GradTile *d_tiles = nullptr, *d_fin_h = nullptr, *d_fin_r = nullptr;
    RowPass* d_otf_rows = nullptr;
    ColPass* d_otf_cols[2] = {nullptr, nullptr};
    ColPass *d_f1s[3], *d_f1[3], *d_f2s[3], *d_f2[3], *d_a1[3], *d_a2[3], *d_gi_h[2], *d_gi_r[2];
    PsfPrep* d_psf = nullptr;
    int *d_first = nullptr, *d_count = nullptr;
    int2 *it_row = nullptr, *it_col[3] = {nullptr, nullptr, nullptr}, *it_otf_row = nullptr;
    int2 *it_otf_col[2] = {nullptr, nullptr}, *it_gi[2] = {nullptr, nullptr}, *it_fin = nullptr;
    // host-side items (sizes known from the geometry)
    std::vector<int2> h_it_row, h_it_col[3], h_it_otf_row, h_it_otf_col[2], h_it_gi[2], h_it_fin;
    std::vector<RowPass> h_otf_rows(2 * ngroups);
    std::vector<ColPass> h_otf_cols[2];
    for (auto& v : h_otf_cols) v.resize(2 * ngroups);
    std::vector<ColPass> h_f1s[3], h_f1[3], h_f2s[3], h_f2[3], h_a1[3], h_a2[3], h_gi_h[2], h_gi_r[2];
    for (int i = 0; i < 3; i++)
        for (auto* v : {&h_f1s[i], &h_f1[i], &h_f2s[i], &h_f2[i], &h_a1[i], &h_a2[i]}) v->resize(ntiles);
    for (int i = 0; i < 2; i++) {
        h_gi_h[i].resize(ntiles);
        h_gi_r[i].resize(ntiles);
    }
    std::vector<GradTile> h_fin_h(ntiles), h_fin_r(ntiles);
    P->h_tiles.assign(ntiles, GradTile{});
    int total_blocks = 0;

    int64_t gh_total = 0;  // floats in the spectral PSF-gradient accumulators of all tiles
    auto layout = [&](Carve& cv) {
        // accumulators of every tile in one block, so that one launch clears them all
        gh_total = 0;
        for (int t = 0; t < ntiles; t++)
            gh_total += 2 * (int64_t)P->geos[t].L[0] * P->geos[t].L[1] * P->geos[t].pitch;
        float2* gh_pool = cv.take<float2>((size_t)gh_total);
        gh_total *= 2;
        int64_t gh_off = 0;
        for (int t = 0; t < ntiles; t++) {
            const Geometry& g = P->geos[t];
            TileMem& m = tm[t];
            const int64_t n = (int64_t)g.n[0] * g.n[1] * g.n[2];
            m.nstride = (n + 63) & ~int64_t(63);
            m.spec_slot = (int64_t)g.L[0] * g.L[1] * g.pitch;
            m.sa = cv.take<float2>((size_t)g.n[0] * g.n[1] * g.pitch);
            m.sb = three ? cv.take<float2>((size_t)g.n[0] * g.L[1] * g.pitch) : nullptr;
            m.sx = cv.take<float2>((size_t)(K + 1) * m.spec_slot);
            m.sr = cv.take<float2>((size_t)K * m.spec_slot);
            m.gh = gh_pool ? gh_pool + gh_off : nullptr;  // gh then grev, contiguous
            gh_off += 2 * m.spec_slot;
            m.grev = m.gh ? m.gh + m.spec_slot : nullptr;
            m.xs = cv.take<float>((size_t)(K + 1) * m.nstride);
            m.us = cv.take<float>((size_t)K * m.nstride);
            m.cs = cv.take<float>((size_t)K * m.nstride);
            m.ck = cv.take<float>((size_t)std::max(S, 1) * m.nstride);
            m.xa = cv.take<float>(m.nstride);
            m.xbar = cv.take<float>(m.nstride);
            m.uT = cv.take<float>(m.nstride);
            m.ta_h = cv.take<float2>((size_t)g.k[0] * g.k[1] * g.pitch);
            m.ta_r = cv.take<float2>((size_t)g.k[0] * g.k[1] * g.pitch);
            m.tb_h = three ? cv.take<float2>((size_t)g.k[0] * g.L[1] * g.pitch) : nullptr;
            m.tb_r = three ? cv.take<float2>((size_t)g.k[0] * g.L[1] * g.pitch) : nullptr;
        }
        for (int gi = 0; gi < ngroups; gi++) {
            const Geometry& g = P->geos[group_tile[gi]];
            const size_t otf = (size_t)g.L[0] * g.L[1] * g.pitch;
            otf1[gi] = cv.take<float2>(otf);
            otf2[gi] = cv.take<float2>(otf);
            pa1[gi] = cv.take<float2>((size_t)g.k[0] * g.k[1] * g.pitch);
            pa2[gi] = cv.take<float2>((size_t)g.k[0] * g.k[1] * g.pitch);
            pb1[gi] = three ? cv.take<float2>((size_t)g.k[0] * g.L[1] * g.pitch) : nullptr;
            pb2[gi] = three ? cv.take<float2>((size_t)g.k[0] * g.L[1] * g.pitch) : nullptr;
        }
        h = cv.take<float>(P->nk);
        rev = cv.take<float>(P->nk);
        P->psf_copy = cv.take<float>(P->nk);
        P->gpsf_buf = cv.take<float>(P->nk);
        gh_all = cv.take<float>((size_t)ntiles * P->nk);
        gr_all = cv.take<float>((size_t)ntiles * P->nk);
        P->loss_buf = cv.take<double>(ntiles);
        double* partials = cv.take<double>(std::max(total_blocks, 1));
        for (int t = 0; t < ntiles; t++) tm[t].partial = partials ? partials + first_block[t] : nullptr;
        d_first = cv.take<int>(ntiles);
        d_count = cv.take<int>(ntiles);
        d_tiles = cv.take<GradTile>(ntiles);
        d_fin_h = cv.take<GradTile>(ntiles);
        d_fin_r = cv.take<GradTile>(ntiles);
        d_otf_rows = cv.take<RowPass>(2 * ngroups);
        for (int i = 0; i < 2; i++) d_otf_cols[i] = cv.take<ColPass>(2 * ngroups);
        for (int i = 0; i < 3; i++) {
            d_f1s[i] = cv.take<ColPass>(ntiles);
            d_f1[i] = cv.take<ColPass>(ntiles);
            d_f2s[i] = cv.take<ColPass>(ntiles);
            d_f2[i] = cv.take<ColPass>(ntiles);
            d_a1[i] = cv.take<ColPass>(ntiles);
            d_a2[i] = cv.take<ColPass>(ntiles);
        }
        for (int i = 0; i < 2; i++) {
            d_gi_h[i] = cv.take<ColPass>(ntiles);
            d_gi_r[i] = cv.take<ColPass>(ntiles);
        }
        d_psf = cv.take<PsfPrep>(1);
        it_row = cv.take<int2>(h_it_row.size());
        for (int i = 0; i < 3; i++) it_col[i] = cv.take<int2>(h_it_col[i].size());
        it_otf_row = cv.take<int2>(h_it_otf_row.size());
        for (int i = 0; i < 2; i++) it_otf_col[i] = cv.take<int2>(h_it_otf_col[i].size());
        for (int i = 0; i < 2; i++) it_gi[i] = cv.take<int2>(h_it_gi[i].size());
        it_fin = cv.take<int2>(h_it_fin.size());
    };

    const float eps = rl_eps(*settings);
    // launch shapes: every tile joins each launch, so a launch's parallelism is the whole batch's
    ShapePolicy gsp{client->num_cus, 0, 0, nullptr};
    const ShapePolicy osp{client->num_cus, 0, 0, nullptr};
    for (int t = 0; t < ntiles; t++) {
        gsp.launch_pairs += ((int64_t)P->geos[t].n[0] * P->geos[t].n[1] + 1) / 2;
        gsp.launch_lines += (int64_t)P->geos[t].Lc * P->geos[t].n[0];
    }
    auto build = [&]() {
        for (int t = 0; t < ntiles; t++) {
            const Geometry& g = P->geos[t];
            TileMem& m = tm[t];
            const float2* tws[3] = {three ? tw[g.L[0]] : nullptr, tw[g.L[1]], tw[g.L[2]]};
            GradTile G{};
            G.r = make_row_desc(g, gsp, tws[2], eps, nullptr, nullptr, nullptr, m.sa, &row_blocks[t]);
            G.kx = g.k[2];
            G.nstride = m.nstride;
            G.xslots = m.xs;
            G.uslots = m.us;
            G.cslots = m.cs;
            G.ckpt = m.ck;
            G.xa = m.xa;
            G.xbar = m.xbar;
            G.partial = m.partial;
            G.uT = m.uT;
            P->h_tiles[t] = G;
            ColPass d[3];
            int blk[3];
            const int gi = group_of[t];
            const int n1 = make_corr_descs(g, gsp, m.sa, m.sb, otf1[gi], tws, d, blk);
            for (int i = 0; i < n1; i++) {
                h_f1[i][t] = d[i];
                h_f1s[i][t] = d[i];
            }
            make_corr_descs(g, gsp, m.sa, m.sb, otf2[gi], tws, d, blk);
            for (int i = 0; i < n1; i++) {
                h_f2[i][t] = d[i];
                h_f2s[i][t] = d[i];
            }
            const int fi = three ? 1 : 0;  // the fused pass saves the padded input's spectrum
            h_f1s[fi][t].save = m.sx;
            h_f1s[fi][t].slot_stride = m.spec_slot;
            h_f2s[fi][t].save = m.sr;
            h_f2s[fi][t].slot_stride = m.spec_slot;
            // transposed correlations: crop-transpose gather, conj transfer, pad-transpose fold
            for (int w = 0; w < 2; w++) {
                const float2* otf = w == 0 ? otf1[gi] : otf2[gi];
                ColPass a[3];
                make_corr_descs(g, gsp, m.sa, m.sb, otf, tws, a, blk);
                const int ax_last = three ? 0 : 1;  // the fused axis
                ColPass& fu = a[fi];
                fu.op = COL_FUSED_T;
                fu.left = 0;
                fu.mode = LUCEM_PAD_CONSTANT;
                fu.fold_left = g.k[ax_last] / 2;
                fu.fold_mode = g.mode;
                fu.fold_n = g.n[ax_last];
                fu.fold_end = g.n[ax_last] + g.k[ax_last] - 1;
                fu.save = w == 0 ? m.sx : m.sr;
                fu.gacc = w == 0 ? m.grev : m.gh;
                fu.slot_stride = m.spec_slot;
                if (three) {
                    a[0].left = 0;  // y forward of the crop transpose
                    a[0].mode = LUCEM_PAD_CONSTANT;
                    a[2].op = COL_INV_FOLD;
                    a[2].fold_left = g.k[1] / 2;
                    a[2].fold_mode = g.mode;
                    a[2].fold_n = g.n[1];
                    a[2].fold_end = g.n[1] + g.k[1] - 1;
                }
                for (int i = 0; i < n1; i++) (w == 0 ? h_a1 : h_a2)[i][t] = a[i];
            }
            // inverse transforms of the PSF-gradient spectra, cropped to the PSF
            for (int w = 0; w < 2; w++) {
                float2* gsrc = w == 0 ? m.gh : m.grev;
                float2* ta = w == 0 ? m.ta_h : m.ta_r;
                float2* tb = w == 0 ? m.tb_h : m.tb_r;
                auto& gi_arr = w == 0 ? h_gi_h : h_gi_r;
                int var = 0;
                ColPass c{};
                c.op = COL_INV;
                c.Lc = g.Lc;
                if (three) {
                    c.f = make_plan(g.L[0]);
                    c.tw = tws[0];
                    c.n_src = g.L[0];
                    c.n_dst = g.k[0];
                    c.src = gsrc;
                    c.src_outer = g.pitch;
                    c.src_axis = (int64_t)g.L[1] * g.pitch;
                    c.dst = tb;
                    c.dst_outer = g.pitch;
                    c.dst_axis = (int64_t)g.L[1] * g.pitch;
                    c.B = choose_B(g.L[0], g.Lc, g.L[1], SHAPE_COL_INV, osp, &c.wide, &var);
                    c.f = make_plan(g.L[0], var);
                    c.nkb = (g.Lc + c.B - 1) / c.B;
                    gi_arr[0][t] = c;
                    ColPass y = c;
                    y.f = make_plan(g.L[1]);
                    y.tw = tws[1];
                    y.n_src = g.L[1];
                    y.n_dst = g.k[1];
                    y.src = tb;
                    y.src_outer = (int64_t)g.L[1] * g.pitch;
                    y.src_axis = g.pitch;
                    y.dst = ta;
                    y.dst_outer = (int64_t)g.k[1] * g.pitch;
                    y.dst_axis = g.pitch;
                    y.B = choose_B(g.L[1], g.Lc, g.k[0], SHAPE_COL_INV, osp, &y.wide, &var);
                    y.f = make_plan(g.L[1], var);
                    y.nkb = (g.Lc + y.B - 1) / y.B;
                    gi_arr[1][t] = y;
                } else {
                    c.f = make_plan(g.L[1]);
                    c.tw = tws[1];
                    c.n_src = g.L[1];
                    c.n_dst = g.k[1];
                    c.src = gsrc;
                    c.src_outer = 0;
                    c.src_axis = g.pitch;
                    c.dst = ta;
                    c.dst_outer = 0;
                    c.dst_axis = g.pitch;
                    c.B = choose_B(g.L[1], g.Lc, 1, SHAPE_COL_INV, osp, &c.wide, &var);
                    c.f = make_plan(g.L[1], var);
                    c.nkb = (g.Lc + c.B - 1) / c.B;
                    gi_arr[0][t] = c;
                }
                GradTile F{};
                Geometry kg = g;
                kg.n[0] = g.k[0];
                kg.n[1] = g.k[1];
                kg.n[2] = g.k[2];
                int fb = 0;
                F.r = make_row_desc(kg, osp, tws[2], eps, nullptr, nullptr, nullptr, ta, &fb);
                F.out = (w == 0 ? gh_all : gr_all) ? (w == 0 ? gh_all : gr_all) + (int64_t)t * P->nk : nullptr;
                F.scale = (float)(0.25 / ((double)g.L[0] * (double)g.L[1] * (double)g.L[2]));
                (w == 0 ? h_fin_h : h_fin_r)[t] = F;
            }
        }
        for (int gi = 0; gi < ngroups; gi++) {
            const Geometry& g = P->geos[group_tile[gi]];
            const float2* tws[3] = {three ? tw[g.L[0]] : nullptr, tw[g.L[1]], tw[g.L[2]]};
            for (int w = 0; w < 2; w++) {
                RowPass r;
                int rb;
                ColPass cols[2];
                int cb[2];
                make_otf_descs(g, osp, w == 0 ? rev : h, w == 0 ? pa1[gi] : pa2[gi], w == 0 ? pb1[gi] : pb2[gi],
                               w == 0 ? otf1[gi] : otf2[gi], tws, &r, &rb, cols, cb);
                h_otf_rows[2 * gi + w] = r;
                for (int i = 0; i < (three ? 2 : 1); i++) h_otf_cols[i][2 * gi + w] = cols[i];
            }
        }
    };

    // items (from a sizing pass of the descriptors)
    {
        Carve sizing;
        layout(sizing);
        build();
        total_blocks = 0;
        for (int t = 0; t < ntiles; t++) {
            first_block[t] = total_blocks;
            total_blocks += row_blocks[t];
            for (int q = 0; q < row_blocks[t]; q++) h_it_row.push_back(make_int2(t, q));
            const Geometry& g = P->geos[t];
            for (int i = 0; i < ncorr; i++) {
                const ColPass& c = h_f1[i][t];
                const int outers = !three ? 1 : (i == 1 ? g.L[1] : g.n[0]);
                for (int q = 0; q < outers * c.nkb; q++) h_it_col[i].push_back(make_int2(t, q));
            }
            const int ngi = three ? 2 : 1;
            for (int i = 0; i < ngi; i++) {
                const ColPass& c = h_gi_h[i][t];
                const int outers = !three ? 1 : (i == 0 ? g.L[1] : g.k[0]);
                for (int q = 0; q < outers * c.nkb; q++) h_it_gi[i].push_back(make_int2(t, q));
            }
            const int frows = g.k[0] * g.k[1];
            const int fblocks = (((frows + 1) / 2) + h_fin_h[t].r.P - 1) / h_fin_h[t].r.P;
            for (int q = 0; q < fblocks; q++) h_it_fin.push_back(make_int2(t, q));
        }
        for (int gi = 0; gi < ngroups; gi++) {
            const Geometry& g = P->geos[group_tile[gi]];
            for (int w = 0; w < 2; w++) {
                const RowPass& r = h_otf_rows[2 * gi + w];
                const int rb = (((r.rows + 1) / 2) + r.P - 1) / r.P;
                for (int q = 0; q < rb; q++) h_it_otf_row.push_back(make_int2(2 * gi + w, q));
                for (int i = 0; i < (three ? 2 : 1); i++) {
                    const ColPass& c = h_otf_cols[i][2 * gi + w];
                    const int outers = !three ? 1 : (i == 0 ? g.k[0] : g.L[1]);
                    for (int q = 0; q < outers * c.nkb; q++) h_it_otf_col[i].push_back(make_int2(2 * gi + w, q));
                }
            }
        }
    }
    Carve sizing;
    layout(sizing);
    P->mem_bytes = sizing.off;
    {
        void* m = nullptr;
        if (lucem_status st = dev_malloc(client->device, &m, P->mem_bytes, "gradient plan")) return st;
        P->mem = m;
    }
    Carve real;
    real.base = static_cast<char*>(P->mem);
    layout(real);
    build();
    P->d_tiles = d_tiles;

    auto up = [&](void* dst, const void* src, size_t bytes) -> lucem_status {
        if (!bytes) return LUCEM_OK;
        hipError_t err = upload_sync(dst, src, bytes, client->stream);
        return err == hipSuccess ? LUCEM_OK : hip_fail(err, "hipMemcpy(gradient descriptors)");
    };
    lucem_status st;
    std::vector<int> counts(ntiles);
    for (int t = 0; t < ntiles; t++) counts[t] = row_blocks[t];
    PsfPrep pp{P->nk, P->psf_copy, h, rev};
    if ((st = up(d_first, first_block.data(), sizeof(int) * ntiles)) || (st = up(d_count, counts.data(), sizeof(int) * ntiles)) ||
        (st = up(d_fin_h, h_fin_h.data(), sizeof(GradTile) * ntiles)) ||
        (st = up(d_fin_r, h_fin_r.data(), sizeof(GradTile) * ntiles)) ||
        (st = up(d_otf_rows, h_otf_rows.data(), sizeof(RowPass) * h_otf_rows.size())) || (st = up(d_psf, &pp, sizeof pp)))
        return st;
    for (int i = 0; i < (three ? 2 : 1); i++)
        if ((st = up(d_otf_cols[i], h_otf_cols[i].data(), sizeof(ColPass) * h_otf_cols[i].size()))) return st;
    for (int i = 0; i < ncorr; i++)
        if ((st = up(d_f1s[i], h_f1s[i].data(), sizeof(ColPass) * ntiles)) || (st = up(d_f1[i], h_f1[i].data(), sizeof(ColPass) * ntiles)) ||
            (st = up(d_f2s[i], h_f2s[i].data(), sizeof(ColPass) * ntiles)) || (st = up(d_f2[i], h_f2[i].data(), sizeof(ColPass) * ntiles)) ||
            (st = up(d_a1[i], h_a1[i].data(), sizeof(ColPass) * ntiles)) || (st = up(d_a2[i], h_a2[i].data(), sizeof(ColPass) * ntiles)))
            return st;
    for (int i = 0; i < (three ? 2 : 1); i++)
        if ((st = up(d_gi_h[i], h_gi_h[i].data(), sizeof(ColPass) * ntiles)) || (st = up(d_gi_r[i], h_gi_r[i].data(), sizeof(ColPass) * ntiles)))
            return st;
    if ((st = up(it_row, h_it_row.data(), sizeof(int2) * h_it_row.size())) || (st = up(it_otf_row, h_it_otf_row.data(), sizeof(int2) * h_it_otf_row.size())) ||
        (st = up(it_fin, h_it_fin.data(), sizeof(int2) * h_it_fin.size())))
        return st;
    for (int i = 0; i < ncorr; i++)
        if ((st = up(it_col[i], h_it_col[i].data(), sizeof(int2) * h_it_col[i].size()))) return st;
    for (int i = 0; i < (three ? 2 : 1); i++)
        if ((st = up(it_otf_col[i], h_it_otf_col[i].data(), sizeof(int2) * h_it_otf_col[i].size())) ||
            (st = up(it_gi[i], h_it_gi[i].data(), sizeof(int2) * h_it_gi[i].size())))
            return st;

    // ---- launch sequences
    auto maxlds_row = [&](const auto& descs, const std::vector<int2>& items, auto get) {
        size_t mx = 0;
        for (const auto& it : items) {
            const RowPass& r = get(descs[it.x]);
            mx = std::max(mx, row_pass_lds_bytes(r.f.L, r.P));
        }
        return mx;
    };
    auto maxlds_col = [&](const std::vector<ColPass>& descs, const std::vector<int2>& items, int* wide) {
        size_t mx = 0;
        int w = 0;
        for (const auto& it : items) {
            mx = std::max(mx, col_pass_lds_bytes(descs[it.x].f.L, descs[it.x].B));
            w |= descs[it.x].wide;
        }
        *wide = w;
        return mx;
    };
    auto id = [](const RowPass& r) -> const RowPass& { return r; };  // OTF rows are plain RowPass
    auto gr = [](const GradTile& g) -> const RowPass& { return g.r; };
    // one thread count per launch: the widest tier any of its items needs (the others fit in it)
    auto wide_row = [&](const auto& descs, const std::vector<int2>& items, auto get) {
        int w = 0;
        for (const auto& it : items) w = std::max(w, get(descs[it.x]).wide);
        return w;
    };
    const size_t lds_row = maxlds_row(P->h_tiles, h_it_row, gr);
    const size_t lds_fin = maxlds_row(h_fin_h, h_it_fin, gr);
    const int wrow = wide_row(P->h_tiles, h_it_row, gr);
    const int wfin = std::max(wide_row(h_fin_h, h_it_fin, gr), wide_row(h_fin_r, h_it_fin, gr));
    int wide_col[3], wide_gi[2], wide_otf[2];
    size_t lds_col[3], lds_gi[2], lds_otf[2];
    for (int i = 0; i < ncorr; i++) lds_col[i] = maxlds_col(h_a1[i], h_it_col[i], &wide_col[i]);
    for (int i = 0; i < (three ? 2 : 1); i++) {
        lds_gi[i] = maxlds_col(h_gi_h[i], h_it_gi[i], &wide_gi[i]);
        lds_otf[i] = maxlds_col(h_otf_cols[i], h_it_otf_col[i], &wide_otf[i]);
    }
    const size_t lds_otf_row = maxlds_row(h_otf_rows, h_it_otf_row, id);
    const int wotf_row = wide_row(h_otf_rows, h_it_otf_row, id);
    const int nrow = (int)h_it_row.size();

    auto grow = [&](int op, int sx, int sx1, int su, int sxm, int scm, int ck, int flags) {
        GLaunch L{};
        L.kind = GLaunch::GROW;
        L.g = GradRowBatch{d_tiles, it_row, nrow, op, sx, sx1, su, sxm, scm, ck, flags, wrow};
        L.lds = lds_row;
        return L;
    };
    auto col = [&](ColPass* const* descs, int i, int op, int slot) {
        GLaunch L{};
        L.kind = GLaunch::COL;
        L.op = op;
        L.descs = descs[i];
        L.items = it_col[i];
        L.nitems = (int)h_it_col[i].size();
        L.lds = lds_col[i];
        L.wide = wide_col[i];
        L.slot = slot;
        return L;
    };
    const int fwd_ops3[3] = {COL_FWD, COL_FUSED, COL_INV};
    const int adj_ops3[3] = {COL_FWD, COL_FUSED_T, COL_INV_FOLD};
    auto corr = [&](std::vector<GLaunch>& seq, ColPass* const* descs, bool adj, int slot) {
        for (int i = 0; i < ncorr; i++)
            seq.push_back(col(descs, i, three ? (adj ? adj_ops3[i] : fwd_ops3[i]) : (adj ? COL_FUSED_T : COL_FUSED), slot));
    };
    auto xs = [&](int t) { return t % (K + 1); };
    auto us = [&](int t) { return t % K; };
    // one forward iteration t (save: store what the backward needs)
    auto fwd_iter = [&](std::vector<GLaunch>& seq, int t, bool save, int extra) {
        corr(seq, save ? d_f1s : d_f1, false, us(t));
        seq.push_back(grow(G_RATIO, 0, 0, us(t), 0, 0, 0, save ? GF_SAVE : 0));
        corr(seq, save ? d_f2s : d_f2, false, us(t));
        seq.push_back(grow(G_UPDATE, xs(t), xs(t + 1), us(t), 0, 0, (t + 1) / K, (save ? GF_SAVE : 0) | extra));
    };

    std::vector<GLaunch> fwd, bwd;
    // forward: transfer functions, segments, loss
    {
        GLaunch L{};
        L.kind = GLaunch::ROW;
        L.op = ROW_LOAD;
        L.descs = d_otf_rows;
        L.items = it_otf_row;
        L.nitems = (int)h_it_otf_row.size();
        L.lds = lds_otf_row;
        L.wide = wotf_row;
        fwd.push_back(L);
        for (int i = 0; i < (three ? 2 : 1); i++) {
            GLaunch c{};
            c.kind = GLaunch::COL;
            c.op = (i + 1 == (three ? 2 : 1)) ? COL_FWD_OTF : COL_FWD;
            c.descs = d_otf_cols[i];
            c.items = it_otf_col[i];
            c.nitems = (int)h_it_otf_col[i].size();
            c.lds = lds_otf[i];
            c.wide = wide_otf[i];
            fwd.push_back(c);
        }
    }
    fwd.push_back(grow(G_LOAD_X, xs(0), 0, 0, 0, 0, 0, GF_FROM_X0));
    for (int t = 0; t < T; t++) {
        const int s = t / K;
        const bool save = s == S - 1;
        int extra = 0;
        if ((t + 1) % K == 0 && (t + 1) / K < S) extra |= GF_CKPT;
        if (t + 1 == T) extra |= GF_LAST;
        fwd_iter(fwd, t, save, extra);
    }
    corr(fwd, d_f1s, false, K);  // u_T = A x_T, saving the spectrum of the padded x_T in slot K
    fwd.push_back(grow(G_LOSS, 0, 0, 0, 0, 0, 0, 0));

    // backward
    bwd.push_back(grow(G_LOSS_ADJ, 0, 0, 0, 0, 0, 0, 0));
    corr(bwd, d_a1, true, K);
    bwd.push_back(grow(G_ADJ_X, 0, 0, 0, xs(T - 1), us(T - 1), 0, GF_NO_XA | GF_SEED_NEXT));
    for (int s = S - 1; s >= 0; s--) {
        const int t0 = s * K, t1 = std::min(T, (s + 1) * K);
        if (s < S - 1) {
            bwd.push_back(grow(G_LOAD_X, xs(t0), 0, 0, 0, 0, s, s == 0 ? GF_FROM_X0 : 0));
            for (int t = t0; t < t1; t++) fwd_iter(bwd, t, true, 0);
            bwd.push_back(grow(G_SEED, 0, 0, 0, xs(t1 - 1), us(t1 - 1), 0, 0));
        }
        for (int t = t1 - 1; t >= t0; t--) {
            corr(bwd, d_a2, true, us(t));
            bwd.push_back(grow(G_ADJ_R, 0, 0, us(t), 0, 0, 0, 0));
            corr(bwd, d_a1, true, us(t));
            int flags = 0;
            if (t > t0)
                flags = GF_SEED_NEXT;
            else if (s == 0)
                flags = GF_TO_GX0;
            bwd.push_back(grow(G_ADJ_X, 0, 0, 0, t > t0 ? xs(t - 1) : 0, t > t0 ? us(t - 1) : 0, 0, flags));
        }
    }
    // PSF gradient: spectra -> real, cropped to the PSF
    for (int w = 0; w < 2; w++) {
        for (int i = 0; i < (three ? 2 : 1); i++) {
            GLaunch c{};
            c.kind = GLaunch::COL;
            c.op = COL_INV;
            c.descs = w == 0 ? d_gi_h[i] : d_gi_r[i];
            c.items = it_gi[i];
            c.nitems = (int)h_it_gi[i].size();
            c.lds = lds_gi[i];
            c.wide = wide_gi[i];
            bwd.push_back(c);
        }
        GLaunch L{};
        L.kind = GLaunch::GROW;
        L.g = GradRowBatch{w == 0 ? d_fin_h : d_fin_r, it_fin, (int)h_it_fin.size(), G_C2R, 0, 0, 0, 0, 0, 0, 0, wfin};
        L.lds = lds_fin;
        bwd.push_back(L);
    }

    // ---- capture
    if ((e = hipStreamCreateWithFlags(&P->capture, hipStreamNonBlocking)) != hipSuccess)
        return hip_fail(e, "hipStreamCreate(capture)");
    auto capture = [&](const std::vector<GLaunch>& seq, bool forward, hipGraph_t* graph, hipGraphExec_t* exec) -> lucem_status {
        hipError_t ce = hipStreamBeginCapture(P->capture, hipStreamCaptureModeThreadLocal);
        if (ce != hipSuccess) return hip_fail(ce, "hipStreamBeginCapture");
        int err = 0;
        if (forward) {
            err = launch_psf_batch(PsfBatch{d_psf, 1}, P->capture);
        } else {
            // the spectral accumulators of every tile start at zero
            err = launch_zero_floats(reinterpret_cast<float*>(tm[0].gh), gh_total, P->capture);
        }
        lucem_status rs = err ? hip_fail(err, "capture") : run_seq(P->capture, seq, nullptr);
        if (!rs && forward) {
            err = launch_loss_reduce(tm[0].partial, d_first, d_count, P->loss_buf, ntiles, P->capture);
            if (err) rs = hip_fail(err, "loss reduce");
        }
        if (!rs && !forward) {
            err = launch_psf_grad(gh_all, gr_all, ntiles, P->nk, P->psf_copy, P->gpsf_buf, P->capture);
            if (err) rs = hip_fail(err, "psf gradient");
        }
        hipGraph_t gph = nullptr;
        ce = hipStreamEndCapture(P->capture, &gph);
        if (rs) {
            if (gph) (void)hipGraphDestroy(gph);
            return rs;
        }
        if (ce != hipSuccess) return hip_fail(ce, "hipStreamEndCapture");
        *graph = gph;
        ce = hipGraphInstantiate(exec, gph, nullptr, nullptr, 0);
        return ce == hipSuccess ? LUCEM_OK : hip_fail(ce, "hipGraphInstantiate");
    };
    if ((st = capture(fwd, true, &P->g_fwd, &P->x_fwd))) return st;
    if ((st = capture(bwd, false, &P->g_bwd, &P->x_bwd))) return st;
    P->launches_fwd = fwd.size() + 2;
    P->launches_bwd = bwd.size() + ntiles + 1;

    if ((e = hipHostMalloc((void**)&P->h_stage, sizeof(GradTile) * ntiles, hipHostMallocDefault)) != hipSuccess)
        return hip_fail(e, "hipHostMalloc");
    if ((e = hipEventCreateWithFlags(&P->staged, hipEventDisableTiming)) != hipSuccess)
        return hip_fail(e, "hipEventCreate");
    *out = P.release();
    return LUCEM_OK;
}

static lucem_status upload_tiles(lucem_rl_grad* P) {
    hipError_t e;
    if (P->staged_pending) {
        if ((e = hipEventSynchronize(P->staged)) != hipSuccess) return hip_fail(e, "hipEventSynchronize");
        P->staged_pending = false;
    }
    std::memcpy(P->h_stage, P->h_tiles.data(), sizeof(GradTile) * P->h_tiles.size());
    if ((e = hipMemcpyAsync(P->d_tiles, P->h_stage, sizeof(GradTile) * P->h_tiles.size(), hipMemcpyHostToDevice,
                            P->client->stream)) != hipSuccess)
        return hip_fail(e, "hipMemcpyAsync(descriptors)");
    if ((e = hipEventRecord(P->staged, P->client->stream)) != hipSuccess) return hip_fail(e, "hipEventRecord");
    P->staged_pending = true;
    return LUCEM_OK;
}

extern "C" lucem_status lucem_rl_grad_forward_target(lucem_rl_grad* P, const float* const* x0, const float* const* y,
                                                     const float* const* target, const float* psf,
                                                     float* const* x_out, double* loss) {
    if (!P) return fail(LUCEM_INVALID_HANDLE, "gradient plan is NULL");
    if (!x0 || !y || !psf) return fail(LUCEM_INVALID_HANDLE, "x0, y or psf is NULL");
    lucem_client* c = P->client;
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    for (size_t t = 0; t < P->h_tiles.size(); t++) {
        if (!x0[t] || !y[t]) return fail(LUCEM_INVALID_HANDLE, "tile pointer is NULL");
        P->h_tiles[t].x0 = x0[t];
        P->h_tiles[t].r.obs = y[t];
        P->h_tiles[t].target = target ? target[t] : nullptr;
        P->h_tiles[t].x_out = x_out ? x_out[t] : nullptr;
    }
    lucem_status st = upload_tiles(P);
    if (st) return st;
    if ((e = hipMemcpyAsync(P->psf_copy, psf, sizeof(float) * P->nk, hipMemcpyDeviceToDevice, c->stream)) != hipSuccess)
        return hip_fail(e, "hipMemcpyAsync(psf)");
    if ((e = hipGraphLaunch(P->x_fwd, c->stream)) != hipSuccess) return hip_fail(e, "hipGraphLaunch");
    if (loss &&
        (e = hipMemcpyAsync(loss, P->loss_buf, sizeof(double) * P->h_tiles.size(), hipMemcpyDeviceToDevice, c->stream)) !=
            hipSuccess)
        return hip_fail(e, "hipMemcpyAsync(loss)");
    P->forward_done = true;
    return LUCEM_OK;
}

extern "C" lucem_status lucem_rl_grad_forward(lucem_rl_grad* P, const float* const* x0, const float* const* y,
                                              const float* psf, float* const* x_out, double* loss) {
    return lucem_rl_grad_forward_target(P, x0, y, nullptr, psf, x_out, loss);
}

extern "C" lucem_status lucem_rl_grad_backward(lucem_rl_grad* P, const float* loss_weight, float* const* grad_x0,
                                               float* grad_psf) {
    if (!P) return fail(LUCEM_INVALID_HANDLE, "gradient plan is NULL");
    if (!P->forward_done) return fail(LUCEM_INVALID_PARAMETER, "invalid parameter `backward`: no forward to differentiate");
    if (!grad_x0) return fail(LUCEM_INVALID_HANDLE, "grad_x0 is NULL");
    lucem_client* c = P->client;
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    for (size_t t = 0; t < P->h_tiles.size(); t++) {
        if (!grad_x0[t]) return fail(LUCEM_INVALID_HANDLE, "grad_x0 tile pointer is NULL");
        P->h_tiles[t].gx0 = grad_x0[t];
        P->h_tiles[t].weight = loss_weight ? loss_weight + t : nullptr;
    }
    lucem_status st = upload_tiles(P);
    if (st) return st;
    if ((e = hipGraphLaunch(P->x_bwd, c->stream)) != hipSuccess) return hip_fail(e, "hipGraphLaunch");
    if (grad_psf &&
        (e = hipMemcpyAsync(grad_psf, P->gpsf_buf, sizeof(float) * P->nk, hipMemcpyDeviceToDevice, c->stream)) != hipSuccess)
        return hip_fail(e, "hipMemcpyAsync(grad_psf)");
    return LUCEM_OK;
}

extern "C" lucem_status lucem_rl_grad_destroy(lucem_rl_grad* P) {
    delete P;
    return LUCEM_OK;
}

extern "C" uint64_t lucem_rl_grad_device_bytes(const lucem_rl_grad* P) { return P ? P->mem_bytes : 0; }
