// Host side of the one-launch RL iteration for 2-D arrays (rl_tiled.hip): tile geometry, the
// check that every tile's boundary folds stay inside the tile, and the calls.
//
// Same operator as rl_launch (restore/deconvolve.rs:73-147): PSF normalisation and flip, the two
// transfer functions, then num_iter iterations from the constant 0.5, the last into `output` with
// the optional clip. Only the transform domain differs: each output tile is correlated on its own
// L0 x L1 domain (overlap-save) instead of the whole array's n + k - 1, and one launch carries
// both correlations, the ratio and the update. Values agree with the whole-array transforms to
// rounding.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/lucem_hip.h"
#include "fft_lds.hpp"
#include "lucem_internal_host.hpp"
#include "rl_internal.hpp"

using namespace lucem;

namespace {
// Even 7-smooth lengths the tile kernel takes (at most kTileMaxLen).
const int kTileLens[] = {16, 18, 20, 24, 28, 30, 32, 36, 40, 42, 48, 50, 54, 56, 60,
                         64, 70, 72, 80, 84, 90, 96, 98, 100, 108, 112, 120, 126, 128};

// Every pad target of every tile's rel window lies in that window: rel at a reflected position is
// then a value the tile computed itself. A tile's window covers its own outputs (the last tile's
// may be shorter than T) with the PSF's reach on either side. Interior windows map to themselves;
// only folds can fail (Wrap on a multi-tile axis, or an axis shorter than a fold).
bool axis_admits(int n, int k, int T, int mode) {
    const int left = k / 2;
    for (int t0 = 0; t0 < n; t0 += T) {
        const int m = std::min(T, n - t0) + k - 1;
        const int q0 = t0 - left;
        if (q0 >= 0 && q0 + m <= n) continue;
        for (int i = 0; i < m; i++) {
            const int s = pad_index(q0 + i, n, mode);
            if (s >= 0 && (s < q0 || s >= q0 + m)) return false;
        }
    }
    return true;
}

struct AxisChoice {
    int L, T;
};

// Candidate (L, T) of one axis: the longest tile each length allows, or the whole axis in the
// shortest length that holds it.
std::vector<AxisChoice> axis_candidates(int n, int k, int mode) {
    std::vector<AxisChoice> out;
    const int halo = 2 * (k - 1);
    bool whole = false;
    for (int L : kTileLens) {
        int T = L - halo;
        if (T < 1) continue;
        if (T >= n) {
            if (whole) continue;
            whole = true;
            T = n;
        }
        if (!axis_admits(n, k, T, mode)) continue;
        out.push_back({L, T});
    }
    return out;
}

// Both transforms of an L0 x L1 tile fit the kernel's threads (lines x length <= threads x values)
// and its row pairs fit the sweeps (at most kTileMaxLen / 2 pairs, columns no more than threads).
bool tile_fits(int L0, int L1) {
    const int64_t cap = (int64_t)kTileThreads * kTileElems;
    return L0 % 2 == 0 && L0 <= kTileMaxLen && L1 <= kTileMaxLen && (int64_t)L1 * tile_lines(L0 / 2) <= cap &&
           (int64_t)L0 * tile_lines(L1 / 2 + 1) <= cap;
}

bool tile_len_ok(int L) {
    for (int x : kTileLens)
        if (x == L) return true;
    return false;
}

bool tiled_mode_env_impl(int* mode) {
    const char* v = std::getenv("LUCEM_RL_TILED");
    if (!v) return false;
    *mode = std::atoi(v);
    return true;
}
}  // namespace

// Modelled time of one iteration on a grid: workgroups run in rounds of (CUs x workgroups per CU);
// a round costs a fixed latency (launch, the global reads and writes, the phase barriers) plus the
// tile's transform work, and two workgroups sharing a CU each run about 1.85x slower. Measured on
// MI355X (30-iteration calls): with constant lengths a 64 x 64 tile takes 16.5 us and a 112 x 112
// one 31 us (9.5 us + 1.7 ns per element-pass); the run-time-length build takes 27 us on 32 x 32
// and 46 us on 112 x 112 (26 us + 1.6 ns per element-pass).
static double grid_cost(int L0, int L1, int64_t tiles, int num_cus) {
    const size_t lds_cu = 160 * 1024;
    const int per_cu = tile_iter_lds_bytes(L0, L1) * 2 <= lds_cu ? 2 : 1;
    const int64_t slots = (int64_t)num_cus * per_cu;
    const int64_t rounds = (tiles + slots - 1) / slots;
    const double elem_passes = (double)L0 * L1 * (make_plan(L0).npass + make_plan(L1).npass + 2) / 6.0;
    const double round_us = tile_len_static(L0, L1) ? 9.5 + 1.72e-3 * elem_passes : 26.0 + 1.6e-3 * elem_passes;
    const bool shared = tiles > (int64_t)num_cus && per_cu == 2;
    return (double)rounds * round_us * (shared ? 1.85 : 1.0);
}

bool lucem::tile_plan(const Geometry& g, int num_cus, TileGrid* out) {
    if (g.ndim != 2 || g.origin[1] != 0 || g.origin[2] != 0) return false;
    const int n0 = g.n[1], n1 = g.n[2], k0 = g.k[1], k1 = g.k[2];
    const std::vector<AxisChoice> a0 = axis_candidates(n0, k0, g.mode), a1 = axis_candidates(n1, k1, g.mode);
    double best = 0.0;
    bool found = false;
    for (const AxisChoice& y : a0) {
        if (y.L % 2) continue;
        for (const AxisChoice& x : a1) {
            if (!tile_fits(y.L, x.L)) continue;
            const int64_t tiles = (int64_t)((n0 + y.T - 1) / y.T) * ((n1 + x.T - 1) / x.T);
            const double cost = grid_cost(y.L, x.L, tiles, num_cus);
            if (!found || cost < best) {
                best = cost;
                found = true;
                *out = TileGrid{y.L, x.L, y.T, x.T, (int)((n0 + y.T - 1) / y.T), (int)((n1 + x.T - 1) / x.T)};
            }
        }
    }
    return found;
}

// The canonical algorithmic bytes of one iteration on the whole array's minimal domain
// (24 n + 80 N_c), the figure the whole-array passes are measured against.
static double canonical_iteration_bytes(const Geometry& g) {
    const double Ly = g.n[1] + g.k[1] - 1, Lx = g.n[2] + g.k[2] - 1;
    const double Nc = Ly * (std::floor(Lx / 2) + 1);
    return 24.0 * g.n[1] * g.n[2] + 80.0 * Nc;
}

// Whether the plain 2-D path takes the tiled iteration: LUCEM_RL_TILED=0 never, =1 whenever a tile
// grid is admissible; otherwise when its modelled time is below that of the whole-array passes.
// Those take four launches of about 12 us each plus the canonical bytes at the 2.7 TB/s they reach
// (measured on MI355X from 256 x 256 to 3072 x 3072), so the tiled form takes every size measured.
bool lucem::use_tiled(const Geometry& g, int num_cus, TileGrid* plan) {
    if (!tile_plan(g, num_cus, plan)) return false;
    // LUCEM_TILE_GRID="L0,L1,T0,T1" replaces the chosen grid when it is admissible (measurement)
    if (const char* v = std::getenv("LUCEM_TILE_GRID")) {
        int L0 = 0, L1 = 0, T0 = 0, T1 = 0;
        if (std::sscanf(v, "%d,%d,%d,%d", &L0, &L1, &T0, &T1) == 4 && tile_len_ok(L0) && tile_len_ok(L1) && T0 >= 1 &&
            T1 >= 1 && T0 + 2 * (g.k[1] - 1) <= L0 && T1 + 2 * (g.k[2] - 1) <= L1 && tile_fits(L0, L1) &&
            axis_admits(g.n[1], g.k[1], T0, g.mode) && axis_admits(g.n[2], g.k[2], T1, g.mode))
            *plan = TileGrid{L0, L1, T0, T1, (g.n[1] + T0 - 1) / T0, (g.n[2] + T1 - 1) / T1};
    }
    int mode = 0;
    if (tiled_mode_env_impl(&mode)) return mode != 0;
    const double tiled_us = grid_cost(plan->L0, plan->L1, (int64_t)plan->tiles0 * plan->tiles1, num_cus);
    const double whole_us = 4 * 12.0 + canonical_iteration_bytes(g) / 2.7e6;
    if (const char* v = std::getenv("LUCEM_TUNE_LOG"))
        if (std::atoi(v) == 1)
            std::fprintf(stderr, "lucem tiled: %dx%d psf %dx%d -> L %dx%d, tiles %dx%d of %dx%d (model %.1f us; whole %.1f us)\n",
                         g.n[1], g.n[2], g.k[1], g.k[2], plan->L0, plan->L1, plan->tiles0, plan->tiles1, plan->T0,
                         plan->T1, tiled_us, whole_us);
    return tiled_us < whole_us;
}


TileIter lucem::tile_desc(const Geometry& g, const TileGrid& tp, const lucem_rl_settings& s, const float2* otf1,
                         const float2* otf2, const float2* tw0, const float2* tw1) {
    TileIter d{};
    d.n0 = g.n[1];
    d.n1 = g.n[2];
    d.T0 = tp.T0;
    d.T1 = tp.T1;
    d.left0 = g.k[1] / 2;
    d.left1 = g.k[2] / 2;
    d.k0 = g.k[1];
    d.k1 = g.k[2];
    d.L0 = tp.L0;
    d.L1 = tp.L1;
    d.tiles1 = tp.tiles1;
    d.mode = g.mode;
    d.pitch = ((tp.L1 / 2 + 1) + 15) & ~15;
    d.clip = s.clip ? 1 : 0;
    {
        const char* v = std::getenv("LUCEM_TILE_GENERIC");  // the run-time-length build, for comparison
        d.generic = v && std::atoi(v) != 0;
    }
    d.eps = rl_eps(s);
    d.init = 0.5f;  // deconvolve.rs:93
    d.f0 = make_plan(tp.L0);
    d.f1 = make_plan(tp.L1);
    d.otf1 = otf1;
    d.otf2 = otf2;
    d.tw0 = tw0;
    d.tw1 = tw1;
    d.own0y = d.own0x = 0;
    d.own1y = g.n[1];
    d.own1x = g.n[2];
    return d;
}

// The tile domain's geometry, on which the transfer functions are built: (1, L0, L1).
Geometry lucem::tile_geometry(const Geometry& g, const TileGrid& tp) {
    Geometry gt = g;
    gt.L[0] = 1;
    gt.L[1] = tp.L0;
    gt.L[2] = tp.L1;
    gt.Lc = tp.L1 / 2 + 1;
    gt.pitch = (gt.Lc + 15) & ~15;
    return gt;
}

TileSizes lucem::tile_sizes(const Geometry& g, const TileGrid& tp) {
    const Geometry gt = tile_geometry(g, tp);
    TileSizes z{};
    z.est = sizeof(float) * (size_t)g.n[1] * g.n[2];
    z.otf = sizeof(float2) * (size_t)tp.L0 * gt.pitch;
    z.psf = sizeof(float) * (size_t)g.k[1] * g.k[2];
    z.pa = sizeof(float2) * (size_t)g.k[1] * gt.pitch;
    return z;
}

lucem_status lucem::rl_tiled_launch(lucem_client* c, const Geometry& g, const lucem_rl_settings& s, const float* input,
                                    const float* psf, float* output, const TileBufs& b, const float2* tw0,
                                    const float2* tw1, const TileGrid& tp) {
    const Geometry gt = tile_geometry(g, tp);
    const float2* tws[3] = {nullptr, tw0, tw1};
    lucem_status st;
    // PSF normalisation and flip (deconvolve.rs:73-90), the transfer functions of rev (:120) and
    // h (:133) on the tile domain
    const int64_t nk = (int64_t)g.k[1] * g.k[2];
    PsfPrep pp{nk, psf, b.h, b.rev};
    const hipError_t e = (hipError_t)launch_psf_prep(pp, c->stream);
    if (e != hipSuccess) return hip_fail(e, "psf prep");
    if ((st = run_otf(c, gt, b.rev, b.pa, nullptr, b.otf1, tws))) return st;
    if ((st = run_otf(c, gt, b.h, b.pa, nullptr, b.otf2, tws))) return st;

    TileIter d = tile_desc(g, tp, s, b.otf1, b.otf2, tw0, tw1);
    d.obs = input;
    d.out = output;
    d.obs_rs = d.out_rs = g.n[2];
    const int blocks = tp.tiles0 * tp.tiles1;
    const double bytes = canonical_iteration_bytes(g);
    // ping-pong: a tile reads its neighbours' estimates while it writes its own
    float* cur = b.est_a;
    float* nxt = b.est_b;
    for (uint64_t it = 0; it < s.num_iter; it++) {
        d.first = it == 0;
        d.last = it + 1 == s.num_iter;
        d.est_in = cur;
        d.est_out = nxt;
        if ((st = run_tile(c, d, blocks, bytes, "tile_iter"))) return st;
        std::swap(cur, nxt);
    }
    return LUCEM_OK;
}

lucem_status lucem::rl_run_tiled(lucem_client* c, lucem_workspace* ws, const float* input, const float* psf,
                                 float* output, const Geometry& g, const lucem_rl_settings& s, const TileGrid& tp) {
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    const TileSizes z = tile_sizes(g, tp);
    TileBufs b{};
    lucem_status st;
    if ((st = ws_slot(ws, "rl_est", z.est, (void**)&b.est_a))) return st;
    if ((st = ws_slot(ws, "rl_est_b", z.est, (void**)&b.est_b))) return st;
    if ((st = ws_slot(ws, "rl_otf_conv", z.otf, (void**)&b.otf1))) return st;
    if ((st = ws_slot(ws, "rl_otf_mirror", z.otf, (void**)&b.otf2))) return st;
    if ((st = ws_slot(ws, "rl_psf_h", z.psf, (void**)&b.h))) return st;
    if ((st = ws_slot(ws, "rl_psf_rev", z.psf, (void**)&b.rev))) return st;
    if ((st = ws_slot(ws, "rl_psf_spec_a", z.pa, (void**)&b.pa))) return st;
    const float2 *tw0 = nullptr, *tw1 = nullptr;
    if ((st = twiddles_for(c, tp.L0, &tw0))) return st;
    if ((st = twiddles_for(c, tp.L1, &tw1))) return st;
    return rl_tiled_launch(c, g, s, input, psf, output, b, tw0, tw1, tp);
}

bool lucem::tile_batch_plan(const std::vector<Geometry>& gs, int num_cus, std::vector<TileGrid>* grids,
                            double* tiled_us, double* whole_us) {
    static const int kLens[] = {32, 48, 64, 80, 96, 112, 120, 128};
    // The per-pass batch streams its canonical bytes at about 3.3 TB/s (measured on the 300-tile
    // slide workload), four launches per iteration.
    double bytes = 0.0;
    for (const Geometry& g : gs) bytes += canonical_iteration_bytes(g);
    *whole_us = 4 * 8.0 + bytes / 3.3e6;
    double best = 0.0;
    bool found = false;
    for (int L : kLens) {
        if (!tile_fits(L, L) || !tile_len_static(L, L)) continue;
        std::vector<TileGrid> gr;
        int64_t items = 0;
        bool ok = true;
        for (const Geometry& g : gs) {
            if (g.ndim != 2 || g.origin[1] != 0 || g.origin[2] != 0) return false;
            int T[2];
            for (int a = 0; a < 2 && ok; a++) {
                const int n = g.n[1 + a], k = g.k[1 + a];
                const int t = std::min(L - 2 * (k - 1), n);
                ok = t >= 1 && axis_admits(n, k, t, g.mode);
                T[a] = t;
            }
            if (!ok) break;
            const TileGrid tg{L, L, T[0], T[1], (g.n[1] + T[0] - 1) / T[0], (g.n[2] + T[1] - 1) / T[1]};
            items += (int64_t)tg.tiles0 * tg.tiles1;
            gr.push_back(tg);
        }
        if (!ok) continue;
        // The 300-tile slide (8828 workgroups of 128 x 128, 34 rounds) measured 0.9x the
        // single-array model: the dispatcher starts a CU's next workgroup while others still run.
        const double cost = 0.9 * grid_cost(L, L, items, num_cus);
        if (!found || cost < best) {
            best = cost;
            found = true;
            *grids = gr;
        }
    }
    *tiled_us = best;
    return found;
}

int lucem::tiled_mode_env() {
    int mode = 0;
    return tiled_mode_env_impl(&mode) ? (mode != 0 ? 1 : 0) : -1;
}
