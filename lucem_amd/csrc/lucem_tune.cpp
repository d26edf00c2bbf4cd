// Autotuner of the FFT launch shapes: for each pass kind (row pass, y/z line passes) and FFT
// length it times candidate shapes on the device and keeps the fastest:
//   - lines per workgroup (row pairs P, or columns B of a line pass), near the largest count that
//     fills the workgroup;
//   - the 256- or 512-thread kernel;
//   - the radix sequence (greedy composite radices, the same in reverse order, or single radices).
// Row shapes are timed on a ratio pass followed by an update pass (the plan's estimate stands in
// for the observed image), line shapes on the pass itself.
// Candidates run on the caller's own plan buffers (plan-owned workspace: the tuning launches
// leave garbage that the captured sequence overwrites before it reads), each one warm-up launch
// then the best of three timed launches (HIP events on the client's stream). Results are kept
// per process, keyed by (kind, L, op, launch size, schedule), so later plans of the same shapes
// skip the measurement (LUCEM_TUNE_CACHE=<file> keeps them across processes). LUCEM_AUTOTUNE=0
// turns tuning off (the heuristics of choose_P/choose_B).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/lucem_hip.h"
#include "lucem_internal_host.hpp"
#include "rl_internal.hpp"

using namespace lucem;

namespace {

// Candidate line counts: the largest few that fill a workgroup, the heuristic's choice and its
// neighbours, and a few small counts (launches of few lines want more workgroups).
std::vector<int> line_candidates(int max_lines, int heuristic) {
    std::vector<int> v;
    for (int x : {max_lines, max_lines - 1, max_lines - 2, heuristic, heuristic - 1, heuristic + 1, 1, 2, 4, 8})
        if (x >= 1 && x <= max_lines && std::find(v.begin(), v.end(), x) == v.end()) v.push_back(x);
    return v;
}

using CacheKey = std::tuple<int, int, int, int, int>;  // kind, L, op, log2(lines), mode
std::mutex g_cache_mu;
std::map<CacheKey, ShapeChoice> g_cache;
bool g_cache_loaded = false;

// LUCEM_TUNE_CACHE=<file> keeps the measured choices across processes: read once, appended to
// after each new measurement (one line per key: kind L op log2lines mode lines wide variant).
const char* cache_file() {
    const char* f = std::getenv("LUCEM_TUNE_CACHE");
    return f && *f ? f : nullptr;
}

void load_cache_locked() {
    if (g_cache_loaded) return;
    g_cache_loaded = true;
    const char* path = cache_file();
    if (!path) return;
    FILE* fp = std::fopen(path, "r");
    if (!fp) return;
    int k, L, op, lg, pers, lines, wide, var;
    while (std::fscanf(fp, "%d %d %d %d %d %d %d %d", &k, &L, &op, &lg, &pers, &lines, &wide, &var) == 8)
        if (lines >= 1 && wide >= 0 && wide <= 2 && var >= 0 && var <= kVariantStatic)
            g_cache[CacheKey{k, L, op, lg, pers}] = ShapeChoice{lines, wide, var};
    std::fclose(fp);
}

void store_locked(const CacheKey& key, const ShapeChoice& c) {
    g_cache[key] = c;
    const char* path = cache_file();
    if (!path) return;
    if (FILE* fp = std::fopen(path, "a")) {
        std::fprintf(fp, "%d %d %d %d %d %d %d %d\n", std::get<0>(key), std::get<1>(key), std::get<2>(key),
                     std::get<3>(key), std::get<4>(key), c.lines, c.wide, c.variant);
        std::fclose(fp);
    }
}

int shape_kind(const ColPass& c) {
    if (c.op == COL_FUSED) return SHAPE_COL_FUSED;
    if (c.op == COL_INV) return SHAPE_COL_INV;
    return SHAPE_COL_FWD;
}

int ilog2(int64_t v) {
    int r = 0;
    while (v > 1) {
        v >>= 1;
        r++;
    }
    return r;
}

bool same_plan(const FftPlan& a, const FftPlan& b) {
    return a.npass == b.npass && a.codes_lo == b.codes_lo && a.codes_hi == b.codes_hi;
}

// Device scratch of one tuning session: descriptors, work items and a queue counter.
struct Scratch {
    char* p = nullptr;
    size_t bytes = 0;
    ~Scratch() {
        if (p) (void)hipFree(p);
    }
    lucem_status reserve(size_t b) {
        if (b <= bytes) return LUCEM_OK;
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        hipError_t e = hipMalloc(&p, b);
        if (e != hipSuccess) return hip_fail(e, "hipMalloc(autotune scratch)");
        bytes = b;
        return LUCEM_OK;
    }
};

struct Timer {
    hipEvent_t a = nullptr, b = nullptr;
    ~Timer() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

// LUCEM_TUNE_LOG=1 prints each choice, 2 also every candidate's time.
int tune_log_level() {
    const char* v = std::getenv("LUCEM_TUNE_LOG");
    return v ? std::atoi(v) : 0;
}
bool tune_log() { return tune_log_level() >= 1; }

bool autotune_enabled() {
    const char* v = std::getenv("LUCEM_AUTOTUNE");
    return !(v && std::string(v) == "0");
}

// Times fn() (which enqueues one launch on `s`): one warm-up run, then the best of three.
template <class F>
lucem_status time_launch(hipStream_t s, Timer& t, F fn, float* ms) {
    hipError_t e;
    if (!t.a && (e = hipEventCreate(&t.a)) != hipSuccess) return hip_fail(e, "hipEventCreate");
    if (!t.b && (e = hipEventCreate(&t.b)) != hipSuccess) return hip_fail(e, "hipEventCreate");
    int err = fn();
    if (err) return hip_fail(err, "autotune launch");
    float best = 1e30f;
    for (int r = 0; r < 3; r++) {
        if ((e = hipEventRecord(t.a, s)) != hipSuccess) return hip_fail(e, "hipEventRecord");
        if ((err = fn())) return hip_fail(err, "autotune launch");
        if ((e = hipEventRecord(t.b, s)) != hipSuccess) return hip_fail(e, "hipEventRecord");
        if ((e = hipEventSynchronize(t.b)) != hipSuccess) return hip_fail(e, "hipEventSynchronize");
        float x = 0.f;
        if ((e = hipEventElapsedTime(&x, t.a, t.b)) != hipSuccess) return hip_fail(e, "hipEventElapsedTime");
        best = std::min(best, x);
    }
    *ms = best;
    return LUCEM_OK;
}

}  // namespace

// Tunes every (kind, L) group of these descriptors that carries at least 3% of its kind's work.
// mode: TUNE_PLAIN times the one-array kernels on the first descriptor; TUNE_BATCH_DIRECT and
// TUNE_BATCH_GRID time the batch kernels over all descriptors of a group, one workgroup per item or
// a persistent grid striding over the items, as the batched form runs them.
lucem_status lucem::tune_shapes(lucem_client* c, const std::vector<RowPass>& rows, const std::vector<TuneCol>& cols,
                                int mode, ShapeTable* out) {
    if (!autotune_enabled()) return LUCEM_OK;
    hipStream_t s = c->stream;
    Scratch scratch;
    Timer timer;

    // ---- row passes, grouped by L
    std::map<int, std::vector<RowPass>> rgroups;
    double rtotal = 0.0;
    for (const RowPass& r : rows) {
        rgroups[r.f.L].push_back(r);
        rtotal += (double)r.rows * r.f.L;
    }
    // One row launch runs one thread count (the batch launches every tile together): the group
    // with the most work decides it and the others are tuned under it.
    std::vector<std::pair<double, int>> rorder;
    for (auto& kv : rgroups) {
        double work = 0.0;
        for (const RowPass& r : kv.second) work += (double)r.rows * kv.first;
        rorder.push_back({-work, kv.first});
    }
    std::sort(rorder.begin(), rorder.end());
    int row_wide = -1;
    for (const auto& ro : rorder) {
        const int L = ro.second;
        std::vector<RowPass>& g = rgroups[L];
        const double work = -ro.first;
        int64_t pairs = 0;
        for (const RowPass& r : g) pairs += (r.rows + 1) / 2;
        if (work < 0.03 * rtotal || work < 2.0e5 || L > kLdsComplex) continue;
        const CacheKey key{SHAPE_ROW * 4 + (row_wide < 0 ? 2 : row_wide), L, ROW_RATIO, ilog2(pairs),
                           mode};
        {
            std::lock_guard<std::mutex> lk(g_cache_mu);
            load_cache_locked();
            auto it = g_cache.find(key);
            if (it != g_cache.end()) {
                (*out)[{SHAPE_ROW, L}] = it->second;
                if (row_wide < 0) row_wide = it->second.wide;
                continue;
            }
        }
        ShapeChoice best{};
        float best_ms = 1e30f;
        std::vector<FftPlan> plans;
        // variant 3: the compile-time ratio / update build, for the pair counts it is built for
        // (not a candidate where LUCEM_ROW_STATIC=0 switches it off)
        const char* rs = std::getenv("LUCEM_ROW_STATIC");
        const bool try_static = mode == TUNE_PLAIN && !(rs && *rs && std::atoi(rs) == 0);
        for (int v = 0; v < 4; v++) {
            const bool stat = v == kVariantStatic;
            if (stat && !try_static) continue;
            FftPlan f = make_plan(L, stat ? 0 : v);
            bool dup = false;
            for (const FftPlan& q : plans) dup = dup || same_plan(q, f);
            if (dup && !stat) continue;
            if (!stat) plans.push_back(f);
            for (int w = 0; w < 2; w++) {
            // 256 threads, or 512 when that fits more row pairs into a workgroup
            if (row_wide >= 0 && w != row_wide) continue;
            if (stat && w != 0) continue;
            const int pmax = std::max(1, std::min(16, (w ? kLdsComplexWide : kLdsComplex) / L));
            if (w == 1 && row_wide < 0 && pmax <= std::max(1, std::min(16, kLdsComplex / L))) continue;
            std::vector<int> cands;
            if (stat) {
                for (int P = 1; P <= pmax; P++)
                    if (row_static_available(L, P)) cands.push_back(P);
            } else {
                cands = line_candidates(pmax, w == 0 ? g[0].P : pmax);
            }
            for (int P : cands) {
                std::vector<RowPass> d = g;
                std::vector<int2> items;
                for (size_t i = 0; i < d.size(); i++) {
                    d[i].P = P;
                    d[i].wide = w;
                    d[i].f = f;
                    d[i].stat = stat ? 1 : 0;
                    d[i].op = ROW_UPDATE;
                    d[i].first = 0;
                    d[i].last = 0;
                    d[i].clip = 0;
                    d[i].obs = d[i].est;  // the ratio pass reads the plan's own estimate as "y"
                    const int nb = ((d[i].rows + 1) / 2 + P - 1) / P;
                    for (int q = 0; q < nb; q++) items.push_back(make_int2((int)i, q));
                }
                const size_t lds = row_pass_lds_bytes(L, P);
                float ms = 0.f;
                lucem_status st;
                if (mode != TUNE_PLAIN) {
                    const size_t bd = sizeof(RowPass) * d.size(), bi = sizeof(int2) * items.size();
                    if ((st = scratch.reserve(bd + bi + 1024))) return st;
                    RowPass* dd = reinterpret_cast<RowPass*>(scratch.p);
                    int2* di = reinterpret_cast<int2*>(scratch.p + ((bd + 255) & ~size_t(255)));
                    int* cnt = reinterpret_cast<int*>(scratch.p + ((bd + 255) & ~size_t(255)) + ((bi + 255) & ~size_t(255)));
                    hipError_t e;
                    if ((e = upload_sync(dd, d.data(), bd, s)) != hipSuccess)
                        return hip_fail(e, "hipMemcpy(autotune)");
                    if ((e = upload_sync(di, items.data(), bi, s)) != hipSuccess)
                        return hip_fail(e, "hipMemcpy(autotune)");
                    const int grid = mode == TUNE_BATCH_DIRECT ? (int)items.size()
                                                               : std::max(1, std::min((int)items.size(), batch_occupancy(true, w, lds) * c->num_cus));
                    // one ratio and one update pass, as an iteration runs them
                    (void)cnt;
                    const int stride = mode == TUNE_BATCH_GRID ? 1 : 0;
                    RowBatch ra{dd, di, (int)items.size(), ROW_RATIO, 0, 0, 0, w, nullptr, stride};
                    RowBatch up{dd, di, (int)items.size(), ROW_UPDATE, 0, 0, 0, w, nullptr, stride};
                    st = time_launch(s, timer, [&]() {
                        int err = launch_row_batch(ra, grid, lds, s);
                        return err ? err : launch_row_batch(up, grid, lds, s);
                    }, &ms);
                } else {
                    const int nb = (int)items.size();
                    RowPass ra = d[0];
                    ra.op = ROW_RATIO;
                    st = time_launch(s, timer, [&]() {
                        int err = launch_row_pass(ra, nb, s);
                        return err ? err : launch_row_pass(d[0], nb, s);
                    }, &ms);
                }
                if (st) return st;
                if (tune_log_level() >= 2)
                    std::fprintf(stderr, "lucem autotune:   row L=%d P=%d wide=%d variant=%d %.4f ms\n", L, P, w, v, ms);
                if (ms < best_ms) {
                    best_ms = ms;
                    best = ShapeChoice{P, w, v};
                }
            }
            }
        }
        (*out)[{SHAPE_ROW, L}] = best;
        if (row_wide < 0) row_wide = best.wide;
        if (tune_log())
            std::fprintf(stderr, "lucem autotune: row L=%d pairs=%lld -> P=%d wide=%d variant=%d (%.4f ms)\n", L,
                         (long long)pairs, best.lines, best.wide, best.variant, best_ms);
        std::lock_guard<std::mutex> lk(g_cache_mu);
        store_locked(key, best);
    }

    // ---- line passes, grouped by (kind, L)
    std::map<std::pair<int, int>, std::vector<TuneCol>> cgroups;
    std::map<int, double> ctotal;
    for (const TuneCol& tc : cols) {
        const int k = shape_kind(tc.d);
        cgroups[{k, tc.d.f.L}].push_back(tc);
        ctotal[k] += (double)tc.outers * tc.d.Lc * tc.d.f.L;
    }
    // one line launch runs one thread count: the widest group decides it, the others follow
    std::map<int, int> kind_wide;
    std::vector<std::pair<double, std::pair<int, int>>> order;
    for (auto& kv : cgroups) {
        double work = 0.0;
        for (const TuneCol& tc : kv.second) work += (double)tc.outers * tc.d.Lc * tc.d.f.L;
        order.push_back({-work, kv.first});
    }
    std::sort(order.begin(), order.end());
    for (const auto& o : order) {
        const int kind = o.second.first, L = o.second.second;
        const double work = -o.first;
        std::vector<TuneCol>& g = cgroups[o.second];
        if (work < 0.03 * ctotal[kind] || work < 2.0e5 || L > kLdsComplexWide) continue;
        int64_t lines = 0;
        for (const TuneCol& tc : g) lines += (int64_t)tc.outers * tc.d.Lc;
        const bool fixed_wide = kind_wide.count(kind) > 0;
        // a line longer than 4096 needs 512 threads: not tunable under a 256-thread launch (the
        // heuristic shape stays, and the launch runs at the widest thread count of its groups)
        if (L > kLdsComplex && fixed_wide && kind_wide[kind] == 0) continue;
        const int wide_key = fixed_wide ? kind_wide[kind] : 2;
        const CacheKey key{kind * 4 + wide_key, L, g[0].d.op, ilog2(lines), mode};
        {
            std::lock_guard<std::mutex> lk(g_cache_mu);
            load_cache_locked();
            auto it = g_cache.find(key);
            if (it != g_cache.end()) {
                (*out)[{kind, L}] = it->second;
                if (!fixed_wide) kind_wide[kind] = it->second.wide;
                continue;
            }
        }
        ShapeChoice best{};
        float best_ms = 1e30f;
        std::vector<FftPlan> plans;
        // variant 3: the compile-time COL_FUSED build, for the column counts it is built for
        // (not a candidate where LUCEM_COL_STATIC=0 switches it off)
        const char* cs = std::getenv("LUCEM_COL_STATIC");
        const int op0 = g[0].d.op;
        const bool try_static = mode == TUNE_PLAIN && (op0 == COL_FUSED || op0 == COL_FWD || op0 == COL_INV) &&
                                !g[0].d.save && !(cs && *cs && std::atoi(cs) == 0);
        const int static_tier = op0 == COL_FUSED ? 0 : 1;  // the thread count the builds have
        for (int v = 0; v < 4; v++) {
            const bool stat = v == kVariantStatic;
            if (stat && !try_static) continue;
            FftPlan f = make_plan(L, stat ? 0 : v);
            bool dup = false;
            for (const FftPlan& q : plans) dup = dup || same_plan(q, f);
            if (dup && !stat) continue;
            if (!stat) plans.push_back(f);
            for (int w = 0; w < 2; w++) {
                if (fixed_wide && w != kind_wide[kind]) continue;
                if (stat && w != static_tier) continue;
                if (w == 0 && L > kLdsComplex) continue;  // the line does not fit 256 threads
                const int cap = w ? kLdsComplexWide : kLdsComplex;
                const int bmax = std::max(1, std::min(32, cap / L));
                if (w == 1 && !fixed_wide && L <= kLdsComplex && bmax <= std::max(1, std::min(32, kLdsComplex / L)))
                    continue;
                std::vector<int> cands;
                if (stat) {
                    for (int B = 1; B <= bmax; B++)
                        if (op0 == COL_FUSED ? col_static_available(L, B) : line_static_available(L, B))
                            cands.push_back(B);
                } else {
                    cands = line_candidates(bmax, g[0].d.wide == w ? g[0].d.B : bmax);
                }
                for (int B : cands) {
                    std::vector<ColPass> d;
                    std::vector<int2> items;
                    for (size_t i = 0; i < g.size(); i++) {
                        ColPass cp = g[i].d;
                        cp.B = B;
                        cp.wide = w;
                        cp.f = f;
                        cp.stat = stat ? 1 : 0;
                        cp.nkb = (cp.Lc + B - 1) / B;
                        for (int q = 0; q < g[i].outers * cp.nkb; q++) items.push_back(make_int2((int)i, q));
                        d.push_back(cp);
                    }
                    const size_t lds = col_pass_lds_bytes(L, B);
                    float ms = 0.f;
                    lucem_status st;
                    if (mode != TUNE_PLAIN) {
                        const size_t bd = sizeof(ColPass) * d.size(), bi = sizeof(int2) * items.size();
                        if ((st = scratch.reserve(bd + bi + 1024))) return st;
                        ColPass* dd = reinterpret_cast<ColPass*>(scratch.p);
                        int2* di = reinterpret_cast<int2*>(scratch.p + ((bd + 255) & ~size_t(255)));
                        int* cnt =
                            reinterpret_cast<int*>(scratch.p + ((bd + 255) & ~size_t(255)) + ((bi + 255) & ~size_t(255)));
                        hipError_t e;
                        if ((e = upload_sync(dd, d.data(), bd, s)) != hipSuccess)
                            return hip_fail(e, "hipMemcpy(autotune)");
                        if ((e = upload_sync(di, items.data(), bi, s)) != hipSuccess)
                            return hip_fail(e, "hipMemcpy(autotune)");
                        const int grid = mode == TUNE_BATCH_DIRECT
                                             ? (int)items.size()
                                             : std::max(1, std::min((int)items.size(), batch_occupancy(false, w, lds) * c->num_cus));
                        (void)cnt;
                        ColBatch cb{dd, di, (int)items.size(), d[0].op, w, nullptr, 0, mode == TUNE_BATCH_GRID ? 1 : 0};
                        st = time_launch(s, timer, [&]() { return launch_col_batch(cb, grid, lds, s); }, &ms);
                    } else {
                        const int nb = (int)items.size();
                        st = time_launch(s, timer, [&]() { return launch_col_pass(d[0], nb, s); }, &ms);
                    }
                    if (st) return st;
                    if (tune_log_level() >= 2)
                        std::fprintf(stderr, "lucem autotune:   line kind=%d L=%d B=%d wide=%d variant=%d %.4f ms\n", kind,
                                     L, B, w, v, ms);
                    if (ms < best_ms) {
                        best_ms = ms;
                        best = ShapeChoice{B, w, v};
                    }
                }
            }
        }
        (*out)[{kind, L}] = best;
        if (!fixed_wide) kind_wide[kind] = best.wide;
        if (tune_log())
            std::fprintf(stderr, "lucem autotune: line kind=%d L=%d lines=%lld -> B=%d wide=%d variant=%d (%.4f ms)\n",
                         kind, L, (long long)lines, best.lines, best.wide, best.variant, best_ms);
        std::lock_guard<std::mutex> lk(g_cache_mu);
        store_locked(key, best);
    }
    return LUCEM_OK;
}

// The plain path: the tuned shapes of this geometry, measured on the first call of the client
// for it (on the workspace buffers) and reused by later calls. Arrays too small to matter keep
// the heuristics (an empty table).
const ShapeTable* lucem::tuned_lookup(lucem_client* c, const Geometry& g) {
    if (!autotune_enabled()) return nullptr;
    const std::vector<int> key{g.ndim, g.n[0], g.n[1], g.n[2], g.k[0], g.k[1], g.k[2]};
    auto it = c->tuned.find(key);
    return it == c->tuned.end() ? nullptr : &it->second;
}

lucem_status lucem::tuned_shapes(lucem_client* c, const Geometry& g, float* est, float2* sa, float2* sb,
                                 const float2* otf, const float2* const tws[3], const ShapeTable** out) {
    *out = nullptr;
    if (!autotune_enabled()) return LUCEM_OK;
    const std::vector<int> key{g.ndim, g.n[0], g.n[1], g.n[2], g.k[0], g.k[1], g.k[2]};
    auto it = c->tuned.find(key);
    if (it == c->tuned.end()) {
        // measuring synchronises the stream: under stream capture keep the heuristics
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing(c->stream, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return LUCEM_OK;
        const ShapePolicy sp{c->num_cus, 0, 0, nullptr};
        int rb = 0;
        const RowPass r = make_row_desc(g, sp, tws[2], 1.0f, nullptr, est, nullptr, sa, &rb);
        ColPass d[3];
        int b[3];
        const int n = make_corr_descs(g, sp, sa, sb, otf, tws, d, b);
        std::vector<TuneCol> cols;
        for (int i = 0; i < n; i++) cols.push_back(TuneCol{d[i], b[i] / std::max(d[i].nkb, 1)});
        ShapeTable t;
        lucem_status st = tune_shapes(c, std::vector<RowPass>{r}, cols, TUNE_PLAIN, &t);
        if (st) return st;
        it = c->tuned.emplace(key, t).first;
    }
    *out = &it->second;
    return LUCEM_OK;
}
