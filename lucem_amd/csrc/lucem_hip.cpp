// C ABI of the MI355X Richardson-Lucy library: client (device + stream), grow-only workspace,
// FFT planning and the plain 2-D / N-D entry points.
//
// Host-side contract mirrored from the reference:
//   richardson_lucy_into   crates/lucem-ops/src/restore/deconvolve.rs:45-149 (check order :53-71)
//   richardson_lucy        deconvolve.rs:152-162
//   Workspace              crates/lucem-core/src/memory/workspace.rs:17-37
//   PadMode::parse         crates/lucem-ops/src/core/pad.rs:48-57
//   element limit          crates/lucem-core/src/compute/limits.rs:11-22
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <memory>
#include <string>
#include <vector>

#include "../../include/lucem_hip.h"
#include "lucem_internal_host.hpp"
#include "rl_internal.hpp"
#include "rl_grad.hpp"

using namespace lucem;

// ---------------------------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------------------------
static thread_local std::string g_last_error;

lucem_status lucem::fail(lucem_status st, const std::string& msg) {
    g_last_error = msg;
    return st;
}

lucem_status lucem::hip_fail(int err, const char* what) {
    char buf[256];
    std::snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString((hipError_t)err));
    if (err == hipErrorOutOfMemory || err == hipErrorMemoryAllocation)
        return fail(LUCEM_OUT_OF_DEVICE_MEMORY, buf);
    return fail(LUCEM_HIP_ERROR, buf);
}

std::string lucem::dims_str(const uint64_t* dims, uint32_t ndim) {
    std::string s = "[";
    for (uint32_t i = 0; i < ndim; i++) {
        if (i) s += ", ";
        s += std::to_string(dims ? dims[i] : 0);
    }
    return s + "]";
}

extern "C" const char* lucem_last_error(void) { return g_last_error.c_str(); }
extern "C" const char* lucem_version(void) { return "lucem-mi355x 0.1.0 (gfx950)"; }

extern "C" void lucem_rl_settings_default(lucem_rl_settings* out) {
    if (!out) return;
    std::memset(out, 0, sizeof *out);
    out->num_iter = 30;
    out->clip = 1;
    out->has_filter_epsilon = 0;
    out->filter_epsilon = 1e-12;
    out->mode = LUCEM_PAD_REFLECT;
}

extern "C" lucem_status lucem_pad_mode_parse(const char* name, uint32_t* mode_out) {
    if (!name || !mode_out) return fail(LUCEM_INVALID_HANDLE, "pad mode name or output is NULL");
    std::string s(name);
    size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
    s = a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
    for (auto& c : s) c = (char)std::tolower((unsigned char)c);
    static const std::pair<const char*, uint32_t> table[] = {
        {"reflect", 0},  {"symmetric", 0}, {"grid-mirror", 0}, {"constant", 1}, {"grid-constant", 1},
        {"zeros", 1},    {"nearest", 2},   {"edge", 2},        {"replicate", 2}, {"mirror", 3},
        {"wrap", 4},     {"grid-wrap", 4}, {"periodic", 4},
    };
    for (const auto& e : table)
        if (s == e.first) {
            *mode_out = e.second;
            return LUCEM_OK;
        }
    return fail(LUCEM_UNKNOWN_PAD_MODE, "boundary pad mode '" + s +
                                            "' is not supported (expected reflect, constant, nearest, mirror, or wrap)");
}

// ---------------------------------------------------------------------------------------------
// client
// ---------------------------------------------------------------------------------------------
static std::atomic<bool> g_exiting{false};

namespace {
// Made on the first client creation, after the HIP runtime has started: its destructor is
// registered after the runtime's own exit-time teardown and so runs before it (exit handlers run
// in reverse order of registration).
struct ExitMark {
    ~ExitMark() { g_exiting.store(true); }
};
}  // namespace

bool lucem::exiting() { return g_exiting.load(); }

void lucem::client_retain(lucem_client* c) {
    if (c) c->refs.fetch_add(1);
}

void lucem::client_release(lucem_client* c) {
    if (c && c->refs.fetch_sub(1) == 1) delete c;
}

extern "C" lucem_status lucem_client_create(int device, void* stream, int32_t own_stream, lucem_client** out) {
    if (!out) return fail(LUCEM_INVALID_HANDLE, "client output pointer is NULL");
    *out = nullptr;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) return fail(LUCEM_HIP_ERROR, "no HIP device is available");
    static ExitMark exit_mark;
    (void)exit_mark;
    if (device < 0 || device >= count) return fail(LUCEM_INVALID_PARAMETER, "device index out of range");
    if ((e = hipSetDevice(device)) != hipSuccess) return hip_fail(e, "hipSetDevice");
    auto c = std::make_unique<lucem_client>();
    c->device = device;
    hipDeviceProp_t prop;
    if ((e = hipGetDeviceProperties(&prop, device)) != hipSuccess) return hip_fail(e, "hipGetDeviceProperties");
    c->num_cus = prop.multiProcessorCount;
    static std::once_flag lds_once;
    static int lds_err = 0;
    std::call_once(lds_once, [] {
        lds_err = allow_big_lds_kernels();
        if (!lds_err) lds_err = allow_big_lds_axis_kernels();
        if (!lds_err) lds_err = allow_big_lds_grad_kernels();
        if (!lds_err) lds_err = allow_tile_lds();
    });
    if (lds_err) return hip_fail(lds_err, "hipFuncSetAttribute(dynamic LDS)");
    noalloc_setup(device);
    if (!own_stream) {
        c->stream = (hipStream_t)stream;
        c->own_stream = false;
    } else {
        if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess)
            return hip_fail(e, "hipStreamCreate");
        c->own_stream = true;
    }
    c->internal_ws = std::make_unique<lucem_workspace>();
    c->internal_ws->client = c.get();
    c->internal_ws->internal = true;
    *out = c.release();
    return LUCEM_OK;
}

lucem_client::~lucem_client() {
    internal_ws.reset();
    if (lucem::exiting()) return;  // the runtime may be gone: leave device memory to the process exit
    (void)hipSetDevice(device);
    (void)hipStreamSynchronize(stream);
    for (auto& r : records) {
        (void)hipEventDestroy(r.start);
        (void)hipEventDestroy(r.stop);
    }
    for (auto ev : event_pool) (void)hipEventDestroy(ev);
    for (auto& kv : twiddles) (void)hipFree(kv.second);
    for (auto& kv : chirps) (void)hipFree(kv.second);
    for (auto& kv : long_twiddles) (void)hipFree(kv.second);
    if (own_stream && stream) (void)hipStreamDestroy(stream);
}

extern "C" lucem_status lucem_client_destroy(lucem_client* client) {
    client_release(client);  // freed now, or by the last workspace or plan made on it
    return LUCEM_OK;
}

extern "C" lucem_status lucem_client_sync(lucem_client* client) {
    if (!client) return fail(LUCEM_INVALID_HANDLE, "client is NULL");
    hipError_t e = hipStreamSynchronize(client->stream);
    if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
    return LUCEM_OK;
}

extern "C" void* lucem_client_stream(lucem_client* client) { return client ? (void*)client->stream : nullptr; }
extern "C" int lucem_client_num_cus(lucem_client* client) { return client ? client->num_cus : 0; }

// ---------------------------------------------------------------------------------------------
// profiling
// ---------------------------------------------------------------------------------------------
static lucem_status take_event(lucem_client* c, hipEvent_t* ev) {
    if (!c->event_pool.empty()) {
        *ev = c->event_pool.back();
        c->event_pool.pop_back();
        return LUCEM_OK;
    }
    hipError_t e = hipEventCreate(ev);
    if (e != hipSuccess) return hip_fail(e, "hipEventCreate");
    return LUCEM_OK;
}

extern "C" lucem_status lucem_client_profile(lucem_client* client, int32_t enable) {
    if (!client) return fail(LUCEM_INVALID_HANDLE, "client is NULL");
    (void)hipStreamSynchronize(client->stream);
    for (auto& r : client->records) {
        client->event_pool.push_back(r.start);
        client->event_pool.push_back(r.stop);
    }
    client->records.clear();
    client->profiling = enable != 0;
    return LUCEM_OK;
}

extern "C" lucem_status lucem_client_profile_read(lucem_client* client, lucem_kernel_stat* out, uint32_t max,
                                                  uint32_t* count) {
    if (!client) return fail(LUCEM_INVALID_HANDLE, "client is NULL");
    hipError_t e = hipStreamSynchronize(client->stream);
    if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
    std::vector<lucem_kernel_stat> agg;
    for (auto& r : client->records) {
        float ms = 0.f;
        e = hipEventElapsedTime(&ms, r.start, r.stop);
        if (e != hipSuccess) return hip_fail(e, "hipEventElapsedTime");
        lucem_kernel_stat* slot = nullptr;
        for (auto& a : agg)
            if (std::strncmp(a.name, r.name, sizeof a.name) == 0) slot = &a;
        if (!slot) {
            agg.emplace_back();
            slot = &agg.back();
            std::memset(slot, 0, sizeof *slot);
            std::snprintf(slot->name, sizeof slot->name, "%s", r.name);
        }
        slot->launches++;
        slot->total_ms += ms;
        slot->bytes += r.bytes;
        client->event_pool.push_back(r.start);
        client->event_pool.push_back(r.stop);
    }
    client->records.clear();
    for (uint32_t i = 0; i < max && i < agg.size(); i++) out[i] = agg[i];
    if (count) *count = (uint32_t)agg.size();
    return LUCEM_OK;
}

double lucem::row_pass_bytes(const RowPass& r) {
    const double R = r.rows, spec = R * r.Lc * 8.0, real = R * r.nx * 4.0;
    switch (r.op) {
        case ROW_LOAD: return real + spec;
        case ROW_LOAD_CONST: return spec;
        case ROW_RATIO: return spec + real + spec;
        default: return spec + (r.first ? 0.0 : real) + real + (r.last ? 0.0 : spec);
    }
}

double lucem::col_pass_bytes(const ColPass& p, int blocks) {
    const double outers = (double)(blocks / p.nkb);
    const double line = p.Lc * 8.0;
    double b = outers * line * (p.n_src + p.n_dst);
    if (p.op == COL_FUSED) b += outers * line * p.f.L;
    return b;
}

// Calls that may not create events (the allocation-free entry point) record nothing.
static thread_local int g_no_profile = 0;
lucem::NoProfileScope::NoProfileScope() { g_no_profile++; }
lucem::NoProfileScope::~NoProfileScope() { g_no_profile--; }

template <class Launch>
static lucem_status profiled(lucem_client* c, const char* name, double bytes, Launch&& launch) {
    hipEvent_t a = nullptr, b = nullptr;
    lucem_status st;
    const bool prof = c->profiling && g_no_profile == 0;
    if (prof) {
        if ((st = take_event(c, &a)) || (st = take_event(c, &b))) return st;
        (void)hipEventRecord(a, c->stream);
    }
    const int err = launch();
    if (err != hipSuccess) return hip_fail(err, name);
    if (prof) {
        (void)hipEventRecord(b, c->stream);
        c->records.push_back(ProfRecord{name, a, b, bytes});
    }
    return LUCEM_OK;
}

lucem_status lucem::run_row(lucem_client* c, const RowPass& r, int blocks, const char* name) {
    return profiled(c, name, row_pass_bytes(r), [&] { return launch_row_pass(r, blocks, c->stream); });
}

lucem_status lucem::run_col(lucem_client* c, const ColPass& p, int blocks, const char* name) {
    return profiled(c, name, col_pass_bytes(p, blocks), [&] { return launch_col_pass(p, blocks, c->stream); });
}

lucem_status lucem::run_tile(lucem_client* c, const TileIter& d, int blocks, double bytes, const char* name) {
    return profiled(c, name, bytes, [&] { return launch_tile_iter(d, blocks, c->stream); });
}

// Twiddle table exp(-2 pi i m / L), m < L, computed in double and rounded once to float.
lucem_status lucem::twiddles_for(lucem_client* c, int L, const float2** out) {
    auto it = c->twiddles.find(L);
    if (it != c->twiddles.end()) {
        *out = it->second;
        return LUCEM_OK;
    }
    std::vector<float2> host(L);
    const double two_pi = 6.283185307179586476925286766559;
    for (int m = 0; m < L; m++) {
        const double a = two_pi * (double)m / (double)L;
        host[m] = make_float2((float)std::cos(a), (float)(-std::sin(a)));
    }
    float2* dev = nullptr;
    hipError_t e = hipMalloc(&dev, sizeof(float2) * (size_t)std::max(L, 1));
    if (e != hipSuccess) return hip_fail(e, "hipMalloc(twiddles)");
    e = upload_sync(dev, host.data(), sizeof(float2) * (size_t)L, c->stream);
    if (e != hipSuccess) {
        (void)hipFree(dev);
        return hip_fail(e, "hipMemcpy(twiddles)");
    }
    c->twiddles[L] = dev;
    *out = dev;
    return LUCEM_OK;
}

// ---------------------------------------------------------------------------------------------
// workspace
// ---------------------------------------------------------------------------------------------
lucem_workspace::~lucem_workspace() {
    const int dev = client ? client->device : 0;
    if (!lucem::exiting()) {
        if (client) (void)hipStreamSynchronize(client->stream);
        for (auto& kv : slots) dev_free(dev, kv.second.ptr, kv.second.bytes);
    } else {
        for (auto& kv : slots) dev_forget(dev, kv.second.bytes);
    }
    if (!internal) client_release(client);
}

lucem_status lucem::ws_slot(lucem_workspace* ws, const char* name, size_t bytes, void** out) {
    Slot& s = ws->slots[name];
    if (s.bytes < bytes) {
        const int dev = ws->client->device;
        // the growth is checked against the budget (workspace.rs:32-35), before the old buffer goes
        if (lucem_status st = budget_check(dev, name, bytes - s.bytes)) return st;
        if (s.ptr) {
            // the old buffer may still be read by queued work
            (void)hipStreamSynchronize(ws->client->stream);
            dev_free(dev, s.ptr, s.bytes);
            s.ptr = nullptr;
            s.bytes = 0;
        }
        void* p = nullptr;
        char buf[160];
        std::snprintf(buf, sizeof buf, "workspace slot '%s' of %zu bytes", name, bytes);
        if (lucem_status st = dev_malloc(dev, &p, bytes, buf)) return st;
        s.ptr = p;
        s.bytes = bytes;
    }
    *out = s.ptr;
    return LUCEM_OK;
}

extern "C" lucem_status lucem_workspace_create(lucem_client* client, lucem_workspace** out) {
    if (!client || !out) return fail(LUCEM_INVALID_HANDLE, "client or workspace output is NULL");
    auto w = std::make_unique<lucem_workspace>();
    w->client = client;
    client_retain(client);
    *out = w.release();
    return LUCEM_OK;
}

extern "C" lucem_status lucem_workspace_destroy(lucem_workspace* ws) {
    delete ws;
    return LUCEM_OK;
}

extern "C" uint64_t lucem_workspace_capacity_bytes(const lucem_workspace* ws) {
    if (!ws) return 0;
    uint64_t total = 0;
    for (const auto& kv : ws->slots) total += kv.second.bytes;
    return total;
}

// ---------------------------------------------------------------------------------------------
// planning
// ---------------------------------------------------------------------------------------------
static bool smooth7(int v) {
    for (int p : {2, 3, 5, 7})
        while (v % p == 0) v /= p;
    return v == 1;
}

// The smallest 7-smooth length >= target, or a longer one within 10% when it needs fewer radix
// passes. Every pass kernel streams its lines through HBM once, which costs about as much as two
// LDS passes, so a length costs about L * (passes + 2): at 1048 that picks 1080 (radices 12, 10,
// 9) over 1050 (10, 7, 5, 3). The search stays within the first length's thread tier and within
// LUCEM_MAX_FFT_LEN.
int lucem::choose_fft_len(int target) {
    int L = std::max(target, 1);
    while (!smooth7(L)) L++;
    if (L > LUCEM_MAX_FFT_LEN) return L;
    const int lim = std::min(LUCEM_MAX_FFT_LEN, L + L / 10);
    int best = L;
    int64_t cost = (int64_t)L * (make_plan(L).npass + 2);
    for (int c = L + 1; c <= lim; c++) {
        if (!smooth7(c) || tier_for_length(c) != tier_for_length(L)) continue;
        const int64_t cc = (int64_t)c * (make_plan(c).npass + 2);
        if (cc < cost) {
            best = c;
            cost = cc;
        }
    }
    return best;
}

FftPlan lucem::make_plan(int L, int variant) {
    // Prime factors, then greedy grouping into the largest supported radix: 16 = 4*4, 12 = 4*3,
    // 10 = 5*2, 9 = 3*3, 8, 7, 6 = 3*2, 5, 4, 3, 2 (variant 2: 8, 7, 5, 4, 3, 2 only). Variant 1
    // runs the passes of variant 0 in reverse order. Radix 15 = 5*3 and 14 = 7*2 are left out:
    // with two butterflies per thread they need 118 VGPRs, where every other pass needs at most
    // 97, and the kernels' register budget is the largest pass's.
    FftPlan p{};
    p.L = L;
    p.npass = 0;
    p.tw = 0;
    int cnt[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int r = L;
    for (int q : {2, 3, 5, 7})
        while (r > 1 && r % q == 0) {
            cnt[q]++;
            r /= q;
        }
    struct Code {
        int r1, r2, n2, n3, n5, n7;
    };
    static const Code composite[] = {
        {4, 4, 4, 0, 0, 0}, {4, 3, 2, 1, 0, 0}, {5, 2, 1, 0, 1, 0},
        {3, 3, 0, 2, 0, 0}, {8, 1, 3, 0, 0, 0}, {7, 1, 0, 0, 0, 1}, {3, 2, 1, 1, 0, 0}, {5, 1, 0, 0, 1, 0},
        {4, 1, 2, 0, 0, 0}, {3, 1, 0, 1, 0, 0}, {2, 1, 1, 0, 0, 0},
    };
    static const Code single[] = {
        {8, 1, 3, 0, 0, 0}, {7, 1, 0, 0, 0, 1}, {5, 1, 0, 0, 1, 0}, {4, 1, 2, 0, 0, 0}, {3, 1, 0, 1, 0, 0},
        {2, 1, 1, 0, 0, 0},
    };
    const Code* codes = variant == 2 ? single : composite;
    const int ncodes = variant == 2 ? (int)(sizeof single / sizeof single[0]) : (int)(sizeof composite / sizeof composite[0]);
    unsigned char seq[32];
    int n = 0;
    for (;;) {
        bool found = false;
        for (int i = 0; i < ncodes; i++) {
            const Code& c = codes[i];
            if (cnt[2] >= c.n2 && cnt[3] >= c.n3 && cnt[5] >= c.n5 && cnt[7] >= c.n7) {
                cnt[2] -= c.n2;
                cnt[3] -= c.n3;
                cnt[5] -= c.n5;
                cnt[7] -= c.n7;
                seq[n++] = (unsigned char)(c.r1 | (c.r2 << 4));
                found = true;
                break;
            }
        }
        if (!found || n == 16) break;
    }
    if (variant == 1) std::reverse(seq, seq + n);
    for (int q = 0; q < n; q++) {
        const unsigned long long code = seq[q];
        if (q < 8)
            p.codes_lo |= code << (8 * q);
        else
            p.codes_hi |= code << (8 * (q - 8));
    }
    p.npass = n;
    return p;
}

// Launch shapes. LUCEM_TUNE (e.g. "row_P=4,col_B=8,col_wide=0") overrides the choices, for
// measurement sweeps.
static int tune_value(const char* key) {
    const char* env = std::getenv("LUCEM_TUNE");
    if (!env) return -1;
    const std::string s(env), k = std::string(key) + "=";
    size_t pos = 0;
    while ((pos = s.find(k, pos)) != std::string::npos) {
        if (pos == 0 || s[pos - 1] == ',') return std::atoi(s.c_str() + pos + k.size());
        pos += k.size();
    }
    return -1;
}

static const ShapeChoice* find_choice(const ShapePolicy& sp, int kind, int L) {
    if (!sp.table) return nullptr;
    auto it = sp.table->find(std::make_pair(kind, L));
    return it == sp.table->end() ? nullptr : &it->second;
}

// Row pairs per workgroup: as many lines of length L as fill the workgroup's 4096 complex
// values (any count, not only powers of two), at most 16, fewer when the launch would not give
// every CU four workgroups.
int lucem::choose_P(int L, int rows, const ShapePolicy& sp, int* tier, int* variant) {
    *variant = 0;
    // lines longer than 4096 take the 512- or 1024-thread kernel, one pair per workgroup
    *tier = tier_for_length(L);
    if (*tier > 0) return 1;
    if (const ShapeChoice* c = find_choice(sp, SHAPE_ROW, L)) {
        *variant = c->variant;
        *tier = c->wide;
        return c->lines;
    }
    const int64_t pairs = std::max<int64_t>((rows + 1) / 2, sp.launch_pairs);
    int P = std::max(1, std::min(16, kLdsComplex / L));
    while (sp.num_cus > 0 && P > 1 && (pairs + P - 1) / P < 4 * sp.num_cus) P--;
    const int t = tune_value("row_P");
    if (t >= 1 && (size_t)L * t <= (size_t)kLdsComplex) P = t;
    return P;
}

// Whether a COL_FUSED pass runs its compile-time build: when the tuner chose it (variant 3), or
// as LUCEM_COL_STATIC=1 / 0 forces, wherever a build of (L, B) at 256 threads exists.
int lucem::col_static_choice(bool tuned, int L, int B, int wide) {
    if (wide != 0 || !col_static_available(L, B)) return 0;
    const char* v = std::getenv("LUCEM_COL_STATIC");
    if (v && *v) return std::atoi(v) != 0 ? 1 : 0;
    return tuned ? 1 : 0;
}

// The same for the y passes (COL_FWD / COL_INV) of 3-D arrays, built at 512 threads; forced by
// LUCEM_COL_STATIC as well.
int lucem::line_static_choice(bool tuned, int L, int B, int wide) {
    if (wide != 1 || !line_static_available(L, B)) return 0;
    const char* v = std::getenv("LUCEM_COL_STATIC");
    if (v && *v) return std::atoi(v) != 0 ? 1 : 0;
    return tuned ? 1 : 0;
}

// Whether the ratio and update row passes run their compile-time build: when the tuner chose it
// (variant 3), or as LUCEM_ROW_STATIC=1 / 0 forces, wherever a build of (L, P) at 256 threads exists.
int lucem::row_static_choice(bool tuned, int L, int P, int wide) {
    if (wide != 0 || !row_static_available(L, P)) return 0;
    const char* v = std::getenv("LUCEM_ROW_STATIC");
    if (v && *v) return std::atoi(v) != 0 ? 1 : 0;
    return tuned ? 1 : 0;
}

// Columns per workgroup of a line pass, any count up to 32, fewer only when the launch would
// not give every CU one workgroup (wide workgroups beat more of them in the measurements). The
// 512-thread shape serves lines too long for 8 columns in 4096 values.
int lucem::choose_B(int L, int Lc, int64_t outers, int kind, const ShapePolicy& sp, int* wide, int* variant) {
    *variant = 0;
    if (L > kLdsComplexWide) {  // the 1024-thread kernel
        *wide = 2;
        return std::max(1, kLdsComplexXL / L);
    }
    const ShapeChoice* c = find_choice(sp, kind, L);
    if (c && (int64_t)L * c->lines <= (int64_t)tier_threads(c->wide) * kElemsPerThread) {  // a valid measured shape
        *variant = c->variant;
        *wide = c->wide;
        return c->lines;
    }
    const int b_small = std::max(1, std::min(32, kLdsComplex / L));
    const int b_big = std::max(1, std::min(32, kLdsComplexWide / L));
    int w = tune_value("col_wide");
    if (w < 0) w = (b_small < 8 && b_big > b_small) ? 1 : 0;
    if (L > kLdsComplex) w = 1;  // a line longer than the 256-thread kernel holds
    const size_t cap = w ? kLdsComplexWide : kLdsComplex;
    int B = w ? b_big : b_small;
    const int64_t lines = std::max<int64_t>(outers * Lc, sp.launch_lines);
    while (sp.num_cus > 0 && B > 4 && lines / B < sp.num_cus) B--;
    const int t = tune_value("col_B");
    if (t >= 1 && (size_t)L * t <= cap) B = t;
    *wide = w;
    return B;
}

bool lucem::make_geometry(const int64_t n[3], const int64_t k[3], int ndim, Geometry* g) {
    std::memset(g, 0, sizeof *g);
    g->ndim = ndim;
    for (int a = 0; a < 3; a++) {
        g->n[a] = (int)n[a];
        g->k[a] = (int)k[a];
        const int64_t need = n[a] + k[a] - 1;
        if (need > LUCEM_MAX_FFT_LEN) return false;
        g->L[a] = (ndim == 2 && a == 0) ? 1 : choose_fft_len((int)need);
        if (g->L[a] > LUCEM_MAX_FFT_LEN) return false;
    }
    g->Lc = g->L[2] / 2 + 1;
    g->pitch = (g->Lc + 15) & ~15;
    return true;
}

// ---------------------------------------------------------------------------------------------
// RL driver for one array (2-D or 3-D)
// ---------------------------------------------------------------------------------------------
static lucem_status launch_check(int err, const char* what) {
    if (err != hipSuccess) return hip_fail(err, what);
    return LUCEM_OK;
}

// Descriptors of the two-step transfer-function build: a real kernel (dims m, at the origin,
// zero-padded to L) -> row R2C -> line FFTs -> conj(v) * scale stored as [L_z][L_y][pitch].
int lucem::make_otf_descs(const Geometry& g, const ShapePolicy& sp, const float* w, float2* pa, float2* pb, float2* otf,
                          const float2* const tws[3], RowPass* row, int* row_blocks, ColPass cols[2],
                          int col_blocks[2]) {
    const int m[3] = {g.k[0], g.k[1], g.k[2]};
    // The kernel's own R2C pairs rows and carries a factor 2, and so does the data's: the
    // transfer function removes both (1/4) and the unnormalised inverse transforms (1/L).
    const float scale = (float)(0.25 / ((double)g.L[0] * (double)g.L[1] * (double)g.L[2]));
    RowPass r{};
    r.op = ROW_LOAD;
    r.nx = m[2];
    r.rows = m[0] * m[1];
    r.left = 0;
    r.mode = LUCEM_PAD_CONSTANT;
    r.Lc = g.Lc;
    r.pitch = g.pitch;
    int var = 0;
    r.P = choose_P(g.L[2], r.rows, sp, &r.wide, &var);
    r.f = make_plan(g.L[2], var);
    r.src = w;
    r.spec = pa;
    r.tw = tws[2];
    *row = r;
    *row_blocks = (((r.rows + 1) / 2) + r.P - 1) / r.P;

    ColPass y{};
    y.n_src = m[1];
    y.left = 0;
    y.mode = LUCEM_PAD_CONSTANT;
    y.n_dst = g.L[1];
    y.Lc = g.Lc;
    y.tw = tws[1];
    y.src = pa;
    y.src_outer = (int64_t)m[1] * g.pitch;
    y.src_axis = g.pitch;
    if (g.ndim == 2) {
        y.op = COL_FWD_OTF;
        y.scale = scale;
        y.dst = otf;
        y.dst_outer = 0;
        y.dst_axis = g.pitch;
        y.B = choose_B(g.L[1], g.Lc, 1, SHAPE_COL_FWD, sp, &y.wide, &var);
        y.f = make_plan(g.L[1], var);
        y.nkb = (g.Lc + y.B - 1) / y.B;
        cols[0] = y;
        col_blocks[0] = y.nkb;
        return 1;
    }
    y.op = COL_FWD;
    y.dst = pb;
    y.dst_outer = (int64_t)g.L[1] * g.pitch;
    y.dst_axis = g.pitch;
    y.B = choose_B(g.L[1], g.Lc, m[0], SHAPE_COL_FWD, sp, &y.wide, &var);
    y.f = make_plan(g.L[1], var);
    y.nkb = (g.Lc + y.B - 1) / y.B;
    cols[0] = y;
    col_blocks[0] = m[0] * y.nkb;

    ColPass z{};
    z.op = COL_FWD_OTF;
    z.n_src = m[0];
    z.left = 0;
    z.mode = LUCEM_PAD_CONSTANT;
    z.n_dst = g.L[0];
    z.Lc = g.Lc;
    z.tw = tws[0];
    z.src = pb;
    z.src_outer = g.pitch;
    z.src_axis = (int64_t)g.L[1] * g.pitch;
    z.dst = otf;
    z.dst_outer = g.pitch;
    z.dst_axis = (int64_t)g.L[1] * g.pitch;
    z.scale = scale;
    z.B = choose_B(g.L[0], g.Lc, g.L[1], SHAPE_COL_FWD, sp, &z.wide, &var);
    z.f = make_plan(g.L[0], var);
    z.nkb = (g.Lc + z.B - 1) / z.B;
    cols[1] = z;
    col_blocks[1] = g.L[1] * z.nkb;
    return 2;
}

// Line-pass descriptors of one correlation over spec_a (and spec_b in 3-D) with transfer `otf`:
// 2-D: y fused (in place); 3-D: y forward (a -> b), z fused (in place on b), y inverse (b -> a).
int lucem::make_corr_descs(const Geometry& g, const ShapePolicy& sp, float2* sa, float2* sb, const float2* otf,
                           const float2* const tws[3], ColPass out[3], int blocks[3]) {
    int var = 0;
    ColPass y{};
    y.Lc = g.Lc;
    y.tw = tws[1];
    y.left = g.k[1] / 2 + g.origin[1];
    y.mode = g.mode;
    if (g.ndim == 2) {
        y.op = COL_FUSED;
        y.n_src = g.n[1];
        y.n_dst = g.n[1];
        y.src = sa;
        y.dst = sa;
        y.src_outer = y.dst_outer = 0;
        y.src_axis = y.dst_axis = g.pitch;
        y.otf = otf;
        y.otf_outer = 0;
        y.otf_axis = g.pitch;
        y.B = choose_B(g.L[1], g.Lc, 1, SHAPE_COL_FUSED, sp, &y.wide, &var);
        y.f = make_plan(g.L[1], var);
        y.nkb = (g.Lc + y.B - 1) / y.B;
        y.stat = col_static_choice(var == kVariantStatic, g.L[1], y.B, y.wide);
        out[0] = y;
        blocks[0] = y.nkb;
        return 1;
    }
    y.op = COL_FWD;
    y.n_src = g.n[1];
    y.n_dst = g.L[1];
    y.src = sa;
    y.src_outer = (int64_t)g.n[1] * g.pitch;
    y.src_axis = g.pitch;
    y.dst = sb;
    y.dst_outer = (int64_t)g.L[1] * g.pitch;
    y.dst_axis = g.pitch;
    y.B = choose_B(g.L[1], g.Lc, g.n[0], SHAPE_COL_FWD, sp, &y.wide, &var);
    y.f = make_plan(g.L[1], var);
    y.nkb = (g.Lc + y.B - 1) / y.B;
    y.stat = line_static_choice(var == kVariantStatic, g.L[1], y.B, y.wide);
    out[0] = y;
    blocks[0] = g.n[0] * y.nkb;

    ColPass z{};
    z.op = COL_FUSED;
    z.Lc = g.Lc;
    z.tw = tws[0];
    z.left = g.k[0] / 2 + g.origin[0];
    z.mode = g.mode;
    z.n_src = g.n[0];
    z.n_dst = g.n[0];
    z.src = sb;
    z.dst = sb;
    z.src_outer = z.dst_outer = g.pitch;
    z.src_axis = z.dst_axis = (int64_t)g.L[1] * g.pitch;
    z.otf = otf;
    z.otf_outer = g.pitch;
    z.otf_axis = (int64_t)g.L[1] * g.pitch;
    z.B = choose_B(g.L[0], g.Lc, g.L[1], SHAPE_COL_FUSED, sp, &z.wide, &var);
    z.f = make_plan(g.L[0], var);
    z.nkb = (g.Lc + z.B - 1) / z.B;
    z.stat = col_static_choice(var == kVariantStatic, g.L[0], z.B, z.wide);
    out[1] = z;
    blocks[1] = g.L[1] * z.nkb;

    ColPass yi = y;
    yi.op = COL_INV;
    yi.B = choose_B(g.L[1], g.Lc, g.n[0], SHAPE_COL_INV, sp, &yi.wide, &var);
    yi.f = make_plan(g.L[1], var);
    yi.nkb = (g.Lc + yi.B - 1) / yi.B;
    yi.stat = line_static_choice(var == kVariantStatic, g.L[1], yi.B, yi.wide);
    yi.n_src = g.L[1];
    yi.n_dst = g.n[1];
    yi.src = sb;
    yi.src_outer = (int64_t)g.L[1] * g.pitch;
    yi.dst = sa;
    yi.dst_outer = (int64_t)g.n[1] * g.pitch;
    out[2] = yi;
    blocks[2] = g.n[0] * yi.nkb;
    return 3;
}

RowPass lucem::make_row_desc(const Geometry& g, const ShapePolicy& sp, const float2* tw_x, float eps, const float* obs,
                             float* est, float* out, float2* spec, int* blocks) {
    const int64_t rows = (int64_t)g.n[0] * g.n[1];
    RowPass r{};
    r.nx = g.n[2];
    r.rows = (int)rows;
    r.left = g.k[2] / 2 + g.origin[2];
    r.mode = g.mode;
    r.Lc = g.Lc;
    r.pitch = g.pitch;
    int var = 0;
    r.P = choose_P(g.L[2], (int)rows, sp, &r.wide, &var);
    r.eps = eps;
    r.init = 0.5f;
    r.f = make_plan(g.L[2], var);
    r.stat = row_static_choice(var == kVariantStatic, g.L[2], r.P, r.wide);
    r.obs = obs;
    r.est = est;
    r.out = out;
    r.spec = spec;
    r.tw = tw_x;
    *blocks = (int)((((rows + 1) / 2) + r.P - 1) / r.P);
    return r;
}

static const char* const kCorrNames2[1] = {"line_fused_y"};
static const char* const kCorrNames3[3] = {"line_fwd_y", "line_fused_z", "line_inv_y"};

lucem_status lucem::run_corr(lucem_client* c, const Geometry& g, const ShapePolicy& sp, float2* sa, float2* sb,
                             const float2* otf, const float2* const tws[3]) {
    ColPass d[3];
    int b[3];
    const int n = make_corr_descs(g, sp, sa, sb, otf, tws, d, b);
    for (int i = 0; i < n; i++) {
        lucem_status st = run_col(c, d[i], b[i], g.ndim == 2 ? kCorrNames2[i] : kCorrNames3[i]);
        if (st) return st;
    }
    return LUCEM_OK;
}

lucem_status lucem::run_otf(lucem_client* c, const Geometry& g, const float* w, float2* pa, float2* pb, float2* otf,
                            const float2* const tws[3]) {
    RowPass r;
    int rb;
    ColPass cols[2];
    int cb[2];
    const ShapePolicy sp{c->num_cus, 0, 0, nullptr};
    const int n = make_otf_descs(g, sp, w, pa, pb, otf, tws, &r, &rb, cols, cb);
    lucem_status st = run_row(c, r, rb, "otf_row_load");
    if (st) return st;
    for (int i = 0; i < n; i++)
        if ((st = run_col(c, cols[i], cb[i], i + 1 == n ? "otf_line_fwd_last" : "otf_line_fwd_y"))) return st;
    return LUCEM_OK;
}

float lucem::rl_eps(const lucem_rl_settings& s) {
    return (float)std::max(s.has_filter_epsilon ? s.filter_epsilon : 1.0e-12, 1.0e-15);
}

RlSizes lucem::rl_sizes(const Geometry& g) {
    const size_t c8 = sizeof(float2);
    RlSizes z{};
    z.est = sizeof(float) * (size_t)g.n[0] * g.n[1] * g.n[2];
    z.sa = c8 * (size_t)g.n[0] * g.n[1] * g.pitch;
    z.sb = g.ndim == 3 ? c8 * (size_t)g.n[0] * g.L[1] * g.pitch : 0;
    z.otf = c8 * (size_t)g.L[0] * g.L[1] * g.pitch;
    z.psf = sizeof(float) * (size_t)g.k[0] * g.k[1] * g.k[2];
    z.pa = c8 * (size_t)g.k[0] * g.k[1] * g.pitch;
    z.pb = g.ndim == 3 ? c8 * (size_t)g.k[0] * g.L[1] * g.pitch : 0;
    return z;
}

lucem_status lucem::rl_launch(lucem_client* c, const Geometry& g, const lucem_rl_settings& s, const float* input,
                              const float* psf, float* output, const RlBuffers& b, const float2* const tws[3],
                              const ShapeTable* table) {
    lucem_status st;
    // PSF normalisation and flip (deconvolve.rs:73-90), then the two transfer functions:
    // otf_conv for correlate(est, rev) (:120), otf_mirror for correlate(rel, h) (:133).
    const int64_t nk = (int64_t)g.k[0] * g.k[1] * g.k[2];
    PsfPrep pp{nk, psf, b.h, b.rev};
    if ((st = launch_check(launch_psf_prep(pp, c->stream), "psf prep"))) return st;
    if ((st = run_otf(c, g, b.rev, b.pa, b.pb, b.otf1, tws))) return st;
    if ((st = run_otf(c, g, b.h, b.pa, b.pb, b.otf2, tws))) return st;

    int row_blocks = 0;
    const ShapePolicy sp{c->num_cus, 0, 0, table};
    RowPass r = make_row_desc(g, sp, tws[2], rl_eps(s), input, b.est, output, b.sa, &row_blocks);

    // est_0 = 0.5 everywhere (deconvolve.rs:93): transform the constant directly.
    r.op = ROW_LOAD_CONST;
    if ((st = run_row(c, r, row_blocks, "row_start"))) return st;
    for (uint64_t it = 0; it < s.num_iter; it++) {
        const bool last = it + 1 == s.num_iter;
        if ((st = run_corr(c, g, sp, b.sa, b.sb, b.otf1, tws))) return st;
        r.op = ROW_RATIO;
        if ((st = run_row(c, r, row_blocks, "row_ratio"))) return st;
        if ((st = run_corr(c, g, sp, b.sa, b.sb, b.otf2, tws))) return st;
        r.op = ROW_UPDATE;
        r.first = it == 0;
        r.last = last;
        r.clip = s.clip ? 1 : 0;
        if ((st = run_row(c, r, row_blocks, last ? "row_update_last" : "row_update"))) return st;
    }
    return LUCEM_OK;
}

lucem_status lucem::rl_run(lucem_client* c, lucem_workspace* ws, const float* input, const float* psf, float* output,
                           const Geometry& g, const lucem_rl_settings& s) {
    if (g.ndim == 2) {
        TileGrid tp;
        if (use_tiled(g, c->num_cus, &tp)) return rl_run_tiled(c, ws, input, psf, output, g, s, tp);
    }
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    const RlSizes z = rl_sizes(g);
    RlBuffers b{};
    lucem_status st;
    if ((st = ws_slot(ws, "rl_est", z.est, (void**)&b.est))) return st;
    if ((st = ws_slot(ws, "rl_spec_a", z.sa, (void**)&b.sa))) return st;
    if (g.ndim == 3 && (st = ws_slot(ws, "rl_spec_b", z.sb, (void**)&b.sb))) return st;
    if ((st = ws_slot(ws, "rl_otf_conv", z.otf, (void**)&b.otf1))) return st;
    if ((st = ws_slot(ws, "rl_otf_mirror", z.otf, (void**)&b.otf2))) return st;
    if ((st = ws_slot(ws, "rl_psf_h", z.psf, (void**)&b.h))) return st;
    if ((st = ws_slot(ws, "rl_psf_rev", z.psf, (void**)&b.rev))) return st;
    if ((st = ws_slot(ws, "rl_psf_spec_a", z.pa, (void**)&b.pa))) return st;
    if (g.ndim == 3 && (st = ws_slot(ws, "rl_psf_spec_b", z.pb, (void**)&b.pb))) return st;
    const float2* tws[3] = {nullptr, nullptr, nullptr};
    for (int a = (g.ndim == 2 ? 1 : 0); a < 3; a++)
        if ((st = twiddles_for(c, g.L[a], &tws[a]))) return st;
    // launch shapes measured on this geometry's own buffers (first call of a geometry only)
    const ShapeTable* table = nullptr;
    if ((st = tuned_shapes(c, g, b.est, b.sa, b.sb, b.otf1, tws, &table))) return st;
    return rl_launch(c, g, s, input, psf, output, b, tws, table);
}

// ---------------------------------------------------------------------------------------------
// validation in the reference's order (deconvolve.rs:53-71)
// ---------------------------------------------------------------------------------------------
lucem_status lucem::validate_rl(const void* input, const uint64_t* in_dims, uint32_t in_ndim, const void* psf,
                                const uint64_t* psf_dims, uint32_t psf_ndim, const void* output,
                                const uint64_t* out_dims, uint32_t out_ndim, const lucem_rl_settings* s,
                                uint32_t max_ndim, bool* empty, Geometry* g, size_t elem_size, bool fft) {
    (void)psf;
    *empty = false;
    if (!s) return fail(LUCEM_INVALID_HANDLE, "settings is NULL");
    if ((in_ndim && !in_dims) || (out_ndim && !out_dims) || (psf_ndim && !psf_dims))
        return fail(LUCEM_INVALID_HANDLE, "dims pointer is NULL");
    bool same = in_ndim == out_ndim;
    for (uint32_t i = 0; same && i < in_ndim; i++) same = in_dims[i] == out_dims[i];
    if (!same)
        return fail(LUCEM_SHAPE_MISMATCH, "input shape " + dims_str(in_dims, in_ndim) + " does not match output shape " +
                                              dims_str(out_dims, out_ndim));
    uint64_t n = 1;
    bool overflow = false;
    for (uint32_t i = 0; i < in_ndim; i++) {
        if (in_dims[i] && n > UINT64_MAX / in_dims[i]) overflow = true;
        n *= in_dims[i];
    }
    if (n > 0 && !overflow && input && output) {
        const char* a = (const char*)input;
        const char* b = (const char*)output;
        const uint64_t bytes = n * elem_size;
        if (a < b + bytes && b < a + bytes)
            return fail(LUCEM_ALIASED_BUFFERS, "input and output arrays must not share the same device buffer");
    }
    const std::string expected = max_ndim == 2 ? "2" : "2 or 3";
    if (in_ndim < 2 || in_ndim > max_ndim)
        return fail(LUCEM_INVALID_DIMENSION, "richardson_lucy input expects " + expected + "-D array, got " +
                                                 std::to_string(in_ndim) + "-D shape " + dims_str(in_dims, in_ndim));
    if (psf_ndim < 2 || psf_ndim > max_ndim || (max_ndim > 2 && psf_ndim != in_ndim))
        return fail(LUCEM_INVALID_DIMENSION, "richardson_lucy psf expects " +
                                                 (max_ndim == 2 ? expected : std::to_string(in_ndim)) +
                                                 "-D array, got " + std::to_string(psf_ndim) + "-D shape " +
                                                 dims_str(psf_dims, psf_ndim));
    uint64_t nk = 1;
    for (uint32_t i = 0; i < psf_ndim; i++) nk *= psf_dims[i];
    if (nk == 0) return fail(LUCEM_INVALID_WEIGHTS, "psf: PSF cannot be empty");
    if (s->num_iter == 0) return fail(LUCEM_INVALID_PARAMETER, "invalid parameter `num_iter`: num_iter must be >= 1");
    if (s->mode > 4)
        return fail(LUCEM_UNKNOWN_PAD_MODE, "boundary pad mode id " + std::to_string(s->mode) +
                                                " is not supported (expected 0..4: reflect, constant, nearest, "
                                                "mirror, wrap)");
    if (n == 0) {
        *empty = true;
        return LUCEM_OK;
    }
    if (overflow || n > 0xFFFFFFFFull)
        return fail(LUCEM_TOO_MANY_ELEMENTS, "rl_est: dims " + dims_str(in_dims, in_ndim) +
                                                 " exceed the device element limit of 4294967295");
    int64_t n3[3] = {1, 1, 1}, k3[3] = {1, 1, 1};
    for (uint32_t i = 0; i < in_ndim; i++) n3[3 - in_ndim + i] = (int64_t)in_dims[i];
    for (uint32_t i = 0; i < psf_ndim; i++) k3[3 - psf_ndim + i] = (int64_t)psf_dims[i];
    if (!fft) {  // the direct (f64) form: sizes only
        // the direct kernels index an axis's padded extent n + k - 1 with 32-bit integers
        for (int a = 0; a < 3; a++)
            if (n3[a] + k3[a] - 1 > 0x7FFFFFFF)
                return fail(LUCEM_UNSUPPORTED_SHAPE, "input " + dims_str(in_dims, in_ndim) + " with psf " +
                                                         dims_str(psf_dims, psf_ndim) +
                                                         ": an axis with the PSF is longer than 2147483647");
        std::memset(g, 0, sizeof *g);
        g->ndim = (int)in_ndim;
        for (int a = 0; a < 3; a++) {
            g->n[a] = (int)n3[a];
            g->k[a] = (int)k3[a];
        }
        g->mode = (int)s->mode;
        return LUCEM_OK;
    }
    if (!make_geometry(n3, k3, (int)in_ndim, g))
        return fail(LUCEM_UNSUPPORTED_SHAPE, "input " + dims_str(in_dims, in_ndim) + " with psf " +
                                                 dims_str(psf_dims, psf_ndim) +
                                                 ": an axis needs an FFT longer than 16384 (n + k - 1 per axis)");
    g->mode = (int)s->mode;
    return LUCEM_OK;
}

template <typename T>
static lucem_status rl_run_direct(lucem_client* c, lucem_workspace* ws, const T* input, const T* psf, T* output,
                                  const Geometry& g, const lucem_rl_settings& s);

static lucem_status rl_entry(lucem_client* client, const float* input, const uint64_t* in_dims, uint32_t in_ndim,
                             const float* psf, const uint64_t* psf_dims, uint32_t psf_ndim, float* output,
                             const uint64_t* out_dims, uint32_t out_ndim, const lucem_rl_settings* settings,
                             lucem_workspace* ws, uint32_t max_ndim) {
    // Argument checks come first and touch no device memory, so their order can be tested
    // without a device (and without a client).
    bool empty = false;
    Geometry g;
    lucem_status st = validate_rl(input, in_dims, in_ndim, psf, psf_dims, psf_ndim, output, out_dims, out_ndim,
                                  settings, max_ndim, &empty, &g);
    bool direct = false;
    if (st == LUCEM_UNSUPPORTED_SHAPE && in_ndim == 2) {
        // an axis longer than the FFT passes take: the direct form, in the reference's own order
        st = validate_rl(input, in_dims, in_ndim, psf, psf_dims, psf_ndim, output, out_dims, out_ndim, settings,
                         max_ndim, &empty, &g, sizeof(float), false);
        direct = true;
    }
    if (st != LUCEM_OK || empty) return st;
    if (!client) return fail(LUCEM_INVALID_HANDLE, "client is NULL");
    if (!input || !psf || !output) return fail(LUCEM_INVALID_HANDLE, "input, psf or output pointer is NULL");
    if (!ws) ws = client->internal_ws.get();
    if (ws->client != client) return fail(LUCEM_INVALID_HANDLE, "workspace belongs to another client");
    if (direct) return rl_run_direct<float>(client, ws, input, psf, output, g, *settings);
    return rl_run(client, ws, input, psf, output, g, *settings);
}

extern "C" lucem_status lucem_richardson_lucy_into_f32(lucem_client* client, const float* input,
                                                       const uint64_t* in_dims, uint32_t in_ndim, const float* psf,
                                                       const uint64_t* psf_dims, uint32_t psf_ndim, float* output,
                                                       const uint64_t* out_dims, uint32_t out_ndim,
                                                       const lucem_rl_settings* settings, lucem_workspace* ws) {
    return rl_entry(client, input, in_dims, in_ndim, psf, psf_dims, psf_ndim, output, out_dims, out_ndim, settings,
                    ws, 2);
}

extern "C" lucem_status lucem_richardson_lucy_f32(lucem_client* client, const float* input, const uint64_t* dims,
                                                  uint32_t ndim, const float* psf, const uint64_t* psf_dims,
                                                  uint32_t psf_ndim, float* output,
                                                  const lucem_rl_settings* settings) {
    return rl_entry(client, input, dims, ndim, psf, psf_dims, psf_ndim, output, dims, ndim, settings, nullptr, 2);
}

extern "C" lucem_status lucem_richardson_lucy_nd_into_f32(lucem_client* client, const float* input,
                                                          const uint64_t* in_dims, uint32_t in_ndim,
                                                          const float* psf, const uint64_t* psf_dims,
                                                          uint32_t psf_ndim, float* output, const uint64_t* out_dims,
                                                          uint32_t out_ndim, const lucem_rl_settings* settings,
                                                          lucem_workspace* ws) {
    return rl_entry(client, input, in_dims, in_ndim, psf, psf_dims, psf_ndim, output, out_dims, out_ndim, settings,
                    ws, 3);
}

// ---------------------------------------------------------------------------------------------
// The direct form: the reference iteration (deconvolve.rs:92-147) with direct correlations, in
// FP64 for richardson_lucy_into::<f64> and in FP32 for 2-D f32 arrays whose padded domain is
// longer than the FFT passes take: est_a = 0.5; per iteration conv = correlate(src, rev),
// rel = y / safe(conv), err = correlate(rel, h), dst = src * err (clip on the last), src/dst
// ping-ponging between est_a and est_b and the last iteration writing the output.
// ---------------------------------------------------------------------------------------------
static int psf_prep_direct(const double* psf, double* h, double* rev, int64_t nk, void* stream) {
    return launch_psf_prep_f64(psf, h, rev, nk, stream);
}
static int psf_prep_direct(const float* psf, float* h, float* rev, int64_t nk, void* stream) {
    PsfPrep pp{nk, psf, h, rev};  // f64 sum and norm, h rounded to f32 (deconvolve.rs:73-90)
    return launch_psf_prep(pp, stream);
}

// The launches of the direct form on given buffers: a, b, conv, rel, err of n elements, h and
// rev of the PSF's size.
template <typename T>
static lucem_status rl_direct_launch(lucem_client* c, const T* input, const T* psf, T* output, const Geometry& g,
                                     const lucem_rl_settings& s, T* a, T* b, T* conv, T* rel, T* err, T* h, T* rev) {
    const int H = g.n[1], W = g.n[2], kh = g.k[1], kw = g.k[2];
    const int64_t n = (int64_t)H * W, nk = (int64_t)kh * kw;
    const bool f64 = sizeof(T) == sizeof(double);
    lucem_status st;
    void* sm = c->stream;
    if ((st = launch_check(psf_prep_direct(psf, h, rev, nk, sm), "psf prep (direct)"))) return st;
    if ((st = launch_check(launch_fill_direct(a, T(0.5), n, sm), "fill (direct)"))) return st;
    // the epsilon of deconvolve.rs:104, cast to F
    const T eps = f64 ? (T)std::max(s.has_filter_epsilon ? s.filter_epsilon : 1.0e-12, 1.0e-15) : (T)rl_eps(s);
    bool ping = true;
    for (uint64_t it = 0; it < s.num_iter; it++) {
        const bool last = it + 1 == s.num_iter;
        const T* src = ping ? a : b;
        T* dst = last ? output : (ping ? b : a);
        if ((st = launch_check(launch_correlate2d_direct(src, rev, conv, H, W, kh, kw, kh / 2, kw / 2, g.mode, T(0), sm),
                               "correlate (direct)")) ||
            (st = launch_check(launch_ratio_direct(input, conv, rel, eps, n, sm), "ratio (direct)")) ||
            (st = launch_check(launch_correlate2d_direct(rel, h, err, H, W, kh, kw, kh / 2, kw / 2, g.mode, T(0), sm),
                               "correlate (direct)")) ||
            (st = launch_check(launch_update_direct(src, err, dst, last && s.clip ? 1 : 0, n, sm), "update (direct)")))
            return st;
        ping = !ping;
    }
    return LUCEM_OK;
}

lucem_status lucem::rl_direct_launch_f32(lucem_client* c, const float* input, const float* psf, float* output,
                                         const Geometry& g, const lucem_rl_settings& s, float* const bufs[7]) {
    return rl_direct_launch<float>(c, input, psf, output, g, s, bufs[0], bufs[1], bufs[2], bufs[3], bufs[4], bufs[5],
                                   bufs[6]);
}

template <typename T>
static lucem_status rl_run_direct(lucem_client* c, lucem_workspace* ws, const T* input, const T* psf, T* output,
                                  const Geometry& g, const lucem_rl_settings& s) {
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    const int64_t n = (int64_t)g.n[1] * g.n[2], nk = (int64_t)g.k[1] * g.k[2];
    const bool f64 = sizeof(T) == sizeof(double);
    T *a, *b, *conv, *rel, *err, *h, *rev;
    lucem_status st;
    if ((st = ws_slot(ws, f64 ? "rl64_est_a" : "rl32_est_a", sizeof(T) * n, (void**)&a)) ||
        (st = ws_slot(ws, f64 ? "rl64_est_b" : "rl32_est_b", sizeof(T) * n, (void**)&b)) ||
        (st = ws_slot(ws, f64 ? "rl64_est_conv" : "rl32_est_conv", sizeof(T) * n, (void**)&conv)) ||
        (st = ws_slot(ws, f64 ? "rl64_rel_blur" : "rl32_rel_blur", sizeof(T) * n, (void**)&rel)) ||
        (st = ws_slot(ws, f64 ? "rl64_err_est" : "rl32_err_est", sizeof(T) * n, (void**)&err)) ||
        (st = ws_slot(ws, f64 ? "rl64_psf_h" : "rl32_psf_h", sizeof(T) * nk, (void**)&h)) ||
        (st = ws_slot(ws, f64 ? "rl64_psf_rev" : "rl32_psf_rev", sizeof(T) * nk, (void**)&rev)))
        return st;
    return rl_direct_launch<T>(c, input, psf, output, g, s, a, b, conv, rel, err, h, rev);
}

static lucem_status rl_entry_f64(lucem_client* client, const double* input, const uint64_t* in_dims, uint32_t in_ndim,
                                 const double* psf, const uint64_t* psf_dims, uint32_t psf_ndim, double* output,
                                 const uint64_t* out_dims, uint32_t out_ndim, const lucem_rl_settings* settings,
                                 lucem_workspace* ws) {
    bool empty = false;
    Geometry g;
    lucem_status st = validate_rl(input, in_dims, in_ndim, psf, psf_dims, psf_ndim, output, out_dims, out_ndim,
                                  settings, 2, &empty, &g, sizeof(double), false);
    if (st != LUCEM_OK || empty) return st;
    if (!client) return fail(LUCEM_INVALID_HANDLE, "client is NULL");
    if (!input || !psf || !output) return fail(LUCEM_INVALID_HANDLE, "input, psf or output pointer is NULL");
    if (!ws) ws = client->internal_ws.get();
    if (ws->client != client) return fail(LUCEM_INVALID_HANDLE, "workspace belongs to another client");
    return rl_run_direct<double>(client, ws, input, psf, output, g, *settings);
}

extern "C" lucem_status lucem_richardson_lucy_into_f64(lucem_client* client, const double* input,
                                                       const uint64_t* in_dims, uint32_t in_ndim, const double* psf,
                                                       const uint64_t* psf_dims, uint32_t psf_ndim, double* output,
                                                       const uint64_t* out_dims, uint32_t out_ndim,
                                                       const lucem_rl_settings* settings, lucem_workspace* ws) {
    return rl_entry_f64(client, input, in_dims, in_ndim, psf, psf_dims, psf_ndim, output, out_dims, out_ndim,
                        settings, ws);
}

extern "C" lucem_status lucem_richardson_lucy_f64(lucem_client* client, const double* input, const uint64_t* dims,
                                                  uint32_t ndim, const double* psf, const uint64_t* psf_dims,
                                                  uint32_t psf_ndim, double* output,
                                                  const lucem_rl_settings* settings) {
    return rl_entry_f64(client, input, dims, ndim, psf, psf_dims, psf_ndim, output, dims, ndim, settings, nullptr);
}

// Batched complex line FFTs (contiguous lines of length L, interleaved re/im).
extern "C" lucem_status lucem_fft_c2c_lines_f32(lucem_client* client, const float* src, float* dst, uint32_t L,
                                                uint64_t nlines, int32_t inverse) {
    if (!client) return fail(LUCEM_INVALID_HANDLE, "client is NULL");
    if (L == 0 || nlines == 0) return LUCEM_OK;
    if (L > LUCEM_MAX_FFT_LEN || !smooth7((int)L))
        return fail(LUCEM_UNSUPPORTED_SHAPE, "FFT length must be 7-smooth and at most 16384");
    if (nlines > 0x7FFFFFFFull) return fail(LUCEM_TOO_MANY_ELEMENTS, "too many lines");
    hipError_t e = hipSetDevice(client->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    const float2* tw;
    lucem_status st = twiddles_for(client, (int)L, &tw);
    if (st) return st;
    ColPass p{};
    p.op = inverse ? COL_INV : COL_FWD;
    p.n_src = (int)L;
    p.left = 0;
    p.mode = LUCEM_PAD_WRAP;
    p.n_dst = (int)L;
    p.Lc = 1;
    p.B = 1;
    p.wide = tier_for_length((int)L);
    p.nkb = 1;
    p.src_outer = L;
    p.src_axis = 1;
    p.dst_outer = L;
    p.dst_axis = 1;
    p.f = make_plan((int)L);
    p.tw = tw;
    p.src = (const float2*)src;
    p.dst = (float2*)dst;
    return run_col(client, p, (int)nlines, "fft_lines");
}
