// Host-side internals shared by the C ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/lucem_hip.h"
#include "rl_internal.hpp"

struct lucem_workspace;

namespace lucem {
// One profiled launch: events on the client stream around it and its algorithmic HBM bytes.
struct ProfRecord {
    const char* name;
    hipEvent_t start, stop;
    double bytes;
};
}  // namespace lucem

struct lucem_client {
    int device = 0;
    int num_cus = 256;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::map<int, float2*> twiddles;               // FFT length -> device twiddle table
    std::unique_ptr<lucem_workspace> internal_ws;  // used by the allocating-style entry points
    bool profiling = false;
    std::vector<lucem::ProfRecord> records;        // launches since profiling was enabled
    std::vector<hipEvent_t> event_pool;            // events of read records, reused
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
    std::map<std::string, lucem::Slot> slots;
    ~lucem_workspace();
};

namespace lucem {

// Geometry of one RL array: sizes per axis (z, y, x), FFT lengths and half-spectrum layout.
struct Geometry {
    int ndim;
    int n[3];
    int k[3];
    int L[3];
    int Lc;
    int pitch;
    int mode;
};

lucem_status fail(lucem_status st, const std::string& msg);
lucem_status hip_fail(int err, const char* what);
std::string dims_str(const uint64_t* dims, uint32_t ndim);
lucem_status twiddles_for(lucem_client* c, int L, const float2** out);
lucem_status ws_slot(lucem_workspace* ws, const char* name, size_t bytes, void** out);

int choose_fft_len(int target);
FftPlan make_plan(int L);
int choose_lgP(int L, int nx, int rows, int num_cus);
int choose_lgB(int L, int Lc, int64_t outers, int num_cus, int* wide);
bool make_geometry(const int64_t n[3], const int64_t k[3], int ndim, Geometry* g);

// Launch with optional profiling (events around the launch, algorithmic bytes recorded).
lucem_status run_row(lucem_client* c, const RowPass& r, int blocks, const char* name);
lucem_status run_col(lucem_client* c, const ColPass& p, int blocks, const char* name);
double row_pass_bytes(const RowPass& r);
double col_pass_bytes(const ColPass& p, int blocks);

// Descriptor builders shared by the plain and batched drivers.
int make_otf_descs(const Geometry& g, int num_cus, const float* w, float2* pa, float2* pb, float2* otf,
                   const float2* const tws[3], RowPass* row, int* row_blocks, ColPass cols[2], int col_blocks[2]);
int make_corr_descs(const Geometry& g, int num_cus, float2* sa, float2* sb, const float2* otf,
                    const float2* const tws[3], ColPass out[3], int blocks[3]);
RowPass make_row_desc(const Geometry& g, int num_cus, const float2* tw_x, float eps, const float* obs, float* est,
                      float* out, float2* spec, int* blocks);
float rl_eps(const lucem_rl_settings& s);

lucem_status validate_rl(const float* input, const uint64_t* in_dims, uint32_t in_ndim, const float* psf,
                         const uint64_t* psf_dims, uint32_t psf_ndim, const float* output, const uint64_t* out_dims,
                         uint32_t out_ndim, const lucem_rl_settings* s, uint32_t max_ndim, bool* empty, Geometry* g);
lucem_status rl_run(lucem_client* c, lucem_workspace* ws, const float* input, const float* psf, float* output,
                    const Geometry& g, const lucem_rl_settings& s);

}  // namespace lucem
