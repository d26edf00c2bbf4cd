// Descriptors of the differentiable Richardson-Lucy passes (shared by rl_grad.hip and lucem_grad.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "rl_internal.hpp"

namespace lucem {

enum GradRowOp : int {
    G_LOAD_X = 0,  // x_{sk} (checkpoint or caller's x0) -> slot 0, pad gather, R2C
    G_RATIO = 1,   // C2R, u_t = crop -> slot, y / safe(u_t), pad gather, R2C
    G_UPDATE = 2,  // C2R, c_t -> slot, x_{t+1} = x_t c_t -> slot (and checkpoint / output), R2C
    G_LOSS = 3,    // C2R, u_T -> uT and the loss partials (no forward transform)
    G_ADJ_R = 4,   // C2R, fold, rbar -> ubar = -rbar y / u_t^2, crop-transpose gather, R2C
    G_ADJ_X = 5,   // C2R, fold, xbar_t = A^T ubar + xa; next: xa, cbar (R2C) or store xbar
    G_SEED = 6,    // xbar (segment boundary) -> xa, cbar, crop-transpose gather, R2C
    G_C2R = 7,     // C2R, crop, scale -> out (PSF-gradient spectra to real)
    G_LOSS_ADJ = 8,// ubar_T = w (1 - y / u_T), crop-transpose gather, R2C (start of the backward)
};

enum GradFlags : int {
    GF_SAVE = 1,       // store u_t / c_t in the segment slots
    GF_CKPT = 2,       // also store x_{t+1} as checkpoint `ck`
    GF_LAST = 4,       // x_{t+1} is x_T: copy it to the caller's output
    GF_FROM_X0 = 8,    // G_LOAD_X reads the caller's x0
    GF_NO_XA = 16,     // G_ADJ_X: no xa term (the loss step)
    GF_SEED_NEXT = 32, // G_ADJ_X: continue into the previous iteration (slot j - 1 is resident)
    GF_TO_GX0 = 64,    // G_ADJ_X: the result is d loss / d x0
};

struct GradTile {
    RowPass r;             // geometry, spectrum, twiddles, obs = y, eps
    int kx;                // PSF extent along x (fold support)
    float scale;           // G_C2R
    int64_t nstride;       // floats per slot
    float* xslots;         // x_{sk + j}, j = 0 .. k
    float* uslots;         // u_{sk + j}
    float* cslots;         // c_{sk + j}
    float* ckpt;           // checkpoints x_{sk}, s = 1 .. S-1 (slot s)
    float* xa;
    float* xbar;
    const float* x0;       // the caller's initial estimate
    float* x_out;          // the caller's final estimate (or null)
    float* gx0;            // the caller's d loss / d x0
    const float* weight;   // d total / d loss of this tile (one float) or null for 1
    double* partial;       // loss partials, one per row block
    float* uT;             // u_T = A x_T
    float* out;            // G_C2R output
    const float* target;   // the counts the loss compares u_T with (null: y, the RL input)
};

struct GradRowBatch {
    const GradTile* tiles;
    const int2* items;
    int nitems;
    int op;
    int sx;     // slot of x_t (x slots are used modulo k + 1)
    int sx1;    // slot of x_{t+1}
    int su;     // slot of u_t, c_t (modulo k)
    int sxm;    // slot of x_{t-1}
    int scm;    // slot of c_{t-1}
    int ck;     // checkpoint index
    int flags;
    int wide;   // thread tier of the launch: 0 = 256, 1 = 512, 2 = 1024 threads
};

int launch_grad_row(const GradRowBatch& b, size_t lds, void* stream);
// Allows the 512- and 1024-thread gradient row kernels their dynamic LDS (once per process).
int allow_big_lds_grad_kernels();
int launch_loss_reduce(const double* partials, const int* first, const int* count, double* loss, int ntiles,
                       void* stream);
int launch_psf_grad(const float* gh, const float* grev, int ntiles, int64_t nk, const float* psf, float* out,
                    void* stream);
int launch_zero_floats(float* p, int64_t n, void* stream);

}  // namespace lucem
