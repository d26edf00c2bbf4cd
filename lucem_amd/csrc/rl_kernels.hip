// HIP kernels of the Richardson-Lucy deconvolution (gfx950).
//
// One RL iteration of the reference (restore/deconvolve.rs:109-147) is
//   est_conv = correlate(est, rev)        filter/kernels/convolve.rs:61-118
//   rel      = y / safe(est_conv)         restore/kernels/deconvolve.rs:7-22
//   err      = correlate(rel, h)
//   est'     = est * err (clip on last)   restore/kernels/deconvolve.rs:24-43
// Here each correlation is an FFT linear correlation of the padded array (see rl_internal.hpp)
// and the pointwise steps are fused into the x-axis passes:
//   row pass (x): C2R of the previous correlation, crop, ratio or update, pad gather, R2C;
//   line passes (y, z): FFT along the axis with the pad gather, multiply by the transfer
//   function, inverse FFT, crop.
// All arithmetic is FP32; the path is bound by HBM bandwidth, not by MFMA.
#include <hip/hip_runtime.h>

#include "fft_lds.hpp"
#include "fft_static.hpp"
#include "rl_internal.hpp"

// Waves per SIMD the FFT kernels are compiled for (their VGPR budget: 4 -> 128, 5 -> 96). The
// one-array kernels run 5 (measured faster despite a few spilled registers); the persistent
// batch kernels run 4 (at 5 they spill in the work loop and were slower).
#ifndef LUCEM_WAVES_PER_EU
#define LUCEM_WAVES_PER_EU 5
#endif
#ifndef LUCEM_BATCH_WAVES_PER_EU
#define LUCEM_BATCH_WAVES_PER_EU 4
#endif

namespace lucem {

// ---------------------------------------------------------------------------------------------
// Row pass. Workgroup g owns row pairs [g*P, g*P + P); pair p is rows (2p, 2p + 1).
// LDS: Z[L * P] (interleaved complex lines, element i of pair b at Z[i*P + b]).
// ---------------------------------------------------------------------------------------------
// Per-launch flags of a row pass (the iteration it belongs to).
struct RowFlags {
    int first, last, clip;
};

// Offset of tile row r in a strided view with row stride rs and plane stride ps.
__device__ __forceinline__ int64_t view_row(const RowPass& d, int r, int64_t rs, int64_t ps) {
    if (d.ny >= d.rows) return (int64_t)r * rs;
    const int z = r / d.ny;
    return (int64_t)z * ps + (int64_t)(r - z * d.ny) * rs;
}

// VIEWS: honour the strided views of the descriptor (the batched form); the one-array kernels
// are compiled without them.
template <int T, int E, int OP, int W, bool VIEWS = false>
__device__ void row_pass_body(const RowPass& d, const RowFlags fl, int g, float2* Z) {
    constexpr int S = E;  // elements per thread in the gather (L * P <= T * E)
    const int P = d.P;
    const Divisor pd = Divisor::of(P);
    const int L = d.f.L;
    const int nx = d.nx;
    const int tid = threadIdx.x;
    const int pair0 = g * P;
    const int total = L * P;
    const Divisor ld = Divisor::of(L);
    const float2* twl = stage_twiddles(Z, total, d.f, d.tw);

    if constexpr (OP == ROW_RATIO || OP == ROW_UPDATE) {
        // Z[k] = A[k] + i B[k] for k < Lc, and the Hermitian mirror for k >= Lc. All loads are
        // issued before the first LDS write (item e = b * Lc + k, P * Lc <= T * KS).
        constexpr int KS = (E + 1) / 2 + 1;
        const int Lc = d.Lc;
        const int nload = Lc * P;
        const float inv_lc = 1.0f / (float)Lc;
        float2 av[KS], bv[KS];
        int kk[KS], bb[KS];
#pragma unroll
        for (int s = 0; s < KS; s++) {
            const int e = tid + s * T;
            int b = __float2int_rz((float)e * inv_lc);
            int k = e - b * Lc;
            if (k >= Lc) { k -= Lc; b++; }
            if (k < 0) { k += Lc; b--; }
            kk[s] = k;
            bb[s] = b;
            av[s] = make_float2(0.f, 0.f);
            bv[s] = make_float2(0.f, 0.f);
            if (e < nload) {
                const int ra = 2 * (pair0 + b), rb = ra + 1;
                if (ra < d.rows) av[s] = d.spec[(int64_t)ra * d.pitch + k];
                if (rb < d.rows) bv[s] = d.spec[(int64_t)rb * d.pitch + k];
            }
        }
#pragma unroll
        for (int s = 0; s < KS; s++) {
            const int e = tid + s * T;
            if (e < nload) {
                const int k = kk[s], b = bb[s];
                float2 A = av[s], Bv = bv[s];
                if (k == 0 || 2 * k == L) {
                    A.y = 0.f;
                    Bv.y = 0.f;
                }
                Z[k * P + b] = make_float2(A.x - Bv.y, A.y + Bv.x);
                const int km = L - k;
                if (k > 0 && km >= Lc) Z[km * P + b] = make_float2(A.x + Bv.y, Bv.x - A.y);
            }
        }
        __syncthreads();
        lds_fft<true, T, E, W>(Z, pd, d.f, d.tw, twl);
        // Gather into registers: padded position i reads the cropped result at src = pad(i - left)
        // and applies the pointwise step there. Every value is read before any is overwritten.
        // Item e = b * L + i runs along the row (x fastest across lanes), so the global reads and
        // writes of the real rows are contiguous; the LDS position of (i, b) is i * P + b.
        float2 val[S];
        int srcs[S];
        // global operands first (observed image or previous estimate at the source column)
#pragma unroll
        for (int s = 0; s < S; s++) {
            const int e = tid + s * T;
            srcs[s] = -1;
            val[s] = make_float2(0.f, 0.f);
            if (e < total) {
                const int b = ld.div(e), i = e - b * L;
                const int src = pad_index(i - d.left, nx, d.mode);
                srcs[s] = src;
                if (src >= 0) {
                    const int ra = 2 * (pair0 + b), rb = ra + 1;
                    if constexpr (OP == ROW_RATIO) {
                        if (VIEWS && d.strided_obs) {
                            if (ra < d.rows) val[s].x = d.obs[view_row(d, ra, d.obs_rs, d.obs_ps) + src];
                            if (rb < d.rows) val[s].y = d.obs[view_row(d, rb, d.obs_rs, d.obs_ps) + src];
                        } else {
                            if (ra < d.rows) val[s].x = d.obs[(int64_t)ra * nx + src];
                            if (rb < d.rows) val[s].y = d.obs[(int64_t)rb * nx + src];
                        }
                    } else {
                        if (!fl.first) {
                            if (ra < d.rows) val[s].x = d.est[(int64_t)ra * nx + src];
                            if (rb < d.rows) val[s].y = d.est[(int64_t)rb * nx + src];
                        } else {
                            val[s] = make_float2(d.init, d.init);
                        }
                    }
                }
            }
        }
        // then the pointwise step with the correlation read from LDS
#pragma unroll
        for (int s = 0; s < S; s++) {
            const int e = tid + s * T;
            const int src = srcs[s];
            if (e < total && src >= 0) {
                const int b = ld.div(e);
                const float2 c = Z[src * P + b];
                if constexpr (OP == ROW_RATIO) {
                    const float da = fabsf(c.x) < d.eps ? d.eps : c.x;
                    const float db = fabsf(c.y) < d.eps ? d.eps : c.y;
                    val[s] = make_float2(div_fast(val[s].x, da), div_fast(val[s].y, db));
                } else {
                    float oa = val[s].x * c.x, ob = val[s].y * c.y;
                    if (fl.last && fl.clip) {
                        oa = fminf(fmaxf(oa, -1.f), 1.f);
                        ob = fminf(fmaxf(ob, -1.f), 1.f);
                    }
                    val[s] = make_float2(oa, ob);
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < S; s++) {
            const int e = tid + s * T;
            if (e < total) {
                const int b = ld.div(e), i = e - b * L;
                if constexpr (OP == ROW_UPDATE) {
                    if (srcs[s] >= 0 && i == srcs[s] + d.left) {
                        const int x = i - d.left;
                        const int ra = 2 * (pair0 + b), rb = ra + 1;
                        if (VIEWS && fl.last && d.window) {
                            if (x >= d.own0[2] && x < d.own1[2]) {
                                const int rr[2] = {ra, rb};
                                const float vv[2] = {val[s].x, val[s].y};
#pragma unroll
                                for (int q = 0; q < 2; q++) {
                                    const int r = rr[q];
                                    if (r >= d.rows) continue;
                                    const int z = d.ny >= d.rows ? 0 : r / d.ny;
                                    const int y = r - z * d.ny;
                                    if (y >= d.own0[1] && y < d.own1[1] && z >= d.own0[0] && z < d.own1[0])
                                        d.out[(int64_t)(z - d.own0[0]) * d.out_ps + (int64_t)(y - d.own0[1]) * d.out_rs +
                                              (x - d.own0[2])] = vv[q];
                                }
                            }
                        } else {
                            float* dst = fl.last ? d.out : d.est;
                            if (ra < d.rows) dst[(int64_t)ra * nx + x] = val[s].x;
                            if (rb < d.rows) dst[(int64_t)rb * nx + x] = val[s].y;
                        }
                    }
                }
                Z[i * P + b] = val[s];
            }
        }
        if constexpr (OP == ROW_UPDATE) {
            if (fl.last) return;  // uniform: the whole workgroup leaves
        }
    } else {
        // pad gather straight from global memory (or the constant start value)
        float2 z[S];
#pragma unroll
        for (int s = 0; s < S; s++) {
            const int e = tid + s * T;
            z[s] = make_float2(0.f, 0.f);
            if (e < total) {
                const int b = ld.div(e), i = e - b * L;  // x fastest across lanes
                const int ra = 2 * (pair0 + b), rb = ra + 1;
                const bool va = ra < d.rows, vb = rb < d.rows;
                const int src = pad_index(i - d.left, nx, d.mode);
                if (src >= 0) {
                    if constexpr (OP == ROW_LOAD_CONST) {
                        z[s] = make_float2(va ? d.init : 0.f, vb ? d.init : 0.f);
                    } else {
                        if (va) z[s].x = d.src[(int64_t)ra * nx + src];
                        if (vb) z[s].y = d.src[(int64_t)rb * nx + src];
                    }
                }
            }
        }
#pragma unroll
        for (int s = 0; s < S; s++) {
            const int e = tid + s * T;
            if (e < total) {
                const int b = ld.div(e), i = e - b * L;
                Z[i * P + b] = z[s];
            }
        }
    }
    __syncthreads();
    lds_fft<false, T, E, W>(Z, pd, d.f, d.tw, twl);
    // split the two real transforms: A = Z[k] + conj(Z[L-k]), B = -i (Z[k] - conj(Z[L-k]))
    // (both carry a factor 2; the transfer function is scaled to remove it)
    // item e = b * Lc + k over all pairs, so that every thread has work whatever P and Lc are
    const Divisor lcd = Divisor::of(d.Lc);
    const int nsplit = P * d.Lc;
    for (int e = tid; e < nsplit; e += T) {
        const int b = lcd.div(e), k = e - b * d.Lc;
        const int ra = 2 * (pair0 + b), rb = ra + 1;
        const float2 zk = Z[k * P + b];
        const float2 zm = Z[(k == 0 ? 0 : L - k) * P + b];
        if (ra < d.rows) d.spec[(int64_t)ra * d.pitch + k] = make_float2(zk.x + zm.x, zk.y - zm.y);
        if (rb < d.rows) d.spec[(int64_t)rb * d.pitch + k] = make_float2(zk.y + zm.y, zm.x - zk.x);
    }
}

template <int T, int E, int OP>
__global__ __launch_bounds__(T) __attribute__((amdgpu_waves_per_eu(LUCEM_WAVES_PER_EU))) void row_pass_kernel(RowPass d) {
    extern __shared__ float2 lds[];
    row_pass_body<T, E, OP, LUCEM_WAVES_PER_EU>(d, RowFlags{d.first, d.last, d.clip}, blockIdx.x, lds);
}

// ---------------------------------------------------------------------------------------------
// Line pass along y or z. Workgroup g owns columns [kb*B, kb*B + B) of outer index `outer`.
// ---------------------------------------------------------------------------------------------
// Folds the padded positions of B interleaved lines back onto [0, n) in place (the transpose of
// the pad gather Q[i] = x[pad(i - left)]): every halo position h < left or h >= left + n (below
// n + k - 1 <= L) is added to position left + pad(h - left). One thread per line, halo in order.
__device__ __forceinline__ void fold_lines(float2* Z, int B, int nlines_valid, int L, int left, int n, int mode,
                                           int nhalo_end) {
    const int b = threadIdx.x;
    if (b < B && b < nlines_valid) {
        // the halo only: [0, left), then [left + n, nhalo_end), in increasing order
        const int right = left + n;
        for (int h = left > 0 ? 0 : right; h < nhalo_end; h = h + 1 == left ? right : h + 1) {
            const int t = pad_index(h - left, n, mode);
            if (t < 0) continue;
            const float2 v = Z[h * B + b];
            float2& dst = Z[(t + left) * B + b];
            dst = make_float2(dst.x + v.x, dst.y + v.y);
        }
    }
    (void)L;
}

template <int T, int E, int OP, int W>
__device__ void col_pass_body(const ColPass& d, int g, float2* Z, int slot = 0) {
    const int B = d.B;
    const Divisor bd = Divisor::of(B);
    const int outer = g / d.nkb;
    const int kx0 = (g - outer * d.nkb) * B;
    const int L = d.f.L;
    const int tid = threadIdx.x;
    const float2* src = d.src + outer * d.src_outer + kx0;
    const int total = L * B;
    const float2* twl = stage_twiddles(Z, total, d.f, d.tw);
    const bool full = kx0 + B <= d.Lc;
    constexpr bool INV_IN = OP == COL_INV || OP == COL_INV_FOLD;
    {
        float2 v[E];
#pragma unroll
        for (int s = 0; s < E; s++) {
            const int e = tid + s * T;
            v[s] = make_float2(0.f, 0.f);
            if (e < total) {
                const int i = bd.div(e), b = e - i * B;
                const int sr = INV_IN ? i : pad_index(i - d.left, d.n_src, d.mode);
                if (sr >= 0 && (full || kx0 + b < d.Lc)) v[s] = src[sr * d.src_axis + b];
            }
        }
#pragma unroll
        for (int s = 0; s < E; s++) {
            const int e = tid + s * T;
            if (e < total) Z[e] = v[s];
        }
    }
    __syncthreads();
    if constexpr (INV_IN) {
        lds_fft<true, T, E, W>(Z, bd, d.f, d.tw, twl);
    } else {
        lds_fft<false, T, E, W>(Z, bd, d.f, d.tw, twl);
        if constexpr (OP == COL_FUSED || OP == COL_FUSED_T) {
            const int64_t off = outer * d.otf_outer + kx0;
            const float2* otf = d.otf + off;
            float2 w[E];
#pragma unroll
            for (int s = 0; s < E; s++) {
                const int e = tid + s * T;
                w[s] = make_float2(0.f, 0.f);
                if (e < total) {
                    const int i = bd.div(e), b = e - i * B;
                    if (full || kx0 + b < d.Lc) w[s] = otf[i * d.otf_axis + b];
                }
            }
            if constexpr (OP == COL_FUSED) {
                if (d.save) {  // the padded input's spectrum, for the gradient of the PSF
                    float2* sv = d.save + slot * d.slot_stride + off;
#pragma unroll
                    for (int s = 0; s < E; s++) {
                        const int e = tid + s * T;
                        if (e < total) {
                            const int i = bd.div(e), b = e - i * B;
                            if (full || kx0 + b < d.Lc) sv[i * d.otf_axis + b] = Z[e];
                        }
                    }
                }
            } else {
                if (d.gacc) {  // gacc += saved * conj(this spectrum)
                    const float2* sv = d.save + slot * d.slot_stride + off;
                    float2* ga = d.gacc + off;
#pragma unroll
                    for (int s = 0; s < E; s++) {
                        const int e = tid + s * T;
                        if (e < total) {
                            const int i = bd.div(e), b = e - i * B;
                            if (full || kx0 + b < d.Lc) {
                                const float2 a = sv[i * d.otf_axis + b], z = Z[e];
                                float2 acc = ga[i * d.otf_axis + b];
                                acc.x += a.x * z.x + a.y * z.y;
                                acc.y += a.y * z.x - a.x * z.y;
                                ga[i * d.otf_axis + b] = acc;
                            }
                        }
                    }
                }
            }
#pragma unroll
            for (int s = 0; s < E; s++) {
                const int e = tid + s * T;
                if (e < total) {
                    if constexpr (OP == COL_FUSED_T) {
                        Z[e] = cmul(Z[e], make_float2(w[s].x, -w[s].y));
                    } else {
                        Z[e] = cmul(Z[e], w[s]);
                    }
                }
            }
            __syncthreads();
            lds_fft<true, T, E, W>(Z, bd, d.f, d.tw, twl);
        }
    }
    int row0 = 0;
    if constexpr (OP == COL_FUSED_T || OP == COL_INV_FOLD) {
        fold_lines(Z, B, d.Lc - kx0, L, d.fold_left, d.fold_n, d.fold_mode, min(L, d.fold_end));
        __syncthreads();
        row0 = d.fold_left;
    }
    float2* dst = d.dst + outer * d.dst_outer + kx0;
    const int nstore = d.n_dst * B;
    for (int e = tid; e < nstore; e += T) {
        const int i = bd.div(e), b = e - i * B;
        if (full || kx0 + b < d.Lc) {
            float2 v = Z[e + row0 * B];
            if constexpr (OP == COL_FWD_OTF) v = make_float2(v.x * d.scale, -v.y * d.scale);
            dst[i * d.dst_axis + b] = v;
        }
    }
}

template <int T, int E, int OP>
__global__ __launch_bounds__(T) __attribute__((amdgpu_waves_per_eu(LUCEM_WAVES_PER_EU))) void col_pass_kernel(ColPass d) {
    extern __shared__ float2 lds[];
    col_pass_body<T, E, OP, LUCEM_WAVES_PER_EU>(d, xcd_block(blockIdx.x, gridDim.x), lds);
}

// ---------------------------------------------------------------------------------------------
// PSF preparation: one workgroup. The f64 sum is reduced in a fixed order (per-thread strided
// partial sums, then a tree), so it is deterministic from call to call.
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void psf_prep_kernel(PsfPrep d) {
    __shared__ double part[256];
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < d.nk; i += 256) s += (double)d.psf[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    const double sum = part[0];
    const double norm = fabs(sum) > 1e-12 ? 1.0 / sum : 1.0;
    for (int64_t i = threadIdx.x; i < d.nk; i += 256) {
        const float v = (float)((double)d.psf[i] * norm);
        d.h[i] = v;
        d.rev[d.nk - 1 - i] = v;
    }
}

// ---------------------------------------------------------------------------------------------
// Batched persistent kernels: workgroups pull (tile, block) items from an atomic counter.
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ int next_item(int* counter, int* slot) {
    if (threadIdx.x == 0) *slot = atomicAdd(counter, 1);
    __syncthreads();
    const int it = __builtin_amdgcn_readfirstlane(*slot);
    __syncthreads();
    return it;
}

template <int T, int E, int OP>
__global__ __launch_bounds__(T) __attribute__((amdgpu_waves_per_eu(LUCEM_BATCH_WAVES_PER_EU))) void row_batch_kernel(RowBatch bt) {
    extern __shared__ float2 lds[];
    __shared__ int slot;
    for (int round = 0;; round++) {
        // items from the atomic queue, or a static stride over them (persistent grid), or one
        // item per block (direct)
        const int it = bt.counter  ? next_item(bt.counter, &slot)
                       : bt.stride ? xcd_block(blockIdx.x, gridDim.x) + round * (int)gridDim.x
                                   : (round == 0 ? xcd_block(blockIdx.x, gridDim.x) : bt.nitems);
        if (it >= bt.nitems) break;
        const int2 item = bt.items[it];
        const RowPass d = bt.tiles[__builtin_amdgcn_readfirstlane(item.x)];  // a private copy: no aliasing
        row_pass_body<T, E, OP, LUCEM_BATCH_WAVES_PER_EU, true>(d, RowFlags{bt.first, bt.last, bt.clip}, __builtin_amdgcn_readfirstlane(item.y), lds);
        __syncthreads();
    }
}

template <int T, int E, int OP>
__global__ __launch_bounds__(T) __attribute__((amdgpu_waves_per_eu(LUCEM_BATCH_WAVES_PER_EU))) void col_batch_kernel(ColBatch bt) {
    extern __shared__ float2 lds[];
    __shared__ int slot;
    for (int round = 0;; round++) {
        const int it = bt.counter  ? next_item(bt.counter, &slot)
                       : bt.stride ? xcd_block(blockIdx.x, gridDim.x) + round * (int)gridDim.x
                                   : (round == 0 ? xcd_block(blockIdx.x, gridDim.x) : bt.nitems);
        if (it >= bt.nitems) break;
        const int2 item = bt.items[it];
        const ColPass d = bt.tiles[__builtin_amdgcn_readfirstlane(item.x)];
        col_pass_body<T, E, OP, LUCEM_BATCH_WAVES_PER_EU>(d, __builtin_amdgcn_readfirstlane(item.y), lds, bt.slot);
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void psf_batch_kernel(PsfBatch bt) {
    __shared__ double part[256];
    const PsfPrep d = bt.psfs[blockIdx.x];
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < d.nk; i += 256) s += (double)d.psf[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    const double sum = part[0];
    const double norm = fabs(sum) > 1e-12 ? 1.0 / sum : 1.0;
    for (int64_t i = threadIdx.x; i < d.nk; i += 256) {
        const float v = (float)((double)d.psf[i] * norm);
        d.h[i] = v;
        d.rev[d.nk - 1 - i] = v;
    }
}

__global__ void zero_ints_kernel(int* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// ---------------------------------------------------------------------------------------------
// COL_FUSED with the line length L and the column count B known at compile time (fft_static.hpp).
// The passes are inlined, so the transfer-function factors of the last forward pass, loaded with
// the gather, arrive while the forward transform runs (a call to an out-of-line pass would wait
// for them), and multiply in that pass. LDS: two buffers of L x B values (out-of-place passes, one
// barrier each), then the twiddle table. Same descriptor, grid and results (to rounding) as
// col_pass_kernel<256, 16, COL_FUSED>; no PSF-gradient save.
// ---------------------------------------------------------------------------------------------
template <int L, int B>
constexpr size_t col_static_lds_bytes() {
    return (size_t)(2 * tw_lds_offset(L * B) + L) * sizeof(float2);
}

template <int L, int B>
struct ColStatic {
    static constexpr int T = kThreads;
    static constexpr int n = L * B;
    static constexpr int S = (n + T - 1) / T;                   // gathered values per thread
    static constexpr int RL = static_last_radix(L);             // radix of the last pass
    static constexpr int nbf = (L / RL) * B;                    // butterflies of the last pass
    static_assert(n <= T * kElemsPerThread, "one block of lines per workgroup");
};

// Block g's padded lines (v) and the factors of the first round of its last forward pass (pre):
// output t of butterfly (j, b) = threadIdx.x is element j + t L / RL of column b.
template <int L, int B>
__device__ __forceinline__ void col_static_load(const ColPass& d, int g, float2 (&v)[ColStatic<L, B>::S],
                                                float2 (&pre)[ColStatic<L, B>::RL]) {
    using K = ColStatic<L, B>;
    const int outer = g / d.nkb;
    const int kx0 = (g - outer * d.nkb) * B;
    const int nl = min(B, d.Lc - kx0);
    const int tid = threadIdx.x;
    const float2* src = d.src + outer * d.src_outer + kx0;
    const float2* otf = d.otf + outer * d.otf_outer + kx0;
#pragma unroll
    for (int s = 0; s < K::S; s++) {
        const int e = tid + s * K::T;
        v[s] = make_float2(0.f, 0.f);
        if (K::n % K::T == 0 || e < K::n) {
            const int i = e / B, b = e - i * B;
            const int sr = pad_index(i - d.left, d.n_src, d.mode);
            if (sr >= 0 && b < nl) v[s] = src[sr * d.src_axis + b];
        }
    }
    const int j = tid / B, b = tid - j * B;
#pragma unroll
    for (int t = 0; t < K::RL; t++) {
        pre[t] = make_float2(0.f, 0.f);
        if ((K::nbf >= K::T || tid < K::nbf) && b < nl) pre[t] = otf[(int64_t)(j + t * (L / K::RL)) * d.otf_axis + b];
    }
}

// Block g from its loaded values: into X, forward transform with the factors in its last pass,
// inverse transform, store. Ends with the store; the caller synchronises before X is written again.
template <int L, int B>
__device__ __forceinline__ void col_static_run(const ColPass& d, int g, const float2 (&v)[ColStatic<L, B>::S],
                                               const float2 (&pre)[ColStatic<L, B>::RL], float2* X, float2* Y,
                                               const float2* tw) {
    using K = ColStatic<L, B>;
    const int outer = g / d.nkb;
    const int kx0 = (g - outer * d.nkb) * B;
    const int nl = min(B, d.Lc - kx0);
    const int tid = threadIdx.x;
#pragma unroll
    for (int s = 0; s < K::S; s++) {
        const int e = tid + s * K::T;
        if (K::n % K::T == 0 || e < K::n) X[e] = v[s];
    }
    __syncthreads();
    const OutScale ms{d.otf + outer * d.otf_outer + kx0, (int)d.otf_axis, nl};
    float2* sp = static_fft<L, B, K::T, false, false, true>(X, Y, tw, Divisor{B, 0u}, ms, pre);
    const float2* r = static_fft<L, B, K::T, true>(sp, sp == X ? Y : X, tw);
    float2* dst = d.dst + outer * d.dst_outer + kx0;
    for (int e = tid; e < d.n_dst * B; e += K::T) {
        const int i = e / B, b = e - i * B;
        if (b < nl) dst[i * d.dst_axis + b] = r[e];
    }
}

template <int L, int B>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(LUCEM_WAVES_PER_EU))) void col_fused_static_kernel(
    ColPass d) {
    using K = ColStatic<L, B>;
    extern __shared__ float2 lds[];
    float2* X = lds;
    float2* Y = lds + tw_lds_offset(K::n);
    float2* tw = Y + tw_lds_offset(K::n);
    const int g = xcd_block(blockIdx.x, gridDim.x);
    float2 v[K::S], pre[K::RL];
    col_static_load<L, B>(d, g, v, pre);
    for (int i = threadIdx.x; i < L; i += K::T) tw[i] = d.tw[d.f.tw + i];
    col_static_run<L, B>(d, g, v, pre, X, Y, tw);
}

// COL_FWD (gather with the pad rule, forward transform, all L rows stored) and COL_INV (L rows,
// inverse transform, the first n_dst stored) with the line length and column count known at
// compile time, on 512 threads and one LDS buffer (in-place passes), the launch shape the
// run-time y passes take; same descriptor, grid and results (to rounding).
template <int L, int B, int OP>
__global__ __launch_bounds__(kThreadsWide) __attribute__((amdgpu_waves_per_eu(LUCEM_WAVES_PER_EU))) void col_line_static_kernel(
    ColPass d) {
    constexpr int T = kThreadsWide;
    constexpr int n = L * B;
    constexpr int S = (n + T - 1) / T;
    static_assert(OP == COL_FWD || OP == COL_INV, "the plain line passes");
    static_assert(n <= T * kElemsPerThread, "one block of lines per workgroup");
    extern __shared__ float2 lds[];
    float2* Z = lds;
    float2* tw = lds + tw_lds_offset(n);
    const int g = xcd_block(blockIdx.x, gridDim.x);
    const int outer = g / d.nkb;
    const int kx0 = (g - outer * d.nkb) * B;
    const int nl = min(B, d.Lc - kx0);
    const int tid = threadIdx.x;
    const float2* src = d.src + outer * d.src_outer + kx0;
    float2 v[S];
#pragma unroll
    for (int s = 0; s < S; s++) {
        const int e = tid + s * T;
        v[s] = make_float2(0.f, 0.f);
        if (n % T == 0 || e < n) {
            const int i = e / B, b = e - i * B;
            const int sr = OP == COL_INV ? i : pad_index(i - d.left, d.n_src, d.mode);
            if (sr >= 0 && b < nl) v[s] = src[sr * d.src_axis + b];
        }
    }
    for (int i = tid; i < L; i += T) tw[i] = d.tw[d.f.tw + i];
#pragma unroll
    for (int s = 0; s < S; s++) {
        const int e = tid + s * T;
        if (n % T == 0 || e < n) Z[e] = v[s];
    }
    __syncthreads();
    static_fft_inplace<L, B, T, OP == COL_INV>(Z, tw);
    float2* dst = d.dst + outer * d.dst_outer + kx0;
    for (int e = tid; e < d.n_dst * B; e += T) {
        const int i = e / B, b = e - i * B;
        if (b < nl) dst[i * d.dst_axis + b] = Z[e];
    }
}

// The compile-time builds of COL_FWD / COL_INV at 512 threads: (L, B) pairs, tried by the autotuner
// as variant 3.
#define LUCEM_LINE_STATIC_SHAPES(X) X(540, 12) X(540, 14)

bool line_static_available(int L, int B) {
#define LUCEM_LINE_STATIC_AVAIL(l, b) \
    if (L == l && B == b) return true;
    LUCEM_LINE_STATIC_SHAPES(LUCEM_LINE_STATIC_AVAIL)
#undef LUCEM_LINE_STATIC_AVAIL
    return false;
}

static int launch_line_static(const ColPass& p, int blocks, hipStream_t st) {
#define LUCEM_LINE_STATIC_LAUNCH(l, b)                                                                 \
    if (p.f.L == l && p.B == b) {                                                                      \
        const size_t lds = (size_t)(tw_lds_offset(l * b) + l) * sizeof(float2);                       \
        if (p.op == COL_FWD)                                                                           \
            hipLaunchKernelGGL((col_line_static_kernel<l, b, COL_FWD>), dim3(blocks), dim3(kThreadsWide), lds, st, p); \
        else                                                                                           \
            hipLaunchKernelGGL((col_line_static_kernel<l, b, COL_INV>), dim3(blocks), dim3(kThreadsWide), lds, st, p); \
        return (int)hipGetLastError();                                                                 \
    }
    LUCEM_LINE_STATIC_SHAPES(LUCEM_LINE_STATIC_LAUNCH)
#undef LUCEM_LINE_STATIC_LAUNCH
    return (int)hipErrorInvalidValue;
}

// ROW_RATIO / ROW_UPDATE with the row length L and the pair count P known at compile time (the
// one-array form, no strided views). As in row_pass_body: the pairs' half spectra are merged into
// complex lines, inverse-transformed, the pointwise step runs at each padded position's source
// column, and the padded real pair is transformed forward and split. The pointwise operands (the
// observed rows, or the estimate) are loaded with the spectra, before the inlined inverse
// transform, which does not wait for them. LDS: two buffers of L x P values, then the twiddles.
template <int L, int P>
struct RowStatic {
    static constexpr int T = kThreads;
    static constexpr int n = L * P;
    static constexpr int Lc = L / 2 + 1;
    static constexpr int S = (n + T - 1) / T;            // padded positions per thread
    static constexpr int KS = (P * Lc + T - 1) / T;      // spectrum bins per thread
    static_assert(L % 2 == 0 && n <= T * kElemsPerThread, "row pairs of an even length");
};

template <int L, int P>
constexpr size_t row_static_lds_bytes() {
    return (size_t)(2 * tw_lds_offset(L * P) + L) * sizeof(float2);
}

template <int L, int P, int OP>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(LUCEM_WAVES_PER_EU))) void row_static_kernel(
    RowPass d) {
    using K = RowStatic<L, P>;
    static_assert(OP == ROW_RATIO || OP == ROW_UPDATE, "the iteration's two row passes");
    extern __shared__ float2 lds[];
    float2* X = lds;
    float2* Y = lds + tw_lds_offset(K::n);
    float2* tw = Y + tw_lds_offset(K::n);
    const int tid = threadIdx.x;
    const int pair0 = blockIdx.x * P;
    const int nx = d.nx;
    float2 av[K::KS], bv[K::KS];
#pragma unroll
    for (int s = 0; s < K::KS; s++) {
        const int e = tid + s * K::T;
        av[s] = make_float2(0.f, 0.f);
        bv[s] = make_float2(0.f, 0.f);
        if (e < P * K::Lc) {
            const int b = e / K::Lc, k = e - b * K::Lc;
            const int ra = 2 * (pair0 + b), rb = ra + 1;
            if (ra < d.rows) av[s] = d.spec[(int64_t)ra * d.pitch + k];
            if (rb < d.rows) bv[s] = d.spec[(int64_t)rb * d.pitch + k];
        }
    }
    float2 val[K::S];
    int srcs[K::S];
#pragma unroll
    for (int s = 0; s < K::S; s++) {
        const int e = tid + s * K::T;
        srcs[s] = -1;
        val[s] = make_float2(0.f, 0.f);
        if (K::n % K::T == 0 || e < K::n) {
            const int b = e / L, i = e - b * L;
            const int src = pad_index(i - d.left, nx, d.mode);
            srcs[s] = src;
            if (src >= 0) {
                const int ra = 2 * (pair0 + b), rb = ra + 1;
                if constexpr (OP == ROW_RATIO) {
                    if (ra < d.rows) val[s].x = d.obs[(int64_t)ra * nx + src];
                    if (rb < d.rows) val[s].y = d.obs[(int64_t)rb * nx + src];
                } else if (!d.first) {
                    if (ra < d.rows) val[s].x = d.est[(int64_t)ra * nx + src];
                    if (rb < d.rows) val[s].y = d.est[(int64_t)rb * nx + src];
                } else {
                    val[s] = make_float2(d.init, d.init);
                }
            }
        }
    }
    for (int i = tid; i < L; i += K::T) tw[i] = d.tw[d.f.tw + i];
#pragma unroll
    for (int s = 0; s < K::KS; s++) {
        const int e = tid + s * K::T;
        if (e < P * K::Lc) {
            const int b = e / K::Lc, k = e - b * K::Lc;
            float2 A = av[s], Bv = bv[s];
            if (k == 0 || 2 * k == L) {
                A.y = 0.f;
                Bv.y = 0.f;
            }
            X[k * P + b] = make_float2(A.x - Bv.y, A.y + Bv.x);
            const int km = L - k;
            if (k > 0 && km >= K::Lc) X[km * P + b] = make_float2(A.x + Bv.y, Bv.x - A.y);
        }
    }
    __syncthreads();
    float2* r = static_fft<L, P, K::T, true>(X, Y, tw);
    float2* W = r == X ? Y : X;  // free since the last inverse pass
#pragma unroll
    for (int s = 0; s < K::S; s++) {
        const int e = tid + s * K::T;
        if (K::n % K::T == 0 || e < K::n) {
            const int b = e / L, i = e - b * L;
            const int src = srcs[s];
            if (src >= 0) {
                const float2 c = r[src * P + b];
                if constexpr (OP == ROW_RATIO) {
                    const float da = fabsf(c.x) < d.eps ? d.eps : c.x;
                    const float db = fabsf(c.y) < d.eps ? d.eps : c.y;
                    val[s] = make_float2(div_fast(val[s].x, da), div_fast(val[s].y, db));
                } else {
                    float oa = val[s].x * c.x, ob = val[s].y * c.y;
                    if (d.last && d.clip) {
                        oa = fminf(fmaxf(oa, -1.f), 1.f);
                        ob = fminf(fmaxf(ob, -1.f), 1.f);
                    }
                    val[s] = make_float2(oa, ob);
                    if (i == src + d.left) {
                        const int x = i - d.left;
                        const int ra = 2 * (pair0 + b), rb = ra + 1;
                        float* dst = d.last ? d.out : d.est;
                        if (ra < d.rows) dst[(int64_t)ra * nx + x] = val[s].x;
                        if (rb < d.rows) dst[(int64_t)rb * nx + x] = val[s].y;
                    }
                }
            }
            W[i * P + b] = val[s];
        }
    }
    if constexpr (OP == ROW_UPDATE) {
        if (d.last) return;  // uniform: the whole workgroup leaves
    }
    __syncthreads();
    const float2* f = static_fft<L, P, K::T, false>(W, r, tw);
    for (int e = tid; e < P * K::Lc; e += K::T) {
        const int b = e / K::Lc, k = e - b * K::Lc;
        const int ra = 2 * (pair0 + b), rb = ra + 1;
        const float2 zk = f[k * P + b];
        const float2 zm = f[(k == 0 ? 0 : L - k) * P + b];
        if (ra < d.rows) d.spec[(int64_t)ra * d.pitch + k] = make_float2(zk.x + zm.x, zk.y - zm.y);
        if (rb < d.rows) d.spec[(int64_t)rb * d.pitch + k] = make_float2(zk.y + zm.y, zm.x - zk.x);
    }
}

// The compile-time builds of ROW_RATIO / ROW_UPDATE: (L, P) pairs, tried by the autotuner as variant 3.
#define LUCEM_ROW_STATIC_SHAPES(X) X(270, 4) X(270, 6) X(280, 6) X(540, 2) X(540, 3) X(1080, 1)

bool row_static_available(int L, int P) {
#define LUCEM_ROW_STATIC_AVAIL(l, p) \
    if (L == l && P == p) return true;
    LUCEM_ROW_STATIC_SHAPES(LUCEM_ROW_STATIC_AVAIL)
#undef LUCEM_ROW_STATIC_AVAIL
    return false;
}

static int launch_row_static(const RowPass& r, int blocks, hipStream_t st) {
#define LUCEM_ROW_STATIC_LAUNCH(l, p)                                                                  \
    if (r.f.L == l && r.P == p) {                                                                      \
        const size_t lds = row_static_lds_bytes<l, p>();                                              \
        if (r.op == ROW_RATIO)                                                                         \
            hipLaunchKernelGGL((row_static_kernel<l, p, ROW_RATIO>), dim3(blocks), dim3(kThreads), lds, st, r); \
        else                                                                                           \
            hipLaunchKernelGGL((row_static_kernel<l, p, ROW_UPDATE>), dim3(blocks), dim3(kThreads), lds, st, r); \
        return (int)hipGetLastError();                                                                 \
    }
    LUCEM_ROW_STATIC_SHAPES(LUCEM_ROW_STATIC_LAUNCH)
#undef LUCEM_ROW_STATIC_LAUNCH
    return (int)hipErrorInvalidValue;
}

// COL_FUSED on one LDS buffer (in-place compile-time passes), for wider column blocks than two
// buffers allow: the transfer-function values of this thread's elements are loaded with the gather,
// in flight during the forward transform, and multiplied in a sweep after it.
template <int L, int B>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(LUCEM_WAVES_PER_EU))) void col_fused_inplace_kernel(
    ColPass d) {
    constexpr int T = kThreads;
    constexpr int n = L * B;
    constexpr int S = (n + T - 1) / T;
    static_assert(n <= T * kElemsPerThread, "one block of lines per workgroup");
    extern __shared__ float2 lds[];
    float2* Z = lds;
    float2* tw = lds + tw_lds_offset(n);
    const int g = xcd_block(blockIdx.x, gridDim.x);
    const int outer = g / d.nkb;
    const int kx0 = (g - outer * d.nkb) * B;
    const int nl = min(B, d.Lc - kx0);
    const int tid = threadIdx.x;
    const float2* src = d.src + outer * d.src_outer + kx0;
    const float2* otf = d.otf + outer * d.otf_outer + kx0;
    float2 v[S], w[S];
#pragma unroll
    for (int s = 0; s < S; s++) {
        const int e = tid + s * T;
        v[s] = make_float2(0.f, 0.f);
        w[s] = make_float2(0.f, 0.f);
        if (n % T == 0 || e < n) {
            const int i = e / B, b = e - i * B;
            const int sr = pad_index(i - d.left, d.n_src, d.mode);
            if (b < nl) {
                if (sr >= 0) v[s] = src[sr * d.src_axis + b];
                w[s] = otf[(int64_t)i * d.otf_axis + b];
            }
        }
    }
    for (int i = tid; i < L; i += T) tw[i] = d.tw[d.f.tw + i];
#pragma unroll
    for (int s = 0; s < S; s++) {
        const int e = tid + s * T;
        if (n % T == 0 || e < n) Z[e] = v[s];
    }
    __syncthreads();
    static_fft_inplace<L, B, T, false>(Z, tw);
#pragma unroll
    for (int s = 0; s < S; s++) {
        const int e = tid + s * T;
        if (n % T == 0 || e < n) Z[e] = cmul(Z[e], w[s]);
    }
    __syncthreads();
    static_fft_inplace<L, B, T, true>(Z, tw);
    float2* dst = d.dst + outer * d.dst_outer + kx0;
    for (int e = tid; e < d.n_dst * B; e += T) {
        const int i = e / B, b = e - i * B;
        if (b < nl) dst[i * d.dst_axis + b] = Z[e];
    }
}

#define LUCEM_COL_INPLACE_SHAPES(X) X(160, 16) X(160, 20) X(160, 24)

// The compile-time builds of COL_FUSED: (L, B) pairs, tried by the autotuner as variant 3.
#define LUCEM_COL_STATIC_SHAPES(X) \
    X(64, 32) X(96, 21) X(128, 16) X(160, 8) X(160, 12) X(192, 10) X(256, 8)

bool col_static_available(int L, int B) {
#define LUCEM_COL_STATIC_AVAIL(l, b) \
    if (L == l && B == b) return true;
    LUCEM_COL_STATIC_SHAPES(LUCEM_COL_STATIC_AVAIL)
    LUCEM_COL_INPLACE_SHAPES(LUCEM_COL_STATIC_AVAIL)
#undef LUCEM_COL_STATIC_AVAIL
    return false;
}

static int launch_col_static(const ColPass& p, int blocks, hipStream_t st) {
#define LUCEM_COL_STATIC_LAUNCH(l, b)                                                                 \
    if (p.f.L == l && p.B == b) {                                                                     \
        const size_t lds = col_static_lds_bytes<l, b>();                                             \
        hipLaunchKernelGGL((col_fused_static_kernel<l, b>), dim3(blocks), dim3(kThreads), lds, st, p); \
        return (int)hipGetLastError();                                                                \
    }
    LUCEM_COL_STATIC_SHAPES(LUCEM_COL_STATIC_LAUNCH)
#undef LUCEM_COL_STATIC_LAUNCH
#define LUCEM_COL_INPLACE_LAUNCH(l, b)                                                                \
    if (p.f.L == l && p.B == b) {                                                                     \
        const size_t lds = (size_t)(tw_lds_offset(l * b) + l) * sizeof(float2);                      \
        hipLaunchKernelGGL((col_fused_inplace_kernel<l, b>), dim3(blocks), dim3(kThreads), lds, st, p); \
        return (int)hipGetLastError();                                                                \
    }
    LUCEM_COL_INPLACE_SHAPES(LUCEM_COL_INPLACE_LAUNCH)
#undef LUCEM_COL_INPLACE_LAUNCH
    return (int)hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------------------------
// Launchers. FFT kernels hold 16 complex values per thread (at most 128 VGPRs). The thread tier
// of a descriptor picks 256 threads (4096 values, 32 KiB of LDS), 512 (8192 values, 64 KiB) or
// 1024 (16384 values, 128 KiB) per workgroup; a workgroup planned for fewer threads runs
// correctly on more (its loops stride by the thread count).
// ---------------------------------------------------------------------------------------------
size_t row_pass_lds_bytes(int L, int P) { return (size_t)lds_elems(L * P, L) * sizeof(float2); }

size_t col_pass_lds_bytes(int L, int B) { return (size_t)lds_elems(L * B, L) * sizeof(float2); }

// Dynamic LDS above 64 KiB is allowed once per process for the 512- and 1024-thread kernels
// (called by lucem_client_create, never during stream capture).
template <class K>
static int allow_lds(K kernel) {
    return (int)hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)(lds_elems(kLdsComplexXL, kLdsComplexXL) * sizeof(float2)));
}

template <int T>
static int allow_big_lds_tier() {
    constexpr int E = kElemsPerThread;
    int err = 0;
    for (int e : {allow_lds(row_pass_kernel<T, E, ROW_LOAD>), allow_lds(row_pass_kernel<T, E, ROW_LOAD_CONST>),
                  allow_lds(row_pass_kernel<T, E, ROW_RATIO>), allow_lds(row_pass_kernel<T, E, ROW_UPDATE>),
                  allow_lds(row_batch_kernel<T, E, ROW_LOAD>), allow_lds(row_batch_kernel<T, E, ROW_LOAD_CONST>),
                  allow_lds(row_batch_kernel<T, E, ROW_RATIO>), allow_lds(row_batch_kernel<T, E, ROW_UPDATE>),
                  allow_lds(col_pass_kernel<T, E, COL_FWD>), allow_lds(col_pass_kernel<T, E, COL_FUSED>),
                  allow_lds(col_pass_kernel<T, E, COL_INV>), allow_lds(col_pass_kernel<T, E, COL_FWD_OTF>),
                  allow_lds(col_pass_kernel<T, E, COL_FUSED_T>), allow_lds(col_pass_kernel<T, E, COL_INV_FOLD>),
                  allow_lds(col_batch_kernel<T, E, COL_FWD>), allow_lds(col_batch_kernel<T, E, COL_FUSED>),
                  allow_lds(col_batch_kernel<T, E, COL_INV>), allow_lds(col_batch_kernel<T, E, COL_FWD_OTF>),
                  allow_lds(col_batch_kernel<T, E, COL_FUSED_T>), allow_lds(col_batch_kernel<T, E, COL_INV_FOLD>)})
        if (e && !err) err = e;
    return err;
}

int allow_big_lds_kernels() {
    const int a = allow_big_lds_tier<kThreadsWide>();
    const int b = allow_big_lds_tier<kThreadsXL>();
    return a ? a : b;
}

template <int T, int E>
static int launch_row_e(const RowPass& p, int blocks, void* stream) {
    const size_t lds = row_pass_lds_bytes(p.f.L, p.P);
    hipStream_t st = (hipStream_t)stream;
    switch (p.op) {
        case ROW_LOAD:
            hipLaunchKernelGGL((row_pass_kernel<T, E, ROW_LOAD>), dim3(blocks), dim3(T), lds, st, p);
            break;
        case ROW_LOAD_CONST:
            hipLaunchKernelGGL((row_pass_kernel<T, E, ROW_LOAD_CONST>), dim3(blocks), dim3(T), lds, st, p);
            break;
        case ROW_RATIO:
            hipLaunchKernelGGL((row_pass_kernel<T, E, ROW_RATIO>), dim3(blocks), dim3(T), lds, st, p);
            break;
        default:
            hipLaunchKernelGGL((row_pass_kernel<T, E, ROW_UPDATE>), dim3(blocks), dim3(T), lds, st, p);
            break;
    }
    return (int)hipGetLastError();
}

int launch_row_pass(const RowPass& p, int blocks, void* stream) {
    if (p.stat && p.wide == 0 && (p.op == ROW_RATIO || p.op == ROW_UPDATE))
        return launch_row_static(p, blocks, (hipStream_t)stream);
    switch (p.wide) {
        case 2: return launch_row_e<kThreadsXL, kElemsPerThread>(p, blocks, stream);
        case 1: return launch_row_e<kThreadsWide, kElemsPerThread>(p, blocks, stream);
        default: return launch_row_e<kThreads, kElemsPerThread>(p, blocks, stream);
    }
}

template <int T, int E>
static int launch_col_e(const ColPass& p, int blocks, void* stream) {
    const size_t lds = col_pass_lds_bytes(p.f.L, p.B);
    hipStream_t st = (hipStream_t)stream;
#define LUCEM_COL_CASE(OPC)                                                                    \
    case OPC: \
        hipLaunchKernelGGL((col_pass_kernel<T, E, OPC>), dim3(blocks), dim3(T), lds, st, p); \
        break;
    switch (p.op) {
        LUCEM_COL_CASE(COL_FWD)
        LUCEM_COL_CASE(COL_FUSED)
        LUCEM_COL_CASE(COL_INV)
        LUCEM_COL_CASE(COL_FUSED_T)
        LUCEM_COL_CASE(COL_INV_FOLD)
        default:
            hipLaunchKernelGGL((col_pass_kernel<T, E, COL_FWD_OTF>), dim3(blocks), dim3(T), lds, st, p);
            break;
    }
#undef LUCEM_COL_CASE
    return (int)hipGetLastError();
}

int launch_col_pass(const ColPass& p, int blocks, void* stream) {
    if (p.stat && p.op == COL_FUSED && !p.save && p.wide == 0) return launch_col_static(p, blocks, (hipStream_t)stream);
    if (p.stat && (p.op == COL_FWD || p.op == COL_INV) && p.wide == 1)
        return launch_line_static(p, blocks, (hipStream_t)stream);
    switch (p.wide) {
        case 2: return launch_col_e<kThreadsXL, kElemsPerThread>(p, blocks, stream);
        case 1: return launch_col_e<kThreadsWide, kElemsPerThread>(p, blocks, stream);
        default: return launch_col_e<kThreads, kElemsPerThread>(p, blocks, stream);
    }
}

template <int T, int E>
static int launch_row_batch_e(const RowBatch& b, int grid, size_t lds, void* stream) {
    hipStream_t st = (hipStream_t)stream;
    switch (b.op) {
        case ROW_LOAD:
            hipLaunchKernelGGL((row_batch_kernel<T, E, ROW_LOAD>), dim3(grid), dim3(T), lds, st, b);
            break;
        case ROW_LOAD_CONST:
            hipLaunchKernelGGL((row_batch_kernel<T, E, ROW_LOAD_CONST>), dim3(grid), dim3(T), lds, st, b);
            break;
        case ROW_RATIO:
            hipLaunchKernelGGL((row_batch_kernel<T, E, ROW_RATIO>), dim3(grid), dim3(T), lds, st, b);
            break;
        default:
            hipLaunchKernelGGL((row_batch_kernel<T, E, ROW_UPDATE>), dim3(grid), dim3(T), lds, st, b);
            break;
    }
    return (int)hipGetLastError();
}

int launch_row_batch(const RowBatch& b, int grid, size_t lds, void* stream) {
    switch (b.wide) {
        case 2: return launch_row_batch_e<kThreadsXL, kElemsPerThread>(b, grid, lds, stream);
        case 1: return launch_row_batch_e<kThreadsWide, kElemsPerThread>(b, grid, lds, stream);
        default: return launch_row_batch_e<kThreads, kElemsPerThread>(b, grid, lds, stream);
    }
}

template <int T, int E>
static int launch_col_batch_e(const ColBatch& b, int grid, size_t lds, void* stream) {
    hipStream_t st = (hipStream_t)stream;
#define LUCEM_COLB_CASE(OPC)                                                                   \
    case OPC: \
        hipLaunchKernelGGL((col_batch_kernel<T, E, OPC>), dim3(grid), dim3(T), lds, st, b);  \
        break;
    switch (b.op) {
        LUCEM_COLB_CASE(COL_FWD)
        LUCEM_COLB_CASE(COL_FUSED)
        LUCEM_COLB_CASE(COL_INV)
        LUCEM_COLB_CASE(COL_FUSED_T)
        LUCEM_COLB_CASE(COL_INV_FOLD)
        default:
            hipLaunchKernelGGL((col_batch_kernel<T, E, COL_FWD_OTF>), dim3(grid), dim3(T), lds, st, b);
            break;
    }
#undef LUCEM_COLB_CASE
    return (int)hipGetLastError();
}

int launch_col_batch(const ColBatch& b, int grid, size_t lds, void* stream) {
    switch (b.wide) {
        case 2: return launch_col_batch_e<kThreadsXL, kElemsPerThread>(b, grid, lds, stream);
        case 1: return launch_col_batch_e<kThreadsWide, kElemsPerThread>(b, grid, lds, stream);
        default: return launch_col_batch_e<kThreads, kElemsPerThread>(b, grid, lds, stream);
    }
}

int launch_psf_batch(const PsfBatch& b, void* stream) {
    hipLaunchKernelGGL(psf_batch_kernel, dim3(b.count), dim3(256), 0, (hipStream_t)stream, b);
    return (int)hipGetLastError();
}

int launch_zero_ints(int* p, int n, void* stream) {
    hipLaunchKernelGGL(zero_ints_kernel, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, p, n);
    return (int)hipGetLastError();
}

int batch_occupancy(bool row, int tier, size_t lds) {
    int blocks = 0;
    hipError_t e;
    const int T = tier_threads(tier);
    if (row) {
        if (tier == 2)
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, row_batch_kernel<kThreadsXL, kElemsPerThread, ROW_RATIO>, T, lds);
        else if (tier == 1)
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, row_batch_kernel<kThreadsWide, kElemsPerThread, ROW_RATIO>, T, lds);
        else
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, row_batch_kernel<kThreads, kElemsPerThread, ROW_RATIO>, T, lds);
    } else {
        if (tier == 2)
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, col_batch_kernel<kThreadsXL, kElemsPerThread, COL_FUSED>, T, lds);
        else if (tier == 1)
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, col_batch_kernel<kThreadsWide, kElemsPerThread, COL_FUSED>, T, lds);
        else
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, col_batch_kernel<kThreads, kElemsPerThread, COL_FUSED>, T, lds);
    }
    if (e != hipSuccess) return 1;
    return blocks > 0 ? blocks : 1;
}

int launch_psf_prep(const PsfPrep& p, void* stream) {
    hipLaunchKernelGGL(psf_prep_kernel, dim3(1), dim3(256), 0, (hipStream_t)stream, p);
    return (int)hipGetLastError();
}

}  // namespace lucem
