// Row (x-axis) passes of the differentiable Richardson-Lucy form (gfx950).
//
// Forward model per tile (no clip), for t = 0 .. T-1:
//   u_t = A x_t          A = correlate with rev = flip(h), the reference's deconvolve.rs:120
//   r_t = y / safe(u_t)  restore/kernels/deconvolve.rs:7-22
//   c_t = H r_t          H = correlate with h, deconvolve.rs:133
//   x_{t+1} = x_t * c_t  restore/kernels/deconvolve.rs:24-43
// loss = sum(u_T - y' log u_T), u_T = A x_T, with y' = y or a separate target of counts.
// Backward (reverse mode), with A^T = Pad^T . (correlate with w)^T . Crop^T:
//   ubar_T = w (1 - y / u_T);  xbar_T = A^T ubar_T
//   for t = T-1 .. 0:  cbar = xbar_{t+1} x_t;  xa = xbar_{t+1} c_t;  rbar = H^T cbar;
//                      ubar = -rbar y / u_t^2 (0 where |u_t| < eps);  xbar_t = xa + A^T ubar
// and the PSF gradient accumulates sum_p ubar[p] P(x)[p - l + k] (and the same for H) as
// spectral products inside the transposed line passes.
//
// Each row pass here is: (C2R of the previous line passes, crop or pad-transpose fold) ->
// pointwise step -> gather (pad or crop-transpose) -> R2C, with the same two-rows-per-FFT layout
// as the forward kernels. Storage slots of a checkpoint segment hold x_t, u_t and c_t.
#include <hip/hip_runtime.h>

#include "fft_lds.hpp"
#include "rl_grad.hpp"
#include "rl_internal.hpp"

namespace lucem {

template <int OP>
struct GradOpTraits {
    static constexpr bool c2r = OP != G_LOAD_X && OP != G_SEED && OP != G_LOSS_ADJ;
    static constexpr bool fold = OP == G_ADJ_R || OP == G_ADJ_X;
    static constexpr bool fwd_gather = OP == G_LOAD_X || OP == G_RATIO || OP == G_UPDATE;
};

// Folds the halo columns of P interleaved row-pair lines back onto [0, nx) (pad transpose).
__device__ __forceinline__ void fold_rows(float2* Z, int P, int left, int nx, int mode, int end) {
    const int b = threadIdx.x;
    if (b < P) {
        // the halo only: [0, left), then [left + nx, end), in increasing order
        const int right = left + nx;
        for (int h = left > 0 ? 0 : right; h < end; h = h + 1 == left ? right : h + 1) {
            const int t = pad_index(h - left, nx, mode);
            if (t < 0) continue;
            const float2 v = Z[h * P + b];
            float2& dst = Z[(t + left) * P + b];
            dst = make_float2(dst.x + v.x, dst.y + v.y);
        }
    }
}

template <int T, int E, int OP>
__device__ void grad_row_body(const GradTile& G, const GradRowBatch& L, int item_index, int g, float2* Z,
                              double* red) {
    using TR = GradOpTraits<OP>;
    const RowPass& d = G.r;
    constexpr int S = E;
    const int P = d.P;
    const Divisor pd = Divisor::of(P);
    const int Lf = d.f.L;
    const int nx = d.nx;
    const int tid = threadIdx.x;
    const int pair0 = g * P;
    const int total = Lf * P;
    const float2* twl = stage_twiddles(Z, total, d.f, d.tw);
    const int64_t ns = G.nstride;

    if constexpr (TR::c2r) {
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
                if (k == 0 || 2 * k == Lf) {
                    A.y = 0.f;
                    Bv.y = 0.f;
                }
                Z[k * P + b] = make_float2(A.x - Bv.y, A.y + Bv.x);
                const int km = Lf - k;
                if (k > 0 && km >= Lc) Z[km * P + b] = make_float2(A.x + Bv.y, Bv.x - A.y);
            }
        }
        __syncthreads();
        lds_fft<true, T, E>(Z, pd, d.f, d.tw, twl);
        if constexpr (TR::fold) {
            fold_rows(Z, P, d.left, nx, d.mode, min(Lf, G.kx + nx - 1));
            __syncthreads();
        }
    }

    // pointwise step at every gathered position; owners (i == src + gather_left) store
    const int gleft = TR::fwd_gather ? d.left : 0;
    const int gmode = TR::fwd_gather ? d.mode : 1;  // crop transpose: zero outside [0, nx)
    const int coff = TR::fold ? d.left : 0;         // where the result for column x sits in Z
    float2 val[S];
    int srcs[S];
    double loss = 0.0;
#pragma unroll
    for (int s = 0; s < S; s++) {
        const int e = tid + s * T;
        srcs[s] = -1;
        val[s] = make_float2(0.f, 0.f);
        if (e < total) {
            const int i = pd.div(e), b = e - i * P;
            const int src = pad_index(i - gleft, nx, gmode);
            srcs[s] = src;
            if (src >= 0) {
                const bool own = i == src + gleft;
                float2 c = make_float2(0.f, 0.f);
                if constexpr (TR::c2r) c = Z[(src + coff) * P + b];
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    const int row = 2 * (pair0 + b) + h;
                    if (row >= d.rows) continue;
                    const int64_t at = (int64_t)row * nx + src;
                    const float cv = h ? c.y : c.x;
                    float v = 0.f;
                    if constexpr (OP == G_LOAD_X) {
                        v = L.flags & GF_FROM_X0 ? G.x0[at] : (G.ckpt + L.ck * ns)[at];
                        if (own) (G.xslots + L.sx * ns)[at] = v;
                    } else if constexpr (OP == G_RATIO) {
                        const float den = fabsf(cv) < d.eps ? d.eps : cv;
                        v = div_fast(d.obs[at], den);
                        if (own && (L.flags & GF_SAVE)) (G.uslots + L.su * ns)[at] = cv;
                    } else if constexpr (OP == G_UPDATE) {
                        v = (G.xslots + L.sx * ns)[at] * cv;
                        if (own) {
                            (G.xslots + L.sx1 * ns)[at] = v;
                            if (L.flags & GF_SAVE) (G.cslots + L.su * ns)[at] = cv;
                            if (L.flags & GF_CKPT) (G.ckpt + L.ck * ns)[at] = v;
                            if ((L.flags & GF_LAST) && G.x_out) G.x_out[at] = v;
                        }
                    } else if constexpr (OP == G_LOSS) {
                        if (own) {
                            const float y = G.target ? G.target[at] : d.obs[at];
                            G.uT[at] = cv;
                            loss += (double)cv - (double)y * log((double)cv);
                        }
                    } else if constexpr (OP == G_LOSS_ADJ) {
                        const float wgt = G.weight ? G.weight[0] : 1.0f;
                        v = wgt * (1.0f - (G.target ? G.target[at] : d.obs[at]) / G.uT[at]);
                    } else if constexpr (OP == G_ADJ_R) {
                        const float u = (G.uslots + L.su * ns)[at];
                        v = fabsf(u) < d.eps ? 0.f : -cv * d.obs[at] / (u * u);
                    } else if constexpr (OP == G_ADJ_X) {
                        float xb = cv;
                        if (!(L.flags & GF_NO_XA)) xb += G.xa[at];
                        if (L.flags & GF_SEED_NEXT) {
                            // xbar_t -> the previous iteration: xa = xbar c_{t-1}, cbar = xbar x_{t-1}
                            if (own) G.xa[at] = xb * (G.cslots + L.scm * ns)[at];
                            v = xb * (G.xslots + L.sxm * ns)[at];
                        } else if (own) {
                            if (L.flags & GF_TO_GX0)
                                G.gx0[at] = xb;
                            else
                                G.xbar[at] = xb;
                        }
                    } else if constexpr (OP == G_SEED) {
                        const float xb = G.xbar[at];
                        if (own) G.xa[at] = xb * (G.cslots + L.scm * ns)[at];
                        v = xb * (G.xslots + L.sxm * ns)[at];
                    } else if constexpr (OP == G_C2R) {
                        if (own) G.out[at] = cv * G.scale;
                    }
                    if (h)
                        val[s].y = v;
                    else
                        val[s].x = v;
                }
            }
        }
    }
    if constexpr (OP == G_LOSS) {
        // deterministic per-workgroup partial of the loss
        red[tid] = loss;
        __syncthreads();
        for (int w = T / 2; w > 0; w >>= 1) {
            if (tid < w) red[tid] += red[tid + w];
            __syncthreads();
        }
        if (tid == 0) G.partial[g] = red[0];
    }
    if constexpr (OP == G_C2R || OP == G_LOSS) return;
    if constexpr (OP == G_ADJ_X) {
        if (!(L.flags & GF_SEED_NEXT)) return;  // uniform
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < S; s++) {
        const int e = tid + s * T;
        if (e < total) Z[e] = val[s];
    }
    __syncthreads();
    lds_fft<false, T, E>(Z, pd, d.f, d.tw, twl);
    // item e = b * Lc + k over all pairs, so that every thread has work whatever P and Lc are
    const Divisor lcd = Divisor::of(d.Lc);
    const int nsplit = P * d.Lc;
    for (int e = tid; e < nsplit; e += T) {
        const int b = lcd.div(e), k = e - b * d.Lc;
        const int ra = 2 * (pair0 + b), rb = ra + 1;
        const float2 zk = Z[k * P + b];
        const float2 zm = Z[(k == 0 ? 0 : Lf - k) * P + b];
        if (ra < d.rows) d.spec[(int64_t)ra * d.pitch + k] = make_float2(zk.x + zm.x, zk.y - zm.y);
        if (rb < d.rows) d.spec[(int64_t)rb * d.pitch + k] = make_float2(zk.y + zm.y, zm.x - zk.x);
    }
    (void)item_index;
}

template <int T, int E, int OP>
__global__ __launch_bounds__(T) void grad_row_kernel(GradRowBatch bt) {
    extern __shared__ float2 lds[];
    __shared__ double red[T];
    const int it = xcd_block(blockIdx.x, gridDim.x);
    if (it >= bt.nitems) return;
    const int2 item = bt.items[it];
    const GradTile G = bt.tiles[__builtin_amdgcn_readfirstlane(item.x)];
    grad_row_body<T, E, OP>(G, bt, it, __builtin_amdgcn_readfirstlane(item.y), lds, red);
}

// Sums the loss partials of each tile in a fixed order.
__global__ __launch_bounds__(256) void loss_reduce_kernel(const double* partials, const int* first, const int* count,
                                                          double* loss, int ntiles) {
    __shared__ double red[256];
    const int t = blockIdx.x;
    if (t >= ntiles) return;
    double s = 0.0;
    for (int i = threadIdx.x; i < count[t]; i += 256) s += partials[first[t] + i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss[t] = red[0];
}

// d loss / d psf from the per-tile correlation gradients of h and rev (k elements each):
// dh = sum_tiles (g_h + flip(g_rev)); through the normalisation h = psf / sum(psf) of
// deconvolve.rs:77-79: dpsf = (dh - sum(dh * h)) / sum(psf) when |sum| > 1e-12, else dh.
__global__ __launch_bounds__(256) void psf_grad_kernel(const float* gh, const float* grev, int ntiles, int64_t nk,
                                                       const float* psf, float* out) {
    __shared__ double red[256];
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < nk; i += 256) s += (double)psf[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    const double sum = red[0];
    __syncthreads();
    const bool normed = fabs(sum) > 1e-12;
    const double norm = normed ? 1.0 / sum : 1.0;
    double dot = 0.0;
    for (int64_t i = threadIdx.x; i < nk; i += 256) {
        double dh = 0.0;
        for (int t = 0; t < ntiles; t++) dh += (double)gh[t * nk + i] + (double)grev[t * nk + (nk - 1 - i)];
        dot += dh * (double)psf[i] * norm;
    }
    red[threadIdx.x] = dot;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    const double total = red[0];
    for (int64_t i = threadIdx.x; i < nk; i += 256) {
        double dh = 0.0;
        for (int t = 0; t < ntiles; t++) dh += (double)gh[t * nk + i] + (double)grev[t * nk + (nk - 1 - i)];
        out[i] = (float)(normed ? (dh - total) * norm : dh);
    }
}

// p is 16-byte aligned; grid-stride over float4 stores, then the tail
__global__ void zero_floats_kernel(float* p, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t n4 = n >> 2;
    float4* q = reinterpret_cast<float4*>(p);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride)
        q[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    const int64_t t = (n4 << 2) + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = 0.f;
}

template <int T>
static int launch_grad_row_t(const GradRowBatch& b, size_t lds, hipStream_t st) {
    constexpr int E = kElemsPerThread;
    const dim3 grid(b.nitems), block(T);
    switch (b.op) {
        case G_LOAD_X: hipLaunchKernelGGL((grad_row_kernel<T, E, G_LOAD_X>), grid, block, lds, st, b); break;
        case G_RATIO: hipLaunchKernelGGL((grad_row_kernel<T, E, G_RATIO>), grid, block, lds, st, b); break;
        case G_UPDATE: hipLaunchKernelGGL((grad_row_kernel<T, E, G_UPDATE>), grid, block, lds, st, b); break;
        case G_LOSS: hipLaunchKernelGGL((grad_row_kernel<T, E, G_LOSS>), grid, block, lds, st, b); break;
        case G_ADJ_R: hipLaunchKernelGGL((grad_row_kernel<T, E, G_ADJ_R>), grid, block, lds, st, b); break;
        case G_ADJ_X: hipLaunchKernelGGL((grad_row_kernel<T, E, G_ADJ_X>), grid, block, lds, st, b); break;
        case G_SEED: hipLaunchKernelGGL((grad_row_kernel<T, E, G_SEED>), grid, block, lds, st, b); break;
        case G_LOSS_ADJ: hipLaunchKernelGGL((grad_row_kernel<T, E, G_LOSS_ADJ>), grid, block, lds, st, b); break;
        default: hipLaunchKernelGGL((grad_row_kernel<T, E, G_C2R>), grid, block, lds, st, b); break;
    }
    return (int)hipGetLastError();
}

int launch_grad_row(const GradRowBatch& b, size_t lds, void* stream) {
    hipStream_t st = (hipStream_t)stream;
    switch (b.wide) {
        case 2: return launch_grad_row_t<kThreadsXL>(b, lds, st);
        case 1: return launch_grad_row_t<kThreadsWide>(b, lds, st);
        default: return launch_grad_row_t<kThreads>(b, lds, st);
    }
}

template <int T>
static int allow_grad_tier() {
    constexpr int E = kElemsPerThread;
    const int bytes = (int)(lds_elems(kLdsComplexXL, kLdsComplexXL) * sizeof(float2));
    int err = 0;
    for (const void* k : {(const void*)grad_row_kernel<T, E, G_LOAD_X>, (const void*)grad_row_kernel<T, E, G_RATIO>,
                          (const void*)grad_row_kernel<T, E, G_UPDATE>, (const void*)grad_row_kernel<T, E, G_LOSS>,
                          (const void*)grad_row_kernel<T, E, G_ADJ_R>, (const void*)grad_row_kernel<T, E, G_ADJ_X>,
                          (const void*)grad_row_kernel<T, E, G_SEED>, (const void*)grad_row_kernel<T, E, G_LOSS_ADJ>,
                          (const void*)grad_row_kernel<T, E, G_C2R>}) {
        const int e = (int)hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
        if (e && !err) err = e;
    }
    return err;
}

int allow_big_lds_grad_kernels() {
    const int a = allow_grad_tier<kThreadsWide>();
    const int b = allow_grad_tier<kThreadsXL>();
    return a ? a : b;
}

int launch_loss_reduce(const double* partials, const int* first, const int* count, double* loss, int ntiles,
                       void* stream) {
    hipLaunchKernelGGL(loss_reduce_kernel, dim3(ntiles), dim3(256), 0, (hipStream_t)stream, partials, first, count,
                       loss, ntiles);
    return (int)hipGetLastError();
}

int launch_psf_grad(const float* gh, const float* grev, int ntiles, int64_t nk, const float* psf, float* out,
                    void* stream) {
    hipLaunchKernelGGL(psf_grad_kernel, dim3(1), dim3(256), 0, (hipStream_t)stream, gh, grev, ntiles, nk, psf, out);
    return (int)hipGetLastError();
}

int launch_zero_floats(float* p, int64_t n, void* stream) {
    if (n <= 0) return 0;
    const int64_t blocks = std::min<int64_t>(((n >> 2) + 255) / 256 + 1, 8192);
    hipLaunchKernelGGL(zero_floats_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, p, n);
    return (int)hipGetLastError();
}

}  // namespace lucem
