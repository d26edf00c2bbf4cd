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
#include "rl_internal.hpp"

namespace lucem {

// ---------------------------------------------------------------------------------------------
// Row pass. Workgroup g owns row pairs [g*P, g*P + P); pair p is rows (2p, 2p + 1).
// LDS: Z[L * P] (interleaved complex lines) followed by stage[P][2][nx] (real rows).
// ---------------------------------------------------------------------------------------------
template <int T, int E>
__device__ void row_pass_body(const RowPass& d, int g, float2* lds) {
    const int P = 1 << d.lgP;
    const int L = d.f.L;
    const int nx = d.nx;
    float2* Z = lds;
    float* stage = reinterpret_cast<float*>(lds + (size_t)L * P);
    const int tid = threadIdx.x;
    const int pair0 = g * P;

    if (d.op == ROW_RATIO || d.op == ROW_UPDATE) {
        // Z[k] = A[k] + i B[k] for k < Lc, and the Hermitian mirror for k >= Lc.
        for (int b = 0; b < P; b++) {
            const int ra = 2 * (pair0 + b), rb = ra + 1;
            const bool va = ra < d.rows, vb = rb < d.rows;
            const float2* Ar = d.spec + (int64_t)ra * d.pitch;
            const float2* Br = d.spec + (int64_t)rb * d.pitch;
            for (int k = tid; k < d.Lc; k += T) {
                float2 A = va ? Ar[k] : make_float2(0.f, 0.f);
                float2 Bv = vb ? Br[k] : make_float2(0.f, 0.f);
                if (k == 0 || 2 * k == L) {
                    A.y = 0.f;
                    Bv.y = 0.f;
                }
                Z[k * P + b] = make_float2(A.x - Bv.y, A.y + Bv.x);
                const int km = L - k;
                if (k > 0 && km >= d.Lc) Z[km * P + b] = make_float2(A.x + Bv.y, Bv.x - A.y);
            }
        }
        __syncthreads();
        lds_fft<true, T, E>(Z, d.lgP, d.f, d.tw);
        // crop to [0, nx) and apply the pointwise step
        for (int b = 0; b < P; b++) {
            const int ra = 2 * (pair0 + b), rb = ra + 1;
            const bool va = ra < d.rows, vb = rb < d.rows;
            float* sa = stage + (size_t)(2 * b) * nx;
            float* sb = sa + nx;
            for (int x = tid; x < nx; x += T) {
                const float2 c = Z[x * P + b];
                float oa = 0.f, ob = 0.f;
                if (d.op == ROW_RATIO) {
                    if (va) {
                        const float den = fabsf(c.x) < d.eps ? d.eps : c.x;
                        oa = d.obs[(int64_t)ra * nx + x] / den;
                    }
                    if (vb) {
                        const float den = fabsf(c.y) < d.eps ? d.eps : c.y;
                        ob = d.obs[(int64_t)rb * nx + x] / den;
                    }
                } else {
                    float* dst = d.last ? d.out : d.est;
                    if (va) {
                        const float prev = d.first ? d.init : d.est[(int64_t)ra * nx + x];
                        oa = prev * c.x;
                        if (d.last && d.clip) oa = fminf(fmaxf(oa, -1.f), 1.f);
                        dst[(int64_t)ra * nx + x] = oa;
                    }
                    if (vb) {
                        const float prev = d.first ? d.init : d.est[(int64_t)rb * nx + x];
                        ob = prev * c.y;
                        if (d.last && d.clip) ob = fminf(fmaxf(ob, -1.f), 1.f);
                        dst[(int64_t)rb * nx + x] = ob;
                    }
                }
                sa[x] = oa;
                sb[x] = ob;
            }
        }
        if (d.op == ROW_UPDATE && d.last) return;  // uniform: the whole workgroup leaves
        __syncthreads();
        // pad gather from the staged rows
        for (int b = 0; b < P; b++) {
            const float* sa = stage + (size_t)(2 * b) * nx;
            const float* sb = sa + nx;
            for (int i = tid; i < L; i += T) {
                const int s = pad_index(i - d.left, nx, d.mode);
                Z[i * P + b] = s < 0 ? make_float2(0.f, 0.f) : make_float2(sa[s], sb[s]);
            }
        }
    } else {
        // pad gather straight from global memory (or the constant start value)
        for (int b = 0; b < P; b++) {
            const int ra = 2 * (pair0 + b), rb = ra + 1;
            const bool va = ra < d.rows, vb = rb < d.rows;
            const float* Ar = d.src + (int64_t)ra * nx;
            const float* Br = d.src + (int64_t)rb * nx;
            for (int i = tid; i < L; i += T) {
                const int s = pad_index(i - d.left, nx, d.mode);
                float2 z = make_float2(0.f, 0.f);
                if (s >= 0) {
                    if (d.op == ROW_LOAD_CONST) {
                        z = make_float2(va ? d.init : 0.f, vb ? d.init : 0.f);
                    } else {
                        z = make_float2(va ? Ar[s] : 0.f, vb ? Br[s] : 0.f);
                    }
                }
                Z[i * P + b] = z;
            }
        }
    }
    __syncthreads();
    lds_fft<false, T, E>(Z, d.lgP, d.f, d.tw);
    // split the two real transforms: A = Z[k] + conj(Z[L-k]), B = -i (Z[k] - conj(Z[L-k]))
    // (both carry a factor 2 that the transfer function removes)
    for (int b = 0; b < P; b++) {
        const int ra = 2 * (pair0 + b), rb = ra + 1;
        const bool va = ra < d.rows, vb = rb < d.rows;
        float2* Ar = d.spec + (int64_t)ra * d.pitch;
        float2* Br = d.spec + (int64_t)rb * d.pitch;
        for (int k = tid; k < d.Lc; k += T) {
            const float2 zk = Z[k * P + b];
            const float2 zm = Z[(k == 0 ? 0 : L - k) * P + b];
            if (va) Ar[k] = make_float2(zk.x + zm.x, zk.y - zm.y);
            if (vb) Br[k] = make_float2(zk.y + zm.y, zm.x - zk.x);
        }
    }
}

template <int T, int E>
__global__ __launch_bounds__(T) void row_pass_kernel(RowPass d) {
    extern __shared__ float2 lds[];
    row_pass_body<T, E>(d, blockIdx.x, lds);
}

// ---------------------------------------------------------------------------------------------
// Line pass along y or z. Workgroup g owns columns [kb*B, kb*B + B) of outer index `outer`.
// ---------------------------------------------------------------------------------------------
template <int T, int E>
__device__ void col_pass_body(const ColPass& d, int g, float2* Z) {
    const int B = 1 << d.lgB;
    const int outer = g / d.nkb;
    const int kx0 = (g - outer * d.nkb) * B;
    const int L = d.f.L;
    const int tid = threadIdx.x;
    const float2* src = d.src + outer * d.src_outer + kx0;
    const int total = L << d.lgB;
    for (int e = tid; e < total; e += T) {
        const int b = e & (B - 1), i = e >> d.lgB;
        const int s = d.op == COL_INV ? i : pad_index(i - d.left, d.n_src, d.mode);
        float2 v = make_float2(0.f, 0.f);
        if (s >= 0 && kx0 + b < d.Lc) v = src[s * d.src_axis + b];
        Z[e] = v;
    }
    __syncthreads();
    if (d.op == COL_INV) {
        lds_fft<true, T, E>(Z, d.lgB, d.f, d.tw);
    } else {
        lds_fft<false, T, E>(Z, d.lgB, d.f, d.tw);
        if (d.op == COL_FUSED) {
            const float2* otf = d.otf + outer * d.otf_outer + kx0;
            for (int e = tid; e < total; e += T) {
                const int b = e & (B - 1), i = e >> d.lgB;
                if (kx0 + b < d.Lc) Z[e] = cmul(Z[e], otf[i * d.otf_axis + b]);
            }
            __syncthreads();
            lds_fft<true, T, E>(Z, d.lgB, d.f, d.tw);
        }
    }
    float2* dst = d.dst + outer * d.dst_outer + kx0;
    const int nstore = d.n_dst << d.lgB;
    for (int e = tid; e < nstore; e += T) {
        const int b = e & (B - 1), i = e >> d.lgB;
        if (kx0 + b < d.Lc) {
            float2 v = Z[e];
            if (d.finalize) v = make_float2(v.x * d.scale, -v.y * d.scale);
            dst[i * d.dst_axis + b] = v;
        }
    }
}

template <int T, int E>
__global__ __launch_bounds__(T) void col_pass_kernel(ColPass d) {
    extern __shared__ float2 lds[];
    col_pass_body<T, E>(d, blockIdx.x, lds);
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

template <int T, int E>
__global__ __launch_bounds__(T) void row_batch_kernel(RowBatch bt) {
    extern __shared__ float2 lds[];
    __shared__ int slot;
    for (;;) {
        const int it = next_item(bt.counter, &slot);
        if (it >= bt.nitems) break;
        const int2 item = bt.items[it];
        RowPass d = bt.tiles[__builtin_amdgcn_readfirstlane(item.x)];
        d.op = bt.op;
        d.first = bt.first;
        d.last = bt.last;
        d.clip = bt.clip;
        row_pass_body<T, E>(d, __builtin_amdgcn_readfirstlane(item.y), lds);
        __syncthreads();
    }
}

template <int T, int E>
__global__ __launch_bounds__(T) void col_batch_kernel(ColBatch bt) {
    extern __shared__ float2 lds[];
    __shared__ int slot;
    for (;;) {
        const int it = next_item(bt.counter, &slot);
        if (it >= bt.nitems) break;
        const int2 item = bt.items[it];
        const ColPass d = bt.tiles[__builtin_amdgcn_readfirstlane(item.x)];
        col_pass_body<T, E>(d, __builtin_amdgcn_readfirstlane(item.y), lds);
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
size_t row_pass_lds_bytes(int L, int nx, int lgP) {
    const size_t P = (size_t)1 << lgP;
    return P * (size_t)L * sizeof(float2) + P * 2 * (size_t)nx * sizeof(float);
}

size_t col_pass_lds_bytes(int L, int lgB) { return ((size_t)L << lgB) * sizeof(float2); }

int launch_row_pass(const RowPass& p, int blocks, void* stream) {
    const size_t lds = row_pass_lds_bytes(p.f.L, p.nx, p.lgP);
    hipLaunchKernelGGL((row_pass_kernel<kThreads, kElemsPerThread>), dim3(blocks), dim3(kThreads), lds,
                       (hipStream_t)stream, p);
    return (int)hipGetLastError();
}

int launch_col_pass(const ColPass& p, int blocks, void* stream) {
    const size_t lds = col_pass_lds_bytes(p.f.L, p.lgB);
    hipLaunchKernelGGL((col_pass_kernel<kThreads, kElemsPerThread>), dim3(blocks), dim3(kThreads), lds,
                       (hipStream_t)stream, p);
    return (int)hipGetLastError();
}

int launch_row_batch(const RowBatch& b, int grid, size_t lds, void* stream) {
    hipLaunchKernelGGL((row_batch_kernel<kThreads, kElemsPerThread>), dim3(grid), dim3(kThreads), lds,
                       (hipStream_t)stream, b);
    return (int)hipGetLastError();
}

int launch_col_batch(const ColBatch& b, int grid, size_t lds, void* stream) {
    hipLaunchKernelGGL((col_batch_kernel<kThreads, kElemsPerThread>), dim3(grid), dim3(kThreads), lds,
                       (hipStream_t)stream, b);
    return (int)hipGetLastError();
}

int launch_psf_batch(const PsfBatch& b, void* stream) {
    hipLaunchKernelGGL(psf_batch_kernel, dim3(b.count), dim3(256), 0, (hipStream_t)stream, b);
    return (int)hipGetLastError();
}

int launch_zero_ints(int* p, int n, void* stream) {
    hipLaunchKernelGGL(zero_ints_kernel, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, p, n);
    return (int)hipGetLastError();
}

int row_batch_occupancy(size_t lds) {
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, row_batch_kernel<kThreads, kElemsPerThread>,
                                                     kThreads, lds) != hipSuccess)
        return 1;
    return blocks > 0 ? blocks : 1;
}

int col_batch_occupancy(size_t lds) {
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, col_batch_kernel<kThreads, kElemsPerThread>,
                                                     kThreads, lds) != hipSuccess)
        return 1;
    return blocks > 0 ? blocks : 1;
}

int launch_psf_prep(const PsfPrep& p, void* stream) {
    hipLaunchKernelGGL(psf_prep_kernel, dim3(1), dim3(256), 0, (hipStream_t)stream, p);
    return (int)hipGetLastError();
}

}  // namespace lucem
