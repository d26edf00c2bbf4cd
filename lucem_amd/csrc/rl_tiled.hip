// One Richardson-Lucy iteration of a 2-D array per launch, a workgroup per output tile
// (gfx950: 1024 threads, the tile's whole 2-D transform in LDS).
//
// The iteration of the reference (restore/deconvolve.rs:109-147) is
//   u   = correlate(x, rev)      (:120)      rel = y / safe(u)  (:121-132)
//   c   = correlate(rel, h)      (:133)      x'  = x * c        (:134-145)
// with both correlations padded by the boundary rule (filter/kernels/convolve.rs:61-118). For an
// output tile O = [t0, t0 + T) per axis, c on O needs P(rel) on [t0 - left, t0 + T - 1 + right],
// hence u on the pad targets of that window, which the host checks lie inside the window itself,
// hence P(x) on [t0 - 2 left, t0 + T - 1 + 2 right]. Each correlation is an overlap-save FFT
// correlation on an L0 x L1 domain with L >= T + 2 (k - 1) per axis, so no circular wrap reaches
// the values kept. The whole chain stays in LDS: the iteration reads x with its halo, y on the rel
// window and the two transfer functions, and writes x' on the tile.
//
// LDS: Zr, the tile's rows transformed in pairs (L0 / 2 interleaved lines of length L1, element
// c of pair b at Zr[c * B + b]); Sc, its half spectrum as column lines (L1 / 2 + 1 of length L0,
// element r of column k at Sc[r * C + k]); the twiddle tables of L1 and L0. Line counts B and C
// are made odd with a padding line, so that a sweep along a line's elements (the gathers, the
// split and merge between the two layouts) spreads over the LDS banks instead of striding a
// multiple of 16 words.
#include <hip/hip_runtime.h>

#include "fft_lds.hpp"
#include "rl_internal.hpp"

namespace lucem {

namespace {
constexpr int kT = kTileThreads;
constexpr int kE = kTileElems;
constexpr int kG = kTileMaxLen / 2 / (kT / kTileMaxLen);  // row-pair sweeps of a tile: 8
constexpr int kW = 4;                                     // waves per SIMD: 16 per workgroup

struct TileLds {
    float2* Zr;  // row-pair lines: element c of pair b at Zr[c * B + b]
    float2* Sc;  // column lines: element r of column k at Sc[r * C + k]
    const float2* tw0;
    const float2* tw1;
    Divisor bd;  // B = tile_lines(L0 / 2) lines (odd: a padding line when L0 / 2 is even)
    Divisor cd;  // C = tile_lines(L1 / 2 + 1) lines
    int nb, nc;  // the lines that hold data: L0 / 2 row pairs, L1 / 2 + 1 columns
};

// The real tile in Zr -> its 2-D spectrum in Sc. The row split carries a factor 2, as the
// transfer functions (built by the same row transform) expect.
__device__ void forward_2d(const TileIter& d, const TileLds& t) {
    lds_fft<false, kT, kE, kW>(t.Zr, t.bd, d.f1, d.tw1, t.tw1);
    const int B = t.bd.d, C = t.cd.d, nc = t.nc, L1 = d.L1;
    const Divisor ncd = Divisor::of(nc);
    for (int e = threadIdx.x; e < t.nb * nc; e += kT) {
        const int b = ncd.div(e), k = e - b * nc;
        const float2 zk = t.Zr[k * B + b];
        const float2 zm = t.Zr[(k == 0 ? 0 : L1 - k) * B + b];
        t.Sc[(2 * b) * C + k] = make_float2(zk.x + zm.x, zk.y - zm.y);
        t.Sc[(2 * b + 1) * C + k] = make_float2(zk.y + zm.y, zm.x - zk.x);
    }
    __syncthreads();
    lds_fft<false, kT, kE, kW>(t.Sc, t.cd, d.f0, d.tw0, t.tw0);
}

// Sc *= otf ([L0][pitch] in global memory; conj(FFT(w)) * scale) on the data columns, loads
// issued before the LDS reads.
__device__ void multiply(const TileIter& d, const TileLds& t, const float2* __restrict__ otf) {
    const int C = t.cd.d, n = d.L0 * C;
    float2 w[kE];
#pragma unroll
    for (int s = 0; s < kE; s++) {
        const int e = threadIdx.x + s * kT;
        w[s] = make_float2(0.f, 0.f);
        if (e < n) {
            const int r = t.cd.div(e), k = e - r * C;
            if (k < t.nc) w[s] = otf[(int64_t)r * d.pitch + k];
        }
    }
#pragma unroll
    for (int s = 0; s < kE; s++) {
        const int e = threadIdx.x + s * kT;
        if (e < n) t.Sc[e] = cmul(t.Sc[e], w[s]);
    }
    __syncthreads();
}

// The spectrum in Sc -> the real tile in Zr (unnormalised: the transfer function carries the
// 1 / (4 L0 L1)).
__device__ void inverse_2d(const TileIter& d, const TileLds& t) {
    lds_fft<true, kT, kE, kW>(t.Sc, t.cd, d.f0, d.tw0, t.tw0);
    const int B = t.bd.d, C = t.cd.d, nc = t.nc, L1 = d.L1;
    const Divisor ncd = Divisor::of(nc);
    for (int e = threadIdx.x; e < t.nb * nc; e += kT) {
        const int b = ncd.div(e), k = e - b * nc;
        float2 A = t.Sc[(2 * b) * C + k], Bv = t.Sc[(2 * b + 1) * C + k];
        if (k == 0 || 2 * k == L1) {
            A.y = 0.f;
            Bv.y = 0.f;
        }
        t.Zr[k * B + b] = make_float2(A.x - Bv.y, A.y + Bv.x);
        const int km = L1 - k;
        if (k > 0 && km >= nc) t.Zr[km * B + b] = make_float2(A.x + Bv.y, Bv.x - A.y);
    }
    __syncthreads();
    lds_fft<true, kT, kE, kW>(t.Zr, t.bd, d.f1, d.tw1, t.tw1);
}

// A box of the image (rows y0 + r, columns x0 + c) lies inside it: no boundary rule applies.
__device__ __forceinline__ bool box_inside(int y0, int x0, int h, int w, int n0, int n1) {
    return y0 >= 0 && x0 >= 0 && y0 + h <= n0 && x0 + w <= n1;
}

__global__ __launch_bounds__(kT) __attribute__((amdgpu_waves_per_eu(kW))) void tile_iter_kernel(TileIter d) {
    extern __shared__ float2 lds[];
    const int nb = d.L0 / 2, nc = d.L1 / 2 + 1;
    const int B = tile_lines(nb), C = tile_lines(nc);
    TileLds t;
    t.Zr = lds;
    t.Sc = lds + d.L1 * B;
    float2* tw1 = t.Sc + d.L0 * C;
    float2* tw0 = tw1 + d.L1;
    t.tw1 = tw1;
    t.tw0 = tw0;
    t.bd = Divisor::of(B);
    t.cd = Divisor::of(C);
    t.nb = nb;
    t.nc = nc;
    const int tid = threadIdx.x;
    const int ty = blockIdx.x / d.tiles1, tx = blockIdx.x - ty * d.tiles1;
    const int t0y = ty * d.T0, t0x = tx * d.T1;
    // Tile sweeps: thread (pr, c) handles row pairs pr, pr + pstep, ... at column c (columns
    // fastest across lanes, so global rows are read whole and a pair is one 8-byte LDS access).
    const int pstep = kT / d.L1;
    const int c = tid - (tid / d.L1) * d.L1;
    const int pr0 = tid / d.L1;
    const bool active = pr0 < pstep;

    for (int i = tid; i < d.L1; i += kT) tw1[i] = d.tw1[d.f1.tw + i];
    for (int i = tid; i < d.L0; i += kT) tw0[i] = d.tw0[d.f0.tw + i];
    // padding lines hold zeros (transforms of them stay finite and are never read)
    if (B > nb)
        for (int i = tid; i < d.L1; i += kT) t.Zr[i * B + nb] = make_float2(0.f, 0.f);
    if (C > nc)
        for (int i = tid; i < d.L0; i += kT) t.Sc[i * C + nc] = make_float2(0.f, 0.f);

    // Q1: the estimate with the pad rule, two halos deep
    {
        const int gy0 = t0y - 2 * d.left0, gx0 = t0x - 2 * d.left1;
        const bool inside = box_inside(gy0, gx0, d.L0, d.L1, d.n0, d.n1);
        float2 v[kG];
#pragma unroll
        for (int s = 0; s < kG; s++) {
            const int b = pr0 + s * pstep;
            v[s] = make_float2(0.f, 0.f);
            if (active && b < nb) {
                if (d.first) {
                    if (inside) {
                        v[s] = make_float2(d.init, d.init);
                    } else {
                        const int sx = pad_index(gx0 + c, d.n1, d.mode);
                        const int sa = pad_index(gy0 + 2 * b, d.n0, d.mode), sb = pad_index(gy0 + 2 * b + 1, d.n0, d.mode);
                        v[s] = make_float2(sx >= 0 && sa >= 0 ? d.init : 0.f, sx >= 0 && sb >= 0 ? d.init : 0.f);
                    }
                } else if (inside) {
                    const float* p = d.est_in + (int64_t)(gy0 + 2 * b) * d.n1 + gx0 + c;
                    v[s] = make_float2(p[0], p[d.n1]);
                } else {
                    const int sx = pad_index(gx0 + c, d.n1, d.mode);
                    const int sa = pad_index(gy0 + 2 * b, d.n0, d.mode), sb = pad_index(gy0 + 2 * b + 1, d.n0, d.mode);
                    if (sx >= 0) {
                        if (sa >= 0) v[s].x = d.est_in[(int64_t)sa * d.n1 + sx];
                        if (sb >= 0) v[s].y = d.est_in[(int64_t)sb * d.n1 + sx];
                    }
                }
            }
        }
#pragma unroll
        for (int s = 0; s < kG; s++) {
            const int b = pr0 + s * pstep;
            if (active && b < nb) t.Zr[c * B + b] = v[s];
        }
    }
    __syncthreads();
    forward_2d(d, t);
    multiply(d, t, d.otf1);
    inverse_2d(d, t);

    // Zr now holds u(t0 - left + p) at p. Q2 = P(rel) on the rel window (my x mx), 0 beyond.
    {
        const int my = d.T0 + d.k0 - 1, mx = d.T1 + d.k1 - 1;
        const int qy0 = t0y - d.left0, qx0 = t0x - d.left1;
        const bool inside = box_inside(qy0, qx0, my, mx, d.n0, d.n1);
        const float* Zf = reinterpret_cast<const float*>(t.Zr);
        float2 v[kG];
        if (inside) {
            // every target is its own position: y and u of the pair, rel in place
            float2 u[kG];
#pragma unroll
            for (int s = 0; s < kG; s++) {
                const int b = pr0 + s * pstep;
                v[s] = make_float2(0.f, 0.f);
                if (active && b < nb && c < mx) {
                    const float* p = d.obs + (int64_t)(qy0 + 2 * b) * d.n1 + qx0 + c;
                    if (2 * b < my) v[s].x = p[0];
                    if (2 * b + 1 < my) v[s].y = p[d.n1];
                }
            }
#pragma unroll
            for (int s = 0; s < kG; s++) {
                const int b = pr0 + s * pstep;
                u[s] = make_float2(1.f, 1.f);
                if (active && b < nb && c < mx) u[s] = t.Zr[c * B + b];
            }
#pragma unroll
            for (int s = 0; s < kG; s++) {
                v[s].x = div_fast(v[s].x, fabsf(u[s].x) < d.eps ? d.eps : u[s].x);
                v[s].y = div_fast(v[s].y, fabsf(u[s].y) < d.eps ? d.eps : u[s].y);
            }
        } else {
            int at[2 * kG];
#pragma unroll
            for (int s = 0; s < kG; s++) {
                const int b = pr0 + s * pstep;
                v[s] = make_float2(0.f, 0.f);
                at[2 * s] = at[2 * s + 1] = -1;
                if (active && b < nb && c < mx) {
                    const int sx = pad_index(qx0 + c, d.n1, d.mode);
#pragma unroll
                    for (int q = 0; q < 2; q++) {
                        const int r = 2 * b + q;
                        const int sy = r < my ? pad_index(qy0 + r, d.n0, d.mode) : -1;
                        // the host admits only geometries whose targets lie in the window; the
                        // bound keeps every LDS read inside the tile regardless
                        if (sy >= 0 && sx >= 0 && (unsigned)(sy - qy0) < (unsigned)my &&
                            (unsigned)(sx - qx0) < (unsigned)mx) {
                            const float y = d.obs[(int64_t)sy * d.n1 + sx];
                            if (q == 0) v[s].x = y; else v[s].y = y;
                            const int ry = sy - qy0;
                            at[2 * s + q] = 2 * ((sx - qx0) * B + (ry >> 1)) + (ry & 1);
                        }
                    }
                }
            }
#pragma unroll
            for (int s = 0; s < kG; s++) {
                if (at[2 * s] >= 0) {
                    const float u = Zf[at[2 * s]];
                    v[s].x = div_fast(v[s].x, fabsf(u) < d.eps ? d.eps : u);
                }
                if (at[2 * s + 1] >= 0) {
                    const float u = Zf[at[2 * s + 1]];
                    v[s].y = div_fast(v[s].y, fabsf(u) < d.eps ? d.eps : u);
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < kG; s++) {
            const int b = pr0 + s * pstep;
            if (active && b < nb) t.Zr[c * B + b] = v[s];
        }
    }
    __syncthreads();
    forward_2d(d, t);
    multiply(d, t, d.otf2);
    inverse_2d(d, t);

    // Zr now holds c(t0 + p) at p: the update on the tile's own pixels
    {
        const int h = min(d.T0, d.n0 - t0y), w = min(d.T1, d.n1 - t0x);
        float2 x[kG];
        float* dst = d.last ? d.out : d.est_out;
#pragma unroll
        for (int s = 0; s < kG; s++) {
            const int b = pr0 + s * pstep;
            x[s] = make_float2(d.init, d.init);
            if (!d.first && active && 2 * b < h && c < w) {
                const float* p = d.est_in + (int64_t)(t0y + 2 * b) * d.n1 + t0x + c;
                x[s].x = p[0];
                if (2 * b + 1 < h) x[s].y = p[d.n1];
            }
        }
#pragma unroll
        for (int s = 0; s < kG; s++) {
            const int b = pr0 + s * pstep;
            if (active && 2 * b < h && c < w) {
                const float2 cv = t.Zr[c * B + b];
                float xa = x[s].x * cv.x, xb = x[s].y * cv.y;
                if (d.last && d.clip) {
                    xa = fminf(fmaxf(xa, -1.f), 1.f);
                    xb = fminf(fmaxf(xb, -1.f), 1.f);
                }
                float* p = dst + (int64_t)(t0y + 2 * b) * d.n1 + t0x + c;
                p[0] = xa;
                if (2 * b + 1 < h) p[d.n1] = xb;
            }
        }
    }
}
}  // namespace

size_t tile_iter_lds_bytes(int L0, int L1) {
    return sizeof(float2) * ((size_t)L1 * tile_lines(L0 / 2) + (size_t)L0 * tile_lines(L1 / 2 + 1) + (size_t)L1 +
                             (size_t)L0);
}

int allow_tile_lds() {
    return (int)hipFuncSetAttribute(reinterpret_cast<const void*>(tile_iter_kernel),
                                    hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)tile_iter_lds_bytes(kTileMaxLen, kTileMaxLen));
}

int launch_tile_iter(const TileIter& d, int blocks, void* stream) {
    hipLaunchKernelGGL(tile_iter_kernel, dim3(blocks), dim3(kT), tile_iter_lds_bytes(d.L0, d.L1), (hipStream_t)stream,
                       d);
    return (int)hipGetLastError();
}

}  // namespace lucem
