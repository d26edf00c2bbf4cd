// One Richardson-Lucy iteration of a 2-D array per launch, a workgroup per output tile
// (gfx950: 1024 threads, the tile's whole 2-D transform in LDS).
//
// The iteration of the reference (restore/deconvolve.rs:109-147) is
//   u   = correlate(x, rev)      (:120)      rel = y / safe(u)  (:121-132)
//   c   = correlate(rel, h)      (:133)      x'  = x * c        (:134-145)
// with both correlations padded by the boundary rule (filter/kernels/convolve.rs:61-118). For an
// output tile O = [t0, t0 + T) per axis (fewer on an array's last tile), c on O needs P(rel) on
// [t0 - left, t0 + |O| - 1 + right], hence u on the pad targets of that window, which the host
// checks lie inside the window itself,
// hence P(x) on [t0 - 2 left, t0 + T - 1 + 2 right]. Each correlation is an overlap-save FFT
// correlation on an L0 x L1 domain with L >= T + 2 (k - 1) per axis, so no circular wrap reaches
// the values kept. The whole chain stays in LDS: the iteration reads x with its halo, y on the rel
// window and the two transfer functions, and writes x' on the tile.
//
// LDS: two buffers that the transform passes alternate between (read one, write the other, one
// barrier per pass), each holding the tile either as row-pair lines (L0 / 2 interleaved lines of
// length L1, element c of pair b at [c * B + b]) or as its half spectrum in column lines
// (L1 / 2 + 1 lines of length L0, element r of column k at [r * C + k]); then the twiddle tables
// of L1 and L0. Line counts B and C are made odd with a padding line, so that a sweep along a
// line's elements (the gathers, the split and merge between the two layouts) spreads over the LDS
// banks instead of striding a multiple of 16 words.
//
// The kernel is compiled twice over: with the tile's lengths known at compile time for the common
// domains (constant strides, immediate LDS offsets, a fixed radix sequence: fft_static.hpp), and
// with run-time lengths for every other admissible domain.
//
// Latency: every global load is issued a transform ahead of its use (the transfer functions, the
// observed window, the estimate on the tile), and the twiddle tables are staged under the first
// gather's loads. Work: the row transforms run only over the row pairs that hold data or are read
// next, and the compile-time build multiplies by the transfer function in the last pass of the
// forward column transform rather than in a sweep of its own.
#include <hip/hip_runtime.h>

#include "fft_lds.hpp"
#include "fft_static.hpp"
#include "rl_internal.hpp"

namespace lucem {

namespace {
constexpr int kT = kTileThreads;
constexpr int kE = kTileElems;
constexpr int kG = kTileMaxLen / 2 / (kT / kTileMaxLen);  // row-pair sweeps of a tile: 8
constexpr int kW = 4;                                     // waves per SIMD: 16 per workgroup

// Tile dimensions at run time (SL0 = SL1 = 0) or at compile time.
template <int SL0, int SL1>
struct Dims {
    static constexpr bool is_static = true;
    static constexpr int L0 = SL0, L1 = SL1, nb = SL0 / 2, nc = SL1 / 2 + 1;
    static constexpr int B = tile_lines(nb), C = tile_lines(nc);
    static constexpr int nbuf = tile_buffer_elems(SL0, SL1);
    __device__ explicit Dims(const TileIter&) {}
    __device__ int div_nc(int e) const { return e / nc; }
    __device__ int div_C(int e) const { return e / C; }
};
template <>
struct Dims<0, 0> {
    static constexpr bool is_static = false;
    int L0, L1, nb, nc, B, C, nbuf;
    Divisor ncd, Cd;
    __device__ explicit Dims(const TileIter& d)
        : L0(d.L0), L1(d.L1), nb(d.L0 / 2), nc(d.L1 / 2 + 1), B(tile_lines(d.L0 / 2)), C(tile_lines(d.L1 / 2 + 1)),
          nbuf(tile_buffer_elems(d.L0, d.L1)), ncd(Divisor::of(d.L1 / 2 + 1)), Cd(Divisor::of(tile_lines(d.L1 / 2 + 1))) {}
    __device__ int div_nc(int e) const { return ncd.div(e); }
    __device__ int div_C(int e) const { return Cd.div(e); }
};

struct TileLds {
    float2* X;
    float2* Y;
    const float2* tw0;
    const float2* tw1;
    __device__ float2* other(const float2* p) const { return p == X ? Y : X; }
};

// Length-L1 transforms of the row-pair lines in `a` (rows) or length-L0 transforms of the column
// lines (cols); the result's buffer is returned. Rows: only the lines below nl are needed (the
// compile-time build transforms just those; the others hold what they held before, or their
// transforms, which no caller reads).
template <bool INV, class D>
__device__ float2* rows_fft(const D& dm, const TileIter& d, const TileLds& t, float2* a, int nl) {
    if constexpr (D::is_static)
        return static_fft<D::L1, D::B, kT, INV, true>(a, t.other(a), t.tw1, Divisor::of(nl));
    else
        return lds_fft_oop<INV, kT, kE, kW>(a, t.other(a), Divisor::of(dm.B), d.f1, t.tw1);
}
template <bool INV, class D>
__device__ float2* cols_fft(const D& dm, const TileIter& d, const TileLds& t, float2* a) {
    if constexpr (D::is_static)
        return static_fft<D::L0, D::C, kT, INV>(a, t.other(a), t.tw0);
    else
        return lds_fft_oop<INV, kT, kE, kW>(a, t.other(a), Divisor::of(dm.C), d.f0, t.tw0);
}

// The real tile in `zr` (row-pair lines) -> its 2-D spectrum (column lines); returns the buffer
// that holds it. Row pairs from nl on are zero (rows past the data) and are not transformed. The
// row split carries a factor 2, as the transfer functions (built by the same row transform) expect.
// The compile-time build multiplies the spectrum by the transfer function `otf` in the column
// transform's last pass (`pre`: this thread's first-round factors, from load_last_pass); the
// run-time build leaves that to apply_otf.
template <class D>
__device__ float2* forward_2d(const D& dm, const TileIter& d, const TileLds& t, float2* zr, int nl,
                              const float2* otf, const float2* pre) {
    const float2* r = rows_fft<false>(dm, d, t, zr, nl);
    float2* sc = t.other(r);
    for (int e = threadIdx.x; e < dm.nb * dm.nc; e += kT) {
        const int b = dm.div_nc(e), k = e - b * dm.nc;
        float2 za = make_float2(0.f, 0.f), zb = za;
        if (b < nl) {
            const float2 zk = r[k * dm.B + b];
            const float2 zm = r[(k == 0 ? 0 : dm.L1 - k) * dm.B + b];
            za = make_float2(zk.x + zm.x, zk.y - zm.y);
            zb = make_float2(zk.y + zm.y, zm.x - zk.x);
        }
        sc[(2 * b) * dm.C + k] = za;
        sc[(2 * b + 1) * dm.C + k] = zb;
    }
    __syncthreads();
    if constexpr (D::is_static) {
        const OutScale ms{otf, d.pitch, D::nc};
        return static_fft<D::L0, D::C, kT, false, false, true>(sc, t.other(sc), t.tw0, Divisor{D::C, 0u}, ms, pre);
    } else {
        (void)otf;
        (void)pre;
        return cols_fft<false>(dm, d, t, sc);
    }
}

// The compile-time build's factors for the first round of the forward column transform's last
// pass: output t of butterfly (j, b), b = threadIdx.x mod C, is element j + t L0 / R of column b.
template <class D, int RL>
__device__ __forceinline__ void load_last_pass(const TileIter& d, const float2* __restrict__ otf, float2 (&pre)[RL]) {
    constexpr int nbf = (D::L0 / RL) * D::C;
    const int w = threadIdx.x;
    const int j = w / D::C, b = w - j * D::C;
#pragma unroll
    for (int t = 0; t < RL; t++) {
        pre[t] = make_float2(0.f, 0.f);
        if ((nbf >= kT || w < nbf) && b < D::nc) pre[t] = otf[(int64_t)(j + t * (D::L0 / RL)) * d.pitch + b];
    }
}

// The transfer-function values of this thread's spectrum elements ([L0][pitch] in global memory;
// conj(FFT(w)) * scale; zero on the padding column). Issued a transform ahead of their use, so
// that their latency hides behind it.
template <class D>
__device__ __forceinline__ void load_otf(const D& dm, const TileIter& d, const float2* __restrict__ otf,
                                         float2 (&w)[kE]) {
    const int n = dm.L0 * dm.C;
#pragma unroll
    for (int s = 0; s < kE; s++) {
        const int e = threadIdx.x + s * kT;
        w[s] = make_float2(0.f, 0.f);
        if (e < n) {
            const int r = dm.div_C(e), k = e - r * dm.C;
            if (k < dm.nc) w[s] = otf[(int64_t)r * d.pitch + k];
        }
    }
}

// sc *= w on the spectrum elements (column lines).
template <class D>
__device__ __forceinline__ void apply_otf(const D& dm, float2* sc, const float2 (&w)[kE]) {
    const int n = dm.L0 * dm.C;
#pragma unroll
    for (int s = 0; s < kE; s++) {
        const int e = threadIdx.x + s * kT;
        if (e < n) sc[e] = cmul(sc[e], w[s]);
    }
    __syncthreads();
}

// The spectrum in `sc` -> the real tile as row-pair lines; returns the buffer that holds it
// (unnormalised: the transfer function carries the 1 / (4 L0 L1)). Only the row pairs below nl,
// the ones the caller reads, are produced.
template <class D>
__device__ float2* inverse_2d(const D& dm, const TileIter& d, const TileLds& t, float2* sc, int nl) {
    const float2* r = cols_fft<true>(dm, d, t, sc);
    float2* zr = t.other(r);
    for (int e = threadIdx.x; e < nl * dm.nc; e += kT) {
        const int b = dm.div_nc(e), k = e - b * dm.nc;
        float2 A = r[(2 * b) * dm.C + k], Bv = r[(2 * b + 1) * dm.C + k];
        if (k == 0 || 2 * k == dm.L1) {
            A.y = 0.f;
            Bv.y = 0.f;
        }
        zr[k * dm.B + b] = make_float2(A.x - Bv.y, A.y + Bv.x);
        const int km = dm.L1 - k;
        if (k > 0 && km >= dm.nc) zr[km * dm.B + b] = make_float2(A.x + Bv.y, Bv.x - A.y);
    }
    __syncthreads();
    return rows_fft<true>(dm, d, t, zr, nl);
}

// A box of the image (rows y0 + r, columns x0 + c) lies inside it: no boundary rule applies.
__device__ __forceinline__ bool box_inside(int y0, int x0, int h, int w, int n0, int n1) {
    return y0 >= 0 && x0 >= 0 && y0 + h <= n0 && x0 + w <= n1;
}

// One iteration of output tile `tile` (row-major over the tiles1 columns of tiles) of array d.
template <int SL0, int SL1>
__device__ __forceinline__ void tile_body(const TileIter& d, int tile, float2* lds) {
    const Dims<SL0, SL1> dm(d);
    const int B = dm.B, nb = dm.nb;
    TileLds t;
    t.X = lds;
    t.Y = lds + dm.nbuf;
    float2* tw1 = t.Y + dm.nbuf;
    float2* tw0 = tw1 + dm.L1;
    t.tw1 = tw1;
    t.tw0 = tw0;
    const int tid = threadIdx.x;
    const int ty = tile / d.tiles1, tx = tile - ty * d.tiles1;
    const int t0y = ty * d.T0, t0x = tx * d.T1;
    // Tile sweeps: thread (pr, c) handles row pairs pr, pr + pstep, ... at column c (columns
    // fastest across lanes, so global rows are read whole and a pair is one 8-byte LDS access).
    const int pstep = kT / dm.L1;
    const int pr0 = tid / dm.L1;
    const int c = tid - pr0 * dm.L1;
    const bool active = pr0 < pstep;

    // Q1: the estimate with the pad rule, two halos deep, into X; the twiddle tables are staged
    // while its loads are in flight
    {
        const int gy0 = t0y - 2 * d.left0, gx0 = t0x - 2 * d.left1;
        const bool inside = box_inside(gy0, gx0, dm.L0, dm.L1, d.n0, d.n1);
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
        for (int i = tid; i < dm.L1; i += kT) tw1[i] = d.tw1[d.f1.tw + i];
        for (int i = tid; i < dm.L0; i += kT) tw0[i] = d.tw0[d.f0.tw + i];
        // Every position a transform reads holds a finite value: X's padding row line (the Q1
        // stores fill the other lines), and in both buffers the part of the column layout beyond
        // the row layout, which no row pass writes. Neither overlaps a Q1 store. The padding
        // lines' transforms are never read.
        if (B > nb)
            for (int i = tid; i < dm.L1; i += kT) t.X[i * B + nb] = make_float2(0.f, 0.f);
        for (int i = dm.L1 * B + tid; i < dm.L0 * dm.C; i += kT) {
            t.X[i] = make_float2(0.f, 0.f);
            t.Y[i] = make_float2(0.f, 0.f);
        }
#pragma unroll
        for (int s = 0; s < kG; s++) {
            const int b = pr0 + s * pstep;
            if (active && b < nb) t.X[c * B + b] = v[s];
        }
    }
    // the first transfer function's values, loaded while the tile is transformed
    constexpr bool kStatic = Dims<SL0, SL1>::is_static;
    constexpr int kRL = kStatic ? static_last_radix(SL0) : 1;
    float2 w[kE];     // run-time build: this thread's elements for apply_otf
    float2 pre[kRL];  // compile-time build: the first round of the last column pass
    if constexpr (kStatic) load_last_pass<Dims<SL0, SL1>, kRL>(d, d.otf1, pre);
    else load_otf(dm, d, d.otf1, w);
    __syncthreads();
    float2* sp = forward_2d(dm, d, t, t.X, nb, d.otf1, pre);
    if constexpr (!kStatic) apply_otf(dm, sp, w);

    // Q2 = P(rel) on the rel window (my x mx: the tile's own outputs, fewer than T on the last tile
    // of an axis, and the PSF's reach), 0 beyond. The observed values and the second transfer
    // function are loaded before the inverse transform that gives u, whose latency they hide behind.
    const int my = min(d.T0, d.n0 - t0y) + d.k0 - 1, mx = min(d.T1, d.n1 - t0x) + d.k1 - 1;
    const int qy0 = t0y - d.left0, qx0 = t0x - d.left1;
    const bool q_inside = box_inside(qy0, qx0, my, mx, d.n0, d.n1);
    float2 v[kG];
    int at[2 * kG];  // border tiles: the float offset in U of each value's pad target, or -1
    if (q_inside) {
        // every target is its own position: y of the pair, rel in place
#pragma unroll
        for (int s = 0; s < kG; s++) {
            const int b = pr0 + s * pstep;
            v[s] = make_float2(0.f, 0.f);
            if (active && b < nb && c < mx) {
                const float* p = d.obs + (int64_t)(qy0 + 2 * b) * d.obs_rs + qx0 + c;
                if (2 * b < my) v[s].x = p[0];
                if (2 * b + 1 < my) v[s].y = p[d.obs_rs];
            }
        }
    } else {
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
                        const float y = d.obs[(int64_t)sy * d.obs_rs + sx];
                        if (q == 0) v[s].x = y; else v[s].y = y;
                        const int ry = sy - qy0;
                        at[2 * s + q] = 2 * ((sx - qx0) * B + (ry >> 1)) + (ry & 1);
                    }
                }
            }
        }
    }
    if constexpr (kStatic) load_last_pass<Dims<SL0, SL1>, kRL>(d, d.otf2, pre);
    else load_otf(dm, d, d.otf2, w);
    const int nq = (my + 1) / 2;  // row pairs of the rel window
    const float2* U = inverse_2d(dm, d, t, sp, nq);

    // U holds u(t0 - left + p) at p; Q2 goes into the other buffer (its last reader, the final
    // inverse pass, is behind a barrier).
    float2* Q = t.other(U);
    {
        if (q_inside) {
#pragma unroll
            for (int s = 0; s < kG; s++) {
                const int b = pr0 + s * pstep;
                float2 u = make_float2(1.f, 1.f);
                if (active && b < nq && c < mx) u = U[c * B + b];
                v[s].x = div_fast(v[s].x, fabsf(u.x) < d.eps ? d.eps : u.x);
                v[s].y = div_fast(v[s].y, fabsf(u.y) < d.eps ? d.eps : u.y);
            }
        } else {
            const float* Uf = reinterpret_cast<const float*>(U);
#pragma unroll
            for (int s = 0; s < kG; s++) {
                if (at[2 * s] >= 0) {
                    const float u = Uf[at[2 * s]];
                    v[s].x = div_fast(v[s].x, fabsf(u) < d.eps ? d.eps : u);
                }
                if (at[2 * s + 1] >= 0) {
                    const float u = Uf[at[2 * s + 1]];
                    v[s].y = div_fast(v[s].y, fabsf(u) < d.eps ? d.eps : u);
                }
            }
        }
#pragma unroll
        for (int s = 0; s < kG; s++) {
            const int b = pr0 + s * pstep;
            if (active && b < nb) Q[c * B + b] = v[s];
        }
    }
    __syncthreads();
    sp = forward_2d(dm, d, t, Q, nq, d.otf2, pre);
    if constexpr (!kStatic) apply_otf(dm, sp, w);

    // the estimate on the tile's own pixels, loaded before the last inverse transform
    const int h = min(d.T0, d.n0 - t0y), wd = min(d.T1, d.n1 - t0x);
    float2 x[kG];
#pragma unroll
    for (int s = 0; s < kG; s++) {
        const int b = pr0 + s * pstep;
        x[s] = make_float2(d.init, d.init);
        if (!d.first && active && 2 * b < h && c < wd) {
            const float* p = d.est_in + (int64_t)(t0y + 2 * b) * d.n1 + t0x + c;
            x[s].x = p[0];
            if (2 * b + 1 < h) x[s].y = p[d.n1];
        }
    }
    const float2* Cv = inverse_2d(dm, d, t, sp, (h + 1) / 2);

    // Cv holds c(t0 + p) at p: the update on the tile's own pixels
    {
#pragma unroll
        for (int s = 0; s < kG; s++) {
            const int b = pr0 + s * pstep;
            if (active && 2 * b < h && c < wd) {
                const float2 cv = Cv[c * B + b];
                float xa = x[s].x * cv.x, xb = x[s].y * cv.y;
                if (d.last && d.clip) {
                    xa = fminf(fmaxf(xa, -1.f), 1.f);
                    xb = fminf(fmaxf(xb, -1.f), 1.f);
                }
                const int gy = t0y + 2 * b, gx = t0x + c;
                if (!d.last) {
                    float* p = d.est_out + (int64_t)gy * d.n1 + gx;
                    p[0] = xa;
                    if (2 * b + 1 < h) p[d.n1] = xb;
                } else if (gx >= d.own0x && gx < d.own1x) {
                    // the output window (the whole array unless the caller stores a part of it)
                    float* p = d.out + (int64_t)(gy - d.own0y) * d.out_rs + (gx - d.own0x);
                    if (gy >= d.own0y && gy < d.own1y) p[0] = xa;
                    if (2 * b + 1 < h && gy + 1 >= d.own0y && gy + 1 < d.own1y) p[d.out_rs] = xb;
                }
            }
        }
    }
}

template <int SL0, int SL1>
__global__ __launch_bounds__(kT) __attribute__((amdgpu_waves_per_eu(kW))) void tile_iter_kernel(TileIter d) {
    extern __shared__ float2 lds[];
    tile_body<SL0, SL1>(d, blockIdx.x, lds);
}

template <int SL0, int SL1>
__global__ __launch_bounds__(kT) __attribute__((amdgpu_waves_per_eu(kW))) void tile_batch_kernel(TileBatch bt) {
    extern __shared__ float2 lds[];
    const int2 item = bt.items[blockIdx.x];
    TileIter d = bt.jobs[__builtin_amdgcn_readfirstlane(item.x)];  // a private copy
    d.first = bt.first;
    d.last = bt.last;
    d.clip = bt.clip;
    if (bt.swap) {
        float* t = const_cast<float*>(d.est_in);
        d.est_in = d.est_out;
        d.est_out = t;
    }
    tile_body<SL0, SL1>(d, __builtin_amdgcn_readfirstlane(item.y), lds);
}

// The square domains compiled with their lengths as constants.
#define LUCEM_TILE_STATIC(X) X(32) X(48) X(64) X(80) X(96) X(112) X(120) X(128)
}  // namespace

bool tile_len_static(int L0, int L1) {
    if (L0 != L1) return false;
#define LUCEM_CASE(n) \
    if (L0 == n) return true;
    LUCEM_TILE_STATIC(LUCEM_CASE)
#undef LUCEM_CASE
    return false;
}

size_t tile_iter_lds_bytes(int L0, int L1) {
    return sizeof(float2) * (2 * (size_t)tile_buffer_elems(L0, L1) + (size_t)L1 + (size_t)L0);
}

int allow_tile_lds() {
    const int bytes = (int)tile_iter_lds_bytes(kTileMaxLen, kTileMaxLen);
    int err = (int)hipFuncSetAttribute(reinterpret_cast<const void*>(tile_iter_kernel<0, 0>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
#define LUCEM_CASE(n)                                                                         \
    if (!err)                                                                                 \
        err = (int)hipFuncSetAttribute(reinterpret_cast<const void*>(tile_iter_kernel<n, n>), \
                                       hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    LUCEM_TILE_STATIC(LUCEM_CASE)
#undef LUCEM_CASE
    if (!err)
        err = (int)hipFuncSetAttribute(reinterpret_cast<const void*>(tile_batch_kernel<0, 0>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
#define LUCEM_CASE(n)                                                                          \
    if (!err)                                                                                  \
        err = (int)hipFuncSetAttribute(reinterpret_cast<const void*>(tile_batch_kernel<n, n>), \
                                       hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    LUCEM_TILE_STATIC(LUCEM_CASE)
#undef LUCEM_CASE
    return err;
}

int launch_tile_iter(const TileIter& d, int blocks, void* stream) {
    const size_t lds = tile_iter_lds_bytes(d.L0, d.L1);
    const hipStream_t s = (hipStream_t)stream;
    if (d.generic || !tile_len_static(d.L0, d.L1)) {
        hipLaunchKernelGGL((tile_iter_kernel<0, 0>), dim3(blocks), dim3(kT), lds, s, d);
        return (int)hipGetLastError();
    }
#define LUCEM_CASE(n)                                                                     \
    if (d.L0 == n) {                                                                      \
        hipLaunchKernelGGL((tile_iter_kernel<n, n>), dim3(blocks), dim3(kT), lds, s, d); \
        return (int)hipGetLastError();                                                    \
    }
    LUCEM_TILE_STATIC(LUCEM_CASE)
#undef LUCEM_CASE
    return (int)hipErrorInvalidValue;
}

int launch_tile_batch(const TileBatch& b, void* stream) {
    if (b.nitems <= 0) return 0;
    const size_t lds = tile_iter_lds_bytes(b.L0, b.L1);
    const hipStream_t s = (hipStream_t)stream;
#define LUCEM_CASE(n)                                                                             \
    if (b.L0 == n && b.L1 == n) {                                                                 \
        hipLaunchKernelGGL((tile_batch_kernel<n, n>), dim3(b.nitems), dim3(kT), lds, s, b); \
        return (int)hipGetLastError();                                                            \
    }
    LUCEM_TILE_STATIC(LUCEM_CASE)
#undef LUCEM_CASE
    hipLaunchKernelGGL((tile_batch_kernel<0, 0>), dim3(b.nitems), dim3(kT), lds, s, b);
    return (int)hipGetLastError();
}

}  // namespace lucem
