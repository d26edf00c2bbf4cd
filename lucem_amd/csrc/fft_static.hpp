// LDS FFT passes with the length, the line count and the thread count known at compile time.
//
// The same Stockham passes as fft_lds.hpp (B interleaved lines, element i of line b at
// buf[i * B + b]; radix R = R1 * R2 computed in registers), but out of place between two LDS
// buffers, one barrier per pass, and every stride a constant: loads and stores take immediate LDS
// offsets, the line / butterfly split is a multiply by a constant, and the radix sequence is fixed
// by the length, so a pass costs its arithmetic and little else. The tile kernel instantiates it
// for its common transform lengths.
#pragma once

#include "fft_lds.hpp"

namespace lucem {

// Radix of the first pass over a remaining length: the largest of 16, 12, 10, 9, 8, 7, 6, 5, 4, 3, 2
// that divides it (the composite radices compute two stages per LDS round trip).
constexpr int static_radix(int rem) {
    return rem % 16 == 0 ? 16
           : rem % 12 == 0 ? 12
           : rem % 10 == 0 ? 10
           : rem % 9 == 0  ? 9
           : rem % 8 == 0  ? 8
           : rem % 7 == 0  ? 7
           : rem % 6 == 0  ? 6
           : rem % 5 == 0  ? 5
           : rem % 4 == 0  ? 4
           : rem % 3 == 0  ? 3
           : rem % 2 == 0  ? 2
                           : 1;
}
constexpr bool static_len_ok(int n) { return n == 1 ? true : static_radix(n) > 1 && static_len_ok(n / static_radix(n)); }
// The radix of a length-n transform's last pass.
constexpr int static_last_radix(int n) { return n / static_radix(n) == 1 ? n : static_last_radix(n / static_radix(n)); }

// A factor per output element of a transform's last pass, read from a [row][pitch] table in global
// memory (element i of line b takes tab[i * pitch + b]; lines from nl on take 0). The first round
// of butterflies takes its factors from registers loaded beforehand (pre[t] for output t of
// butterfly threadIdx.x); later rounds, when there are any, read the table.
struct OutScale {
    const float2* tab = nullptr;
    int pitch = 0;
    int nl = 0;
};

template <int R, bool INV>
struct StaticButterfly;
template <bool INV> struct StaticButterfly<16, INV> : Composite<4, 4, INV> {};
template <bool INV> struct StaticButterfly<12, INV> : Composite<4, 3, INV> {};
template <bool INV> struct StaticButterfly<10, INV> : Composite<5, 2, INV> {};
template <bool INV> struct StaticButterfly<9, INV> : Composite<3, 3, INV> {};
template <bool INV> struct StaticButterfly<6, INV> : Composite<3, 2, INV> {};
template <bool INV> struct StaticButterfly<8, INV> : Composite<8, 1, INV> {};
template <bool INV> struct StaticButterfly<7, INV> : Composite<7, 1, INV> {};
template <bool INV> struct StaticButterfly<5, INV> : Composite<5, 1, INV> {};
template <bool INV> struct StaticButterfly<4, INV> : Composite<4, 1, INV> {};
template <bool INV> struct StaticButterfly<3, INV> : Composite<3, 1, INV> {};
template <bool INV> struct StaticButterfly<2, INV> : Composite<2, 1, INV> {};

// One pass: butterfly u = j * B + b (j < N / R, line b) takes elements u + t * (N / R) * B,
// multiplies element t by w^(t k), k = j mod Ns, w = exp(-+2 pi i / (Ns R)), and writes element s
// of its DFT to (j - k) * R * B + b + (k + s * Ns) * B. Twiddles from tw[m] = exp(-2 pi i m / N).
// With PRUNE only the lines b < nl are transformed: the butterflies are numbered w = j * nl + b,
// so that the threads past the last one idle as whole waves.
template <int N, int B, int T, bool INV, int Ns, int R, bool PRUNE, bool MUL>
__device__ __forceinline__ void static_pass(const float2* __restrict__ src, float2* __restrict__ dst,
                                            const float2* __restrict__ tw, const Divisor nl, const OutScale& ms,
                                            const float2* pre) {
    static_assert(!MUL || (Ns * R == N && !PRUNE), "output factors apply to a whole last pass");
    constexpr int nbf = (N / R) * B;
    constexpr int S = (nbf + T - 1) / T;
    constexpr int tws = N / (Ns * R);
    typedef const __attribute__((address_space(3))) lds_vec2* csrc_t;
    typedef __attribute__((address_space(3))) lds_vec2* cdst_t;
    const csrc_t sp = (csrc_t)(size_t)src;
    const cdst_t dp = (cdst_t)(size_t)dst;
    const csrc_t tp = (csrc_t)(size_t)tw;
    const int nact = PRUNE ? (N / R) * nl.d : nbf;
#pragma unroll
    for (int s = 0; s < S; s++) {
        const int w = (int)threadIdx.x + s * T;
        if (PRUNE ? w < nact : (nbf % T == 0 || w < nbf)) {
            int u, j;
            if constexpr (PRUNE) {
                j = nl.div(w);
                u = j * B + (w - j * nl.d);
            } else {
                u = w;
                j = u / B;
            }
            float2 v[R];
#pragma unroll
            for (int t = 0; t < R; t++) {
                const lds_vec2 x = sp[u + t * nbf];
                v[t] = make_float2(x.x, x.y);
            }
            const int k = Ns > 1 ? j % Ns : 0;
            if constexpr (Ns > 1) {
                const int step = k * tws;
#pragma unroll
                for (int t = 1; t < R; t++) {
                    const lds_vec2 x = tp[t * step];
                    v[t] = cmul(v[t], make_float2(x.x, INV ? -x.y : x.y));
                }
            }
            StaticButterfly<R, INV>::run(v);
            if constexpr (MUL) {
                // the last pass writes output t of butterfly (j, b) to element j + t N / R of line b
                const int b = u - j * B;
#pragma unroll
                for (int t = 0; t < R; t++) {
                    float2 f;
                    if (s == 0)
                        f = pre[t];
                    else
                        f = b < ms.nl ? ms.tab[(int64_t)(j + t * (N / R)) * ms.pitch + b] : make_float2(0.f, 0.f);
                    v[t] = cmul(v[t], f);
                }
            }
            const int o = u + (j - k) * (R - 1) * B;
#pragma unroll
            for (int t = 0; t < R; t++) {
                lds_vec2 x;
                x.x = v[t].x;
                x.y = v[t].y;
                dp[o + t * Ns * B] = x;
            }
        }
    }
}

// All passes of a length-N transform of the B lines in `a`, alternating with `b`; returns the
// buffer that holds the result. Each pass ends with a barrier. With PRUNE only the lines below
// nl.d are transformed; the others of the result buffer keep what they held. With MUL the last
// pass multiplies its outputs by the factors of `ms` (the first round's from `pre`).
template <int N, int B, int T, bool INV, bool PRUNE = false, bool MUL = false, int Ns = 1>
__device__ __forceinline__ float2* static_fft(float2* a, float2* b, const float2* tw, const Divisor nl = Divisor{B, 0u},
                                              const OutScale& ms = OutScale{}, const float2* pre = nullptr) {
    static_assert(static_len_ok(N), "transform length must be 7-smooth");
    if constexpr (Ns >= N) {
        (void)b;
        (void)tw;
        (void)nl;
        (void)ms;
        (void)pre;
        return a;
    } else {
        constexpr int R = static_radix(N / Ns);
        static_pass<N, B, T, INV, Ns, R, PRUNE, MUL && Ns * R == N>(a, b, tw, nl, ms, pre);
        __syncthreads();
        return static_fft<N, B, T, INV, PRUNE, MUL, Ns * R>(b, a, tw, nl, ms, pre);
    }
}

// In-place form of the same passes, for a workgroup that holds one buffer of B lines: each pass
// reads every butterfly's inputs into registers, computes them, and writes after a barrier (two
// barriers per pass). Returns nothing: the result is in `a`.
template <int N, int B, int T, bool INV, int Ns, int R>
__device__ __forceinline__ void static_pass_inplace(float2* __restrict__ a, const float2* __restrict__ tw) {
    constexpr int nbf = (N / R) * B;
    constexpr int S = (nbf + T - 1) / T;
    constexpr int tws = N / (Ns * R);
    typedef __attribute__((address_space(3))) lds_vec2* lds_t;
    typedef const __attribute__((address_space(3))) lds_vec2* clds_t;
    const lds_t ap = (lds_t)(size_t)a;
    const clds_t tp = (clds_t)(size_t)tw;
    float2 v[S][R];
#pragma unroll
    for (int s = 0; s < S; s++) {
        const int u = (int)threadIdx.x + s * T;
        if (nbf % T == 0 || u < nbf) {
            const int j = u / B;
#pragma unroll
            for (int t = 0; t < R; t++) {
                const lds_vec2 x = ap[u + t * nbf];
                v[s][t] = make_float2(x.x, x.y);
            }
            if constexpr (Ns > 1) {
                const int step = (j % Ns) * tws;
#pragma unroll
                for (int t = 1; t < R; t++) {
                    const lds_vec2 x = tp[t * step];
                    v[s][t] = cmul(v[s][t], make_float2(x.x, INV ? -x.y : x.y));
                }
            }
            StaticButterfly<R, INV>::run(v[s]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < S; s++) {
        const int u = (int)threadIdx.x + s * T;
        if (nbf % T == 0 || u < nbf) {
            const int j = u / B;
            const int k = Ns > 1 ? j % Ns : 0;
            const int o = u + (j - k) * (R - 1) * B;
#pragma unroll
            for (int t = 0; t < R; t++) {
                lds_vec2 x;
                x.x = v[s][t].x;
                x.y = v[s][t].y;
                ap[o + t * Ns * B] = x;
            }
        }
    }
    __syncthreads();
}

template <int N, int B, int T, bool INV, int Ns = 1>
__device__ __forceinline__ void static_fft_inplace(float2* a, const float2* tw) {
    static_assert(static_len_ok(N), "transform length must be 7-smooth");
    if constexpr (Ns < N) {
        constexpr int R = static_radix(N / Ns);
        static_pass_inplace<N, B, T, INV, Ns, R>(a, tw);
        static_fft_inplace<N, B, T, INV, Ns * R>(a, tw);
    }
}

}  // namespace lucem
