// LDS-resident mixed-radix FFT for gfx950 (wave64).
//
// B lines of length N are stored interleaved in LDS, element i of line b at buf[i*B + b], so
// consecutive lanes touch consecutive LDS words in every pass. Each Stockham pass of radix
// R = R1*R2 loads all of its butterfly inputs into registers, waits at a barrier, then computes
// and writes in place: a pass needs N*B/threads complex values per thread of registers and no
// second LDS buffer. A composite radix R1*R2 is computed in registers as R2 butterflies of size
// R1, internal twiddles, then R1 butterflies of size R2, so two radix stages cost one LDS round
// trip. Inter-pass twiddles come from a table tw[m] = exp(-2*pi*i*m/N) built on the host in
// double precision and rounded once to float; intra-butterfly constants are evaluated at compile
// time from the Taylor series of sin and cos in double precision.
#pragma once

#include <hip/hip_runtime.h>

#include "rl_internal.hpp"

namespace lucem {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// -i * a (forward) or +i * a (inverse)
template <bool INV>
__device__ __forceinline__ float2 mul_mi(float2 a) {
    return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

// a / b from the hardware reciprocal with one Newton correction (the residual a - q b is exact
// with an FMA): within an ulp of the IEEE quotient for the finite, normal divisors the kernels
// divide by (the eps floor keeps them away from 0), at a third of its instructions.
__device__ __forceinline__ float div_fast(float a, float b) {
    const float r = __builtin_amdgcn_rcpf(b);
    const float q = a * r;
    return fmaf(fmaf(-q, b, a), r, q);
}

// Compile-time sin/cos of 2*pi*m/R (Taylor series on [-pi, pi], double precision).
constexpr double kPi = 3.14159265358979323846264338327950288;
constexpr double taylor_sin(double x) {
    double term = x, sum = x;
    for (int n = 1; n < 40; ++n) {
        term *= -x * x / ((2.0 * n) * (2.0 * n + 1.0));
        sum += term;
    }
    return sum;
}
constexpr double taylor_cos(double x) {
    double term = 1.0, sum = 1.0;
    for (int n = 1; n < 40; ++n) {
        term *= -x * x / ((2.0 * n - 1.0) * (2.0 * n));
        sum += term;
    }
    return sum;
}
constexpr double unit_angle(int m, int R) {
    double a = 2.0 * kPi * (double)(m % R) / (double)R;
    return a > kPi ? a - 2.0 * kPi : a;
}
// exp(-+2*pi*i*m/R) as float constants
template <int M, int R, bool INV>
struct Twiddle {
    static constexpr float re = (float)taylor_cos(unit_angle(M, R));
    static constexpr float im = (float)(INV ? taylor_sin(unit_angle(M, R)) : -taylor_sin(unit_angle(M, R)));
};

// Radix butterflies: v[s] <- sum_t v[t] * exp(-+2*pi*i*t*s/R).
template <int R, bool INV>
struct Butterfly;

template <bool INV>
struct Butterfly<1, INV> {
    __device__ __forceinline__ static void run(float2*) {}
};

template <bool INV>
struct Butterfly<2, INV> {
    __device__ __forceinline__ static void run(float2* v) {
        const float2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};

template <bool INV>
struct Butterfly<4, INV> {
    __device__ __forceinline__ static void run(float2* v) {
        const float2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
        const float2 t2 = cadd(v[1], v[3]), t3 = mul_mi<INV>(csub(v[1], v[3]));
        v[0] = cadd(t0, t2);
        v[2] = csub(t0, t2);
        v[1] = cadd(t1, t3);
        v[3] = csub(t1, t3);
    }
};

template <bool INV>
struct Butterfly<8, INV> {
    __device__ __forceinline__ static void run(float2* v) {
        constexpr float c = 0.7071067811865476f;  // sqrt(1/2)
        float2 e[4] = {v[0], v[2], v[4], v[6]};
        float2 o[4] = {v[1], v[3], v[5], v[7]};
        Butterfly<4, INV>::run(e);
        Butterfly<4, INV>::run(o);
        const float s1 = INV ? c : -c;
        o[1] = make_float2(c * o[1].x - s1 * o[1].y, c * o[1].y + s1 * o[1].x);
        o[2] = mul_mi<INV>(o[2]);
        o[3] = make_float2(-c * o[3].x - s1 * o[3].y, -c * o[3].y + s1 * o[3].x);
#pragma unroll
        for (int s = 0; s < 4; s++) {
            v[s] = cadd(e[s], o[s]);
            v[s + 4] = csub(e[s], o[s]);
        }
    }
};

template <bool INV>
struct Butterfly<3, INV> {
    __device__ __forceinline__ static void run(float2* v) {
        constexpr float s3 = 0.8660254037844387f;  // sin(2*pi/3)
        const float2 t = cadd(v[1], v[2]);
        const float2 d = cscale(mul_mi<INV>(csub(v[1], v[2])), s3);
        const float2 m = make_float2(v[0].x - 0.5f * t.x, v[0].y - 0.5f * t.y);
        v[0] = cadd(v[0], t);
        v[1] = cadd(m, d);
        v[2] = csub(m, d);
    }
};

template <bool INV>
struct Butterfly<5, INV> {
    __device__ __forceinline__ static void run(float2* v) {
        constexpr float c1 = 0.30901699437494745f, c2 = -0.8090169943749473f;  // cos(2pi/5), cos(4pi/5)
        constexpr float s1 = 0.9510565162951535f, s2 = 0.5877852522924732f;    // sin(2pi/5), sin(4pi/5)
        const float2 a0 = v[0];
        const float2 t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
        const float2 d1 = csub(v[1], v[4]), d2 = csub(v[2], v[3]);
        const float2 m1 = make_float2(a0.x + c1 * t1.x + c2 * t2.x, a0.y + c1 * t1.y + c2 * t2.y);
        const float2 m2 = make_float2(a0.x + c2 * t1.x + c1 * t2.x, a0.y + c2 * t1.y + c1 * t2.y);
        const float2 n1 = mul_mi<INV>(make_float2(s1 * d1.x + s2 * d2.x, s1 * d1.y + s2 * d2.y));
        const float2 n2 = mul_mi<INV>(make_float2(s2 * d1.x - s1 * d2.x, s2 * d1.y - s1 * d2.y));
        v[0] = cadd(a0, cadd(t1, t2));
        v[1] = cadd(m1, n1);
        v[4] = csub(m1, n1);
        v[2] = cadd(m2, n2);
        v[3] = csub(m2, n2);
    }
};

template <bool INV>
struct Butterfly<7, INV> {
    __device__ __forceinline__ static void run(float2* v) {
        constexpr float C1 = 0.6234898018587336f, C2 = -0.22252093395631434f, C3 = -0.900968867902419f;
        constexpr float S1 = 0.7818314824680298f, S2 = 0.9749279121818236f, S3 = 0.43388373911755823f;
        const float2 a0 = v[0];
        const float2 t1 = cadd(v[1], v[6]), t2 = cadd(v[2], v[5]), t3 = cadd(v[3], v[4]);
        const float2 d1 = csub(v[1], v[6]), d2 = csub(v[2], v[5]), d3 = csub(v[3], v[4]);
        const float2 m1 = make_float2(a0.x + C1 * t1.x + C2 * t2.x + C3 * t3.x, a0.y + C1 * t1.y + C2 * t2.y + C3 * t3.y);
        const float2 m2 = make_float2(a0.x + C2 * t1.x + C3 * t2.x + C1 * t3.x, a0.y + C2 * t1.y + C3 * t2.y + C1 * t3.y);
        const float2 m3 = make_float2(a0.x + C3 * t1.x + C1 * t2.x + C2 * t3.x, a0.y + C3 * t1.y + C1 * t2.y + C2 * t3.y);
        const float2 n1 = mul_mi<INV>(make_float2(S1 * d1.x + S2 * d2.x + S3 * d3.x, S1 * d1.y + S2 * d2.y + S3 * d3.y));
        const float2 n2 = mul_mi<INV>(make_float2(S2 * d1.x - S3 * d2.x - S1 * d3.x, S2 * d1.y - S3 * d2.y - S1 * d3.y));
        const float2 n3 = mul_mi<INV>(make_float2(S3 * d1.x - S1 * d2.x + S2 * d3.x, S3 * d1.y - S1 * d2.y + S2 * d3.y));
        v[0] = cadd(a0, cadd(t1, cadd(t2, t3)));
        v[1] = cadd(m1, n1);
        v[6] = csub(m1, n1);
        v[2] = cadd(m2, n2);
        v[5] = csub(m2, n2);
        v[3] = cadd(m3, n3);
        v[4] = csub(m3, n3);
    }
};

// Multiplies a by the compile-time constant exp(-+2*pi*i*M/R).
template <int M, int R, bool INV>
__device__ __forceinline__ float2 mul_const(float2 a) {
    if constexpr (M % R == 0) {
        return a;
    } else {
        constexpr float c = Twiddle<M, R, INV>::re, s = Twiddle<M, R, INV>::im;
        return make_float2(a.x * c - a.y * s, a.x * s + a.y * c);
    }
}

// DFT of size R1*R2 in registers. Input t = R2*t1 + t2, output s = s1 + R1*s2.
template <int R1, int R2, bool INV>
struct Composite {
    template <int T2, int S1>
    __device__ __forceinline__ static void twiddle_row(float2* a) {
        if constexpr (S1 < R1) {
            a[S1 * R2 + T2] = mul_const<T2 * S1, R1 * R2, INV>(a[S1 * R2 + T2]);
            twiddle_row<T2, S1 + 1>(a);
        }
    }
    template <int T2>
    __device__ __forceinline__ static void stage_a(const float2* v, float2* a) {
        if constexpr (T2 < R2) {
            float2 col[R1];
#pragma unroll
            for (int t1 = 0; t1 < R1; t1++) col[t1] = v[R2 * t1 + T2];
            Butterfly<R1, INV>::run(col);
#pragma unroll
            for (int s1 = 0; s1 < R1; s1++) a[s1 * R2 + T2] = col[s1];
            twiddle_row<T2, 1>(a);
            stage_a<T2 + 1>(v, a);
        }
    }
    __device__ __forceinline__ static void run(float2* v) {
        if constexpr (R2 == 1) {
            Butterfly<R1, INV>::run(v);
        } else {
            float2 a[R1 * R2];
            stage_a<0>(v, a);
#pragma unroll
            for (int s1 = 0; s1 < R1; s1++) {
                float2 row[R2];
#pragma unroll
                for (int t2 = 0; t2 < R2; t2++) row[t2] = a[s1 * R2 + t2];
                Butterfly<R2, INV>::run(row);
#pragma unroll
                for (int s2 = 0; s2 < R2; s2++) v[s1 + R1 * s2] = row[s2];
            }
        }
    }
};

// A twiddle from the LDS-staged table (TWL) or from global memory, with the address space
// explicit so that the non-inlined passes issue ds_read / global_load rather than flat loads.
template <bool TWL>
__device__ __forceinline__ float2 load_twiddle(const float2* tw, int i) {
    const float* f = reinterpret_cast<const float*>(tw);
    if constexpr (TWL) {
        typedef const __attribute__((address_space(3))) float* lds_ptr;
        const lds_ptr q = (lds_ptr)f;
        return make_float2(q[2 * i], q[2 * i + 1]);
    } else {
        typedef const __attribute__((address_space(1))) float* global_ptr;
        const global_ptr q = (global_ptr)f;
        return make_float2(q[2 * i], q[2 * i + 1]);
    }
}

// Exact division by a small positive divisor d through a multiply-high: q = u / d for every
// 0 <= u < 2^17 and d <= 16384 (m = floor((2^32 - 1) / d) + 1, checked exhaustively on the host).
struct Divisor {
    int d;
    unsigned m;
    __device__ __forceinline__ static Divisor of(int d) { return Divisor{d, d > 1 ? 0xFFFFFFFFu / (unsigned)d + 1u : 0u}; }
    __device__ __forceinline__ int div(int u) const { return d > 1 ? (int)__umulhi((unsigned)u, m) : u; }
};

// The arguments of a non-inlined call arrive in VGPRs even when every lane holds the same value.
// Reading them back with readfirstlane puts them in SGPRs, so the index arithmetic on them runs on
// the scalar unit and the branches on them are scalar branches.
__device__ __forceinline__ int uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ unsigned uniform(unsigned x) { return (unsigned)__builtin_amdgcn_readfirstlane((int)x); }
// An empty asm statement that the compiler cannot see through: a scalar offset kept in an SGPR as
// computed (not folded back into a per-lane multiply-add), or a per-lane running sum kept as adds.
__device__ __forceinline__ int in_sgpr(int x) {
    asm("" : "+s"(x));
    return x;
}
__device__ __forceinline__ int in_vgpr(int x) {
    asm("" : "+v"(x));
    return x;
}
// Product of two values below 2^24 on the 24-bit multiplier (full rate).
__device__ __forceinline__ int mul24(int a, int b) { return (int)(((unsigned)a & 0xFFFFFFu) * ((unsigned)b & 0xFFFFFFu)); }
template <class P>
__device__ __forceinline__ P* uniform(P* p) {
    const unsigned long long v = (unsigned long long)p;
    const unsigned long long lo = uniform((unsigned)v), hi = uniform((unsigned)(v >> 32));
    return (P*)((hi << 32) | lo);
}

// One in-place Stockham pass of radix R = R1*R2 over B interleaved lines of length N (B is any
// count >= 1). Input element j + t*N/R of butterfly j is multiplied by w^(t*k), k = j mod Ns,
// w = exp(-+2*pi*i/(Ns*R)), transformed, and written to (j - k)*R + k + t*Ns. With the
// butterfly index u = j*B + b, the inputs sit at u + t*(N/R)*B and the outputs at
// u + (j - k)*(R - 1)*B + t*Ns*B.
// W: waves per SIMD of the calling kernels. Each W gets its own copy of the pass, compiled to
// that register budget (a copy shared by kernels of different budgets gets the smallest).
// Ns = the product of the earlier radices, ns_mag its multiply-high reciprocal (0 when Ns = 1) and
// rem = N / Ns, so that no pass divides by a value known only at run time.
// Index arithmetic: the per-line strides (N/R)*B, Ns*B and (R - 1)*B are uniform (SGPRs); the
// per-lane products take the 24-bit multiplier (every index is below 2^17), and the line data is
// addressed as LDS, so that each access costs one vector add.
typedef float lds_vec2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(3))) lds_vec2 lds_c;
__device__ __forceinline__ float2 lds_get(const lds_c* p, int i) {
    const lds_vec2 x = p[i];
    return make_float2(x.x, x.y);
}
__device__ __forceinline__ void lds_put(lds_c* p, int i, float2 v) {
    lds_vec2 x;
    x.x = v.x;
    x.y = v.y;
    p[i] = x;
}

// OOP: out of place, reading buf_ and writing out_ (another LDS buffer), with no barrier between
// the loads and the stores: a wave stores as soon as its own butterflies are done.
template <int R1, int R2, bool INV, int T, int E, bool TWL, int W, bool OOP = false>
__device__ __noinline__ void stockham_pass(float2* __restrict__ buf_, float2* __restrict__ out_, int N_, Divisor lines_,
                                           int Ns_, unsigned ns_mag_, int rem_, const float2* __restrict__ tw_) {
    constexpr int R = R1 * R2;
    constexpr int S = (E + R - 1) / R;
    lds_c* const buf = (lds_c*)(size_t)uniform((unsigned)(size_t)(lds_c*)buf_);
    lds_c* const out = OOP ? (lds_c*)(size_t)uniform((unsigned)(size_t)(lds_c*)out_) : buf;
    const float2* const tw = uniform(tw_);
    const int N = uniform(N_);
    const Divisor lines{uniform(lines_.d), uniform(lines_.m)};
    const int Ns = uniform(Ns_);
    const unsigned ns_mag = uniform(ns_mag_);
    const int B = lines.d;
    const int nbf = (N / R) * B;
    const int tws = uniform(rem_) / R;  // N / (Ns * R)
    const int ostride = Ns * B;
    const int jstride = (R - 1) * B;
    float2 v[S][R];
#pragma unroll
    for (int s = 0; s < S; s++) {
        const int u = threadIdx.x + s * T;
        if (u < nbf) {
#pragma unroll
            for (int t = 0; t < R; t++) v[s][t] = lds_get(buf, u + in_sgpr(t * nbf));
        }
    }
    if (!OOP) __syncthreads();
#pragma unroll
    for (int s = 0; s < S; s++) {
        const int u = threadIdx.x + s * T;
        if (u < nbf) {
            const int j = lines.div(u);
            const int k = Ns > 1 ? j - mul24((int)__umulhi((unsigned)j, ns_mag), Ns) : 0;
            if (Ns > 1) {
                const int step = mul24(k, tws);
                int widx = 0;
#pragma unroll
                for (int t = 1; t < R; t++) {
                    widx = in_vgpr(widx + step);
                    float2 w = load_twiddle<TWL>(tw, widx);
                    if (INV) w.y = -w.y;
                    v[s][t] = cmul(v[s][t], w);
                }
            }
            Composite<R1, R2, INV>::run(v[s]);
            const int o = u + mul24(j - k, jstride);
#pragma unroll
            for (int t = 0; t < R; t++) lds_put(out, o + in_sgpr(t * ostride), v[s][t]);
        }
    }
    __syncthreads();
}

// Radix codes of a pass: R1 in the low nibble, R2 in the high nibble (R2 = 1 for a simple pass).
constexpr unsigned char radix_code(int r1, int r2) { return (unsigned char)(r1 | (r2 << 4)); }

// Twiddle tables of lengths up to kTwLdsMax are staged in LDS behind the line data (a global
// twiddle load waits about a thousand cycles under load, an LDS read about a hundred).
constexpr int kTwLdsMax = 2048;
__host__ __device__ constexpr int tw_lds_offset(int data_elems) { return (data_elems + 15) & ~15; }
// LDS elements of `data_elems` line values of length L plus the staged table.
__host__ __device__ constexpr int lds_elems(int data_elems, int L) {
    return L <= kTwLdsMax ? tw_lds_offset(data_elems) + L : data_elems;
}

// Copies the twiddle table of f into LDS behind the data (the caller synchronises before the
// first pass); null when the table stays in global memory.
__device__ __forceinline__ const float2* stage_twiddles(float2* Z, int data_elems, const FftPlan& f,
                                                        const float2* __restrict__ twbase) {
    if (f.L > kTwLdsMax) return nullptr;
    float2* t = Z + tw_lds_offset(data_elems);
    const float2* g = twbase + f.tw;
    for (int i = threadIdx.x; i < f.L; i += blockDim.x) t[i] = g[i];
    return t;
}

// Full transform of the interleaved lines. The caller has written buf (and the staged twiddles,
// twl) and synchronised.
// OOP: passes alternate between buf and alt (pass q reads buf when q is even); the result is in
// buf after an even number of passes, in alt after an odd one.
template <bool INV, int T, int E, bool TWL, int W, bool OOP = false>
__device__ void lds_fft_run(float2* __restrict__ buf, Divisor lines, const FftPlan& p, const float2* __restrict__ tw,
                            float2* __restrict__ alt = nullptr) {
    const int N = p.L;
    int Ns = 1, rem = N;
    for (int q = 0; q < p.npass; q++) {
        float2* const src = OOP && (q & 1) ? alt : buf;
        float2* const dst = OOP ? ((q & 1) ? buf : alt) : buf;
        const unsigned long long w = q < 8 ? p.codes_lo : p.codes_hi;
        const int code = (int)((w >> (8 * (q & 7))) & 0xffull);
        // one division per pass (uniform), none per butterfly
        const unsigned mag = __builtin_amdgcn_readfirstlane(Divisor::of(Ns).m);
        switch (code) {
#define LUCEM_PASS(r1, r2)                                                             \
    case radix_code(r1, r2):                                                           \
        stockham_pass<r1, r2, INV, T, E, TWL, W, OOP>(src, dst, N, lines, Ns, mag, rem, tw); \
        Ns *= r1 * r2;                                                                 \
        rem /= r1 * r2;                                                                \
        break;
            LUCEM_PASS(2, 1)
            LUCEM_PASS(3, 1)
            LUCEM_PASS(4, 1)
            LUCEM_PASS(5, 1)
            LUCEM_PASS(7, 1)
            LUCEM_PASS(8, 1)
            LUCEM_PASS(3, 2)
            LUCEM_PASS(5, 2)
            LUCEM_PASS(8, 2)
            LUCEM_PASS(3, 3)
            LUCEM_PASS(4, 3)
            LUCEM_PASS(4, 4)
#undef LUCEM_PASS
            default:
                break;
        }
    }
}

template <bool INV, int T, int E, int W = 4>
__device__ void lds_fft(float2* __restrict__ buf, Divisor lines, const FftPlan& p, const float2* __restrict__ twbase,
                        const float2* twl) {
    if (twl)
        lds_fft_run<INV, T, E, true, W>(buf, lines, p, twl);
    else
        lds_fft_run<INV, T, E, false, W>(buf, lines, p, twbase + p.tw);
}

// Out-of-place transform of the lines in `buf` with `alt` (same size) as the other buffer, the
// twiddles staged in LDS; returns the buffer that holds the result.
template <bool INV, int T, int E, int W = 4>
__device__ float2* lds_fft_oop(float2* __restrict__ buf, float2* __restrict__ alt, Divisor lines, const FftPlan& p,
                               const float2* twl) {
    lds_fft_run<INV, T, E, true, W, true>(buf, lines, p, twl, alt);
    return (p.npass & 1) ? alt : buf;
}

// Reference pad_index (core/kernels/pad.rs:11-34); -1 for Constant outside [0, len). Positions
// within one fold of the array (every position when the window is no longer than the array)
// take the closed forms; farther ones fold by the modulo of the period.
__host__ __device__ __forceinline__ int pad_index(int p, int len, int mode) {
    if (p >= 0 && p < len) return p;
    switch (mode) {
        case 1: return -1;
        case 2: return p < 0 ? 0 : len - 1;
        case 4:
            if (p >= -len && p < 2 * len) return p < 0 ? p + len : p - len;
            return ((p % len) + len) % len;
        case 0: {
            if (p >= -len && p < 2 * len) return p < 0 ? -1 - p : 2 * len - 1 - p;
            const int period = 2 * len;
            const int r = ((p % period) + period) % period;
            return r < len ? r : period - 1 - r;
        }
        default: {
            if (len > 1 && p > -len && p < 2 * len - 1) return p < 0 ? -p : 2 * len - 2 - p;
            const int period = max(2 * len - 2, 1);
            const int r = ((p % period) + period) % period;
            return r < len ? r : period - r;
        }
    }
}

// Blocks are dealt to the 8 XCDs round-robin (block b runs on XCD b % 8). Renumbering them so
// that each XCD gets a contiguous range of logical blocks keeps neighbouring column blocks, which
// share cache lines, in the same L2.
__device__ __forceinline__ int xcd_block(int b, int nblocks) {
    const int per = nblocks >> 3, rem = nblocks & 7;
    const int x = b & 7, slot = b >> 3;
    return x * per + min(x, rem) + slot;
}

}  // namespace lucem
