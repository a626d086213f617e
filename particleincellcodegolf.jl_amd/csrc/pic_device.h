// pic_device.h -- device-side building blocks shared by the gfx950 kernels.
//
// Everything in this header is compiled with -ffp-contract=off: each product
// and sum rounds as written, so the FP64 kernels reproduce the CPU
// restatement of the reference scripts bit for bit.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pic_internal.h"

#pragma clang fp contract(off)

typedef unsigned __int128 u128;
typedef unsigned long long u64;

namespace pic {

__device__ __forceinline__ uint32_t hiword(double d) {
    return (uint32_t)((u64)__double_as_longlong(d) >> 32);
}
__device__ __forceinline__ double from_hiword(uint32_t hi) {
    return __longlong_as_double((long long)((u64)hi << 32));
}

// The exp and erf below restate FreeBSD msun e_exp.c and s_erf.c, whose
// copyright notices are kept here as their licence asks:
//
// ====================================================
// Copyright (C) 1993 by Sun Microsystems, Inc. All rights reserved.
//
// Developed at SunPro, a Sun Microsystems, Inc. business.
// Permission to use, copy, modify, and distribute this
// software is freely granted, provided that this notice
// is preserved.
// ====================================================
//
// ====================================================
// Copyright (C) 2004 by Sun Microsystems, Inc. All rights reserved.
//
// Permission to use, copy, modify, and distribute this
// software is freely granted, provided that this notice
// is preserved.
// ====================================================
// ---------------------------------------------------------------------------
// exp and erf: FreeBSD msun e_exp.c / s_erf.c as shipped by OpenLibm 0.8.1,
// the erf SpecialFunctions 2.2.0 calls in the Gaussian shape
// (src/GaussianFixedPointQuiet.jl:5).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double ol_exp(double x) {
    const double ln2HI = 6.93147180369123816490e-01, ln2LO = 1.90821492927058770002e-10;
    const double invln2 = 1.44269504088896338700e+00;
    const double P1 = 1.66666666666666019037e-01, P2 = -2.77777777770155933842e-03,
                 P3 = 6.61375632143793436117e-05, P4 = -1.65339022054652515390e-06,
                 P5 = 4.13813679705723846039e-08;
    double y, hi = 0.0, lo = 0.0, c, t, twopk;
    int32_t k = 0;
    uint32_t hx = hiword(x);
    int xsb = (int)((hx >> 31) & 1);
    hx &= 0x7fffffffu;
    if (hx >= 0x40862E42u) {
        if (hx >= 0x7ff00000u) {
            if (((hx & 0xfffffu) | (uint32_t)(u64)__double_as_longlong(x)) != 0) return x + x;
            return xsb == 0 ? x : 0.0;
        }
        if (x > 7.09782712893383973096e+02) return __longlong_as_double(0x7ff0000000000000LL);
        if (x < -7.45133219101941108420e+02) return 0.0;
    }
    if (hx > 0x3fd62e42u) {
        if (hx < 0x3FF0A2B2u) {
            hi = xsb ? x + ln2HI : x - ln2HI;
            lo = xsb ? -ln2LO : ln2LO;
            k = 1 - xsb - xsb;
        } else {
            k = (int32_t)(invln2 * x + (xsb ? -0.5 : 0.5));
            t = (double)k;
            hi = x - t * ln2HI;
            lo = t * ln2LO;
        }
        x = hi - lo;
    } else if (hx < 0x3e300000u) {
        return 1.0 + x;
    } else {
        k = 0;
    }
    t = x * x;
    if (k >= -1021)
        twopk = from_hiword(0x3ff00000u + ((uint32_t)k << 20));
    else
        twopk = from_hiword(0x3ff00000u + ((uint32_t)(k + 1000) << 20));
    c = x - t * (P1 + t * (P2 + t * (P3 + t * (P4 + t * P5))));
    if (k == 0) return 1.0 - ((x * c) / (c - 2.0) - x);
    y = 1.0 - ((lo - (x * c) / (2.0 - c)) - hi);
    if (k >= -1021) {
        if (k == 1024) return y * 2.0 * 0x1p1023;
        return y * twopk;
    }
    return y * twopk * 0x1p-1000;
}

__device__ __forceinline__ double ol_erf(double x) {
    const double erx = 8.45062911510467529297e-01, efx = 1.28379167095512586316e-01,
                 efx8 = 1.02703333676410069053e+00;
    const double pp0 = 1.28379167095512558561e-01, pp1 = -3.25042107247001499370e-01,
                 pp2 = -2.84817495755985104766e-02, pp3 = -5.77027029648944159157e-03,
                 pp4 = -2.37630166566501626084e-05;
    const double qq1 = 3.97917223959155352819e-01, qq2 = 6.50222499887672944485e-02,
                 qq3 = 5.08130628187576562776e-03, qq4 = 1.32494738004321644526e-04,
                 qq5 = -3.96022827877536812320e-06;
    const double pa0 = -2.36211856075265944077e-03, pa1 = 4.14856118683748331666e-01,
                 pa2 = -3.72207876035701323847e-01, pa3 = 3.18346619901161753674e-01,
                 pa4 = -1.10894694282396677476e-01, pa5 = 3.54783043256182359371e-02,
                 pa6 = -2.16637559486879084300e-03;
    const double qa1 = 1.06420880400844228286e-01, qa2 = 5.40397917702171048937e-01,
                 qa3 = 7.18286544141962662868e-02, qa4 = 1.26171219808761642112e-01,
                 qa5 = 1.36370839120290507362e-02, qa6 = 1.19844998467991074170e-02;
    const double ra0 = -9.86494403484714822705e-03, ra1 = -6.93858572707181764372e-01,
                 ra2 = -1.05586262253232909814e+01, ra3 = -6.23753324503260060396e+01,
                 ra4 = -1.62396669462573470355e+02, ra5 = -1.84605092906711035994e+02,
                 ra6 = -8.12874355063065934246e+01, ra7 = -9.81432934416914548592e+00;
    const double sa1 = 1.96512716674392571292e+01, sa2 = 1.37657754143519042600e+02,
                 sa3 = 4.34565877475229228821e+02, sa4 = 6.45387271733267880336e+02,
                 sa5 = 4.29008140027567833386e+02, sa6 = 1.08635005541779435134e+02,
                 sa7 = 6.57024977031928170135e+00, sa8 = -6.04244152148580987438e-02;
    const double rb0 = -9.86494292470009928597e-03, rb1 = -7.99283237680523006574e-01,
                 rb2 = -1.77579549177547519889e+01, rb3 = -1.60636384855821916062e+02,
                 rb4 = -6.37566443368389627722e+02, rb5 = -1.02509513161107724954e+03,
                 rb6 = -4.83519191608651397019e+02;
    const double sb1 = 3.03380607434824582924e+01, sb2 = 3.25792512996573918826e+02,
                 sb3 = 1.53672958608443695994e+03, sb4 = 3.19985821950859553908e+03,
                 sb5 = 2.55305040643316442583e+03, sb6 = 4.74528541206955367215e+02,
                 sb7 = -2.24409524465858183362e+01;
    uint32_t hx = hiword(x);
    uint32_t ix = hx & 0x7fffffffu;
    bool neg = (hx >> 31) != 0;
    double R, S, P, Q, s, y, z, r;
    if (ix >= 0x7ff00000u) return (neg ? -1.0 : 1.0) + 1.0 / x;
    if (ix < 0x3feb0000u) {
        if (ix < 0x3e300000u) {
            if (ix < 0x00800000u) return 0.125 * (8.0 * x + efx8 * x);
            return x + efx * x;
        }
        z = x * x;
        r = pp0 + z * (pp1 + z * (pp2 + z * (pp3 + z * pp4)));
        s = 1.0 + z * (qq1 + z * (qq2 + z * (qq3 + z * (qq4 + z * qq5))));
        y = r / s;
        return x + x * y;
    }
    if (ix < 0x3ff40000u) {
        s = fabs(x) - 1.0;
        P = pa0 + s * (pa1 + s * (pa2 + s * (pa3 + s * (pa4 + s * (pa5 + s * pa6)))));
        Q = 1.0 + s * (qa1 + s * (qa2 + s * (qa3 + s * (qa4 + s * (qa5 + s * qa6)))));
        return neg ? -erx - P / Q : erx + P / Q;
    }
    if (ix >= 0x40180000u) return neg ? -1.0 : 1.0;
    x = fabs(x);
    s = 1.0 / (x * x);
    if (ix < 0x4006DB6Eu) {
        R = ra0 + s * (ra1 + s * (ra2 + s * (ra3 + s * (ra4 + s * (ra5 + s * (ra6 + s * ra7))))));
        S = 1.0 + s * (sa1 + s * (sa2 + s * (sa3 + s * (sa4 + s * (sa5 + s * (sa6 + s * (sa7 + s * sa8)))))));
    } else {
        R = rb0 + s * (rb1 + s * (rb2 + s * (rb3 + s * (rb4 + s * (rb5 + s * rb6)))));
        S = 1.0 + s * (sb1 + s * (sb2 + s * (sb3 + s * (sb4 + s * (sb5 + s * (sb6 + s * sb7))))));
    }
    z = __longlong_as_double((long long)((u64)__double_as_longlong(x) & 0xffffffff00000000ull));
    r = ol_exp(-z * z - 0.5625) * ol_exp((z - x) * (z + x) + R / S);
    return neg ? r / x - 1.0 : 1.0 - r / x;
}

// ---------------------------------------------------------------------------
// Julia scalar semantics
// ---------------------------------------------------------------------------
// mod(x, 1) on Float64 (src/NGPFourier.jl:2, src/GaussianFixedPointQuiet.jl:10):
// r = rem(x, 1); r == 0 -> +0; r < 0 -> r + 1 (rounded); else r.
__device__ __forceinline__ double jl_mod1f(double x) {
    double r = fmod(x, 1.0);
    if (r == 0.0) return 0.0;
    if (r < 0.0) return r + 1.0;
    return r;
}
// mod1(i, n) - 1 for a power-of-two n: the 0-based array index of Julia index i.
__device__ __forceinline__ int32_t cell0(int64_t i, int32_t nmask) {
    return (int32_t)((i - 1) & (int64_t)nmask);
}
// f(x) = Int(mod1(round(x*N), N)) (src/NGPFourier.jl:3), 0-based
__device__ __forceinline__ int32_t ngp_cell(double x, double Nd, int32_t nmask) {
    return cell0((int64_t)rint(x * Nd), nmask);
}
// f(g,c) = erf((g-c)*N)/2 at the left edge g = (i-0.5)/N of cell i
// (src/GaussianFixedPointQuiet.jl:5).  The left edge of cell i+1 and the right
// edge of cell i have bit-identical arguments, so 2R+2 edges give all 2R+1
// weights of the stencil exactly as the scripts' 2(2R+1) evaluations do.
__device__ __forceinline__ double edge_half_erf(int64_t i, double c, double Nd) {
    double g = ((double)i - 0.5) / Nd;
    return ol_erf((g - c) * Nd) / 2.0;
}

// ---------------------------------------------------------------------------
// Exact per-cell accumulation: contribution a >= 0 rounded (ties to even) to
// the integer grid 2^-q, summed in integers, rounded once to Float64.
// In LDS a value is held as three 43-bit limbs in separate u64 counters so
// that integer atomics never need a carry.
// ---------------------------------------------------------------------------
constexpr u64 kLimbMask = (1ull << 43) - 1;

__device__ __forceinline__ u128 fxp_from(double a, int q) {
    u64 b = (u64)__double_as_longlong(a);
    int ex = (int)((b >> 52) & 0x7ff);
    u64 m = b & 0xFFFFFFFFFFFFFull;
    int e;
    if (ex == 0x7ff) return 0;  // Inf / NaN add nothing (a defined result; no grid holds them)
    if (ex == 0) {
        if (m == 0) return 0;
        e = -1074;
    } else {
        m |= 1ull << 52;
        e = ex - 1075;
    }
    int s = e + q;
    if (s >= 0) return s <= 74 ? (u128)m << s : ~(u128)0;  // beyond 2^128 of the grid: saturate
    int r = -s;
    if (r >= 54) return 0;
    u64 qv = m >> r;
    u64 rem = m & ((1ull << r) - 1);
    u64 half = 1ull << (r - 1);
    if (rem > half || (rem == half && (qv & 1))) qv++;
    return (u128)qv;
}

__device__ __forceinline__ double fxp_to(u128 S, int q) {
    if (S == 0) return 0.0;
    u64 hi = (u64)(S >> 64), lo = (u64)S;
    int L = hi ? 128 - __clzll((long long)hi) : 64 - __clzll((long long)lo);
    u64 m;
    int e = 0;
    if (L <= 53) {
        m = lo;
    } else {
        int r = L - 53;
        m = (u64)(S >> r);
        u128 rem = S & (((u128)1 << r) - 1);
        u128 half = (u128)1 << (r - 1);
        if (rem > half || (rem == half && (m & 1))) {
            m++;
            if (m == (1ull << 53)) { m >>= 1; r++; }
        }
        e = r;
    }
    return ldexp((double)m, e - q);
}

__device__ __forceinline__ void lds_add_fxp(u64* acc, int32_t N, int32_t cell, u128 v) {
    u64 l0 = (u64)v & kLimbMask;
    u64 l1 = (u64)(v >> 43) & kLimbMask;
    u64 l2 = (u64)(v >> 86);
    if (l0) atomicAdd(&acc[cell], l0);
    if (l1) atomicAdd(&acc[N + cell], l1);
    if (l2) atomicAdd(&acc[2 * N + cell], l2);
}

__device__ __forceinline__ u128 limbs_to_u128(u64 l0, u64 l1, u64 l2) {
    return (u128)l0 + ((u128)l1 << 43) + ((u128)l2 << 86);
}

// ---------------------------------------------------------------------------
// Radix-2 decimation-in-time butterflies of M consecutive stages (s0+1 ..
// s0+M) on a group of 2^M elements spaced 2^s0 apart, held in registers:
// t = b*w (tr = br*wr - bi*wi, ti = br*wi + bi*wr), b = a - t, a = a + t,
// exactly as the stage-by-stage transform of the CPU restatement.  tb is the
// per-pass twiddle table at this pass (entry ((2^t - 1 + jlo) << s0) + lo);
// the inverse transform uses the conjugate twiddles.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t brev(uint32_t i, int bits) {
    return __brev(i) >> (32 - bits);
}
// LDS padding (one slot per 16 doubles) against bank conflicts
__device__ __forceinline__ int padi(int i) { return i + (i >> 4); }

template <int M>
__device__ __forceinline__ void radix2_group(double* ar, double* ai, const double2* __restrict__ tb, int lo,
                                             int s0, bool inverse) {
    constexpr int G = 1 << M;
#pragma unroll
    for (int t = 0; t < M; ++t) {
#pragma unroll
        for (int j1 = 0; j1 < G; ++j1) {
            if (j1 & (1 << t)) continue;
            const int j2 = j1 | (1 << t);
            const double2 w = tb[(((1 << t) - 1 + (j1 & ((1 << t) - 1))) << s0) + lo];
            const double wr = w.x;
            const double wi = inverse ? -w.y : w.y;
            const double br = ar[j2], bi = ai[j2];
            const double tr = br * wr - bi * wi;
            const double tq = br * wi + bi * wr;
            const double ur = ar[j1], ui = ai[j1];
            ar[j1] = ur + tr;
            ai[j1] = ui + tq;
            ar[j2] = ur - tr;
            ai[j2] = ui - tq;
        }
    }
}

// ---------------------------------------------------------------------------
// Order-independent sums of the mixed paths: each value rounded once onto a
// fixed-point grid of 2^-q and summed exactly in 128-bit integers, so a sum
// does not depend on how particles are split over tiles, threads or sorts.
// ---------------------------------------------------------------------------
typedef __int128 i128;

// rint(a * 2^q) (ties to even) as a 128-bit integer; exact below 2^126 in
// magnitude, which saturates.  Non-finite values, and a * 2^q beyond the
// double range, give 0 (a non-finite field is reported by the solve's check).
// Below 2^63 the integer t converts directly; beyond, it is split into
// hi * 2^64 + lo by a floor.  lo = t - hi * 2^64 is exact there for either
// sign: t's lowest bit is at least 2^11, so lo < 2^64 needs at most 53 bits
// (below 2^63 a negative t's lo could need 54).
__host__ __device__ __forceinline__ i128 fix_round128(double a, int q) {
    const double t = rint(ldexp(a, q));
    if (fabs(t) < 0x1p63) return (i128)(long long)t;
    if (t != t || __builtin_isinf(t)) return 0;
    if (!(fabs(t) < 0x1p126)) return t > 0 ? ((i128)1 << 126) : -((i128)1 << 126);
    const double hi = floor(t * 0x1p-64);  // exact: t is an integer beyond 2^53
    const double lo = t - hi * 0x1p64;     // exact, in [0, 2^64)
    return (i128)(long long)hi * ((i128)1 << 64) + (i128)(u64)lo;
}

__device__ __forceinline__ double i128_to_double(i128 v, int q) {
    return v < 0 ? -fxp_to((u128)(-v), q) : fxp_to((u128)v, q);
}

// Sum over a wave (lane 0 receives it).
__device__ __forceinline__ i128 wave_sum_i128(i128 v) {
#pragma unroll
    for (int h = 32; h > 0; h >>= 1) {
        const u64 lo = __shfl_down((u64)v, h, 64);
        const u64 hi = __shfl_down((u64)((u128)v >> 64), h, 64);
        v += (i128)(((u128)hi << 64) | lo);
    }
    return v;
}

// ---------------------------------------------------------------------------
// Canonical reduction tree (defined in the CPU restatement): within a chunk of
// at most 4096 values, lane t < 256 sums elements t, t+256, ... in order from
// +0.0, then a halving tree a[t] += a[t+h] for h = 128..1.  The h = 128, 64
// steps go through LDS, h = 32..1 through wave shuffles (lane t adds lane
// t+h: the same operands as the LDS form).  Requires blockDim.x >= 256; all
// threads of the block must call.  `red` is LDS of >= 256 doubles.
// ---------------------------------------------------------------------------
// NaN-propagating maximum: a if a is NaN, else b if b is NaN, else
// (a > b ? a : b).  Written as three selects so that the compiler emits
// compares and v_cndmask instead of a branch per call (the norm and tree
// loops call it thousands of times per launch).
__device__ __forceinline__ double nanmax(double a, double b) {
    const double m = a > b ? a : b;
    const double r = b != b ? b : m;
    return a != a ? a : r;
}

template <bool MAX>
__device__ __forceinline__ double tree256(double* red, double v) {
    const int t = threadIdx.x;
    if (t < kTreeLanes) red[t] = v;
    __syncthreads();
    if (t < 128) red[t] = MAX ? nanmax(red[t], red[t + 128]) : red[t] + red[t + 128];
    __syncthreads();
    if (t < 64) {
        double r = MAX ? nanmax(red[t], red[t + 64]) : red[t] + red[t + 64];
#pragma unroll
        for (int h = 32; h > 0; h >>= 1) {
            const double o = __shfl_down(r, h, 64);
            r = MAX ? nanmax(r, o) : r + o;
        }
        if (t == 0) red[kTreeLanes] = r;
    }
    __syncthreads();
    const double r = red[kTreeLanes];
    __syncthreads();
    return r;
}

template <typename F>
__device__ __forceinline__ double block_chunk_sum(double* red, int64_t n, F&& elem) {
    const int t = threadIdx.x;
    double s = 0.0;
    if (t < kTreeLanes) {
        if (n > 0 && n <= kTreeChunk) {
            // a lane's (at most 16) values are all loaded before its sum, so
            // their latencies overlap: past n the element of index n-1 is
            // read (no branch around the load) and -0.0 taken in its place;
            // x + (-0.0) is x for every x, signed zeros and NaN included
            constexpr int K = kTreeChunk / kTreeLanes;
            double x[K];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const int64_t i = t + (int64_t)k * kTreeLanes;
                x[k] = elem(i < n ? i : n - 1);
            }
#pragma unroll
            for (int k = 0; k < K; ++k) s = s + (t + (int64_t)k * kTreeLanes < n ? x[k] : -0.0);
        } else {
            for (int64_t i = t; i < n; i += kTreeLanes) s = s + elem(i);
        }
    }
    return tree256<false>(red, s);
}

// NaN-propagating max of |x| over n values (order-independent).
// Canonical tree over n partials vals[stride*i] (consumed in place).
__device__ inline double tree_strided(double* red, double* vals, int stride, int64_t n) {
    while (n > kTreeChunk) {
        const int64_t C = (n + kTreeChunk - 1) / kTreeChunk;
        for (int64_t c = 0; c < C; ++c) {
            const int64_t m = n - c * kTreeChunk < kTreeChunk ? n - c * kTreeChunk : kTreeChunk;
            const double s = block_chunk_sum(red, m, [&](int64_t i) { return vals[stride * (c * kTreeChunk + i)]; });
            if (threadIdx.x == 0) vals[stride * c] = s;
            __syncthreads();
        }
        n = C;
    }
    return block_chunk_sum(red, n, [&](int64_t i) { return vals[stride * i]; });
}

template <typename F>
__device__ __forceinline__ double block_maxabs(double* red, int64_t n, F&& elem) {
    const int t = threadIdx.x;
    double m = 0.0;
    if (t < kTreeLanes) {
        if (n > 0 && n <= kTreeChunk) {
            // loads first, as in block_chunk_sum; +0.0 past n leaves the
            // maximum (>= +0 or NaN) as it is
            constexpr int K = kTreeChunk / kTreeLanes;
            double x[K];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const int64_t i = t + (int64_t)k * kTreeLanes;
                x[k] = elem(i < n ? i : n - 1);
            }
#pragma unroll
            for (int k = 0; k < K; ++k) m = nanmax(m, t + (int64_t)k * kTreeLanes < n ? fabs(x[k]) : 0.0);
        } else {
            for (int64_t i = t; i < n; i += kTreeLanes) m = nanmax(m, fabs(elem(i)));
        }
    }
    return tree256<true>(red, m);
}

// LinearAlgebra.norm (generic_norm2) with the canonical sum, n <= 4096.
template <typename F>
__device__ __forceinline__ double block_norm2(double* red, int64_t n, F&& elem) {
    double m = block_maxabs(red, n, elem);
    if (m == 0.0 || isinf(m)) return m;
    if (isfinite((double)n * m * m) && m * m != 0.0) {
        double s = block_chunk_sum(red, n, [&](int64_t i) { double x = elem(i); return x * x; });
        return sqrt(s);
    }
    double s = block_chunk_sum(red, n, [&](int64_t i) { double q = fabs(elem(i)) / m; return q * q; });
    return m * sqrt(s);
}

// Three canonical trees at once (sum or NaN-propagating max); red holds
// >= 3*256 + 3 doubles.  Same operands per tree as tree256.
template <bool MAX>
__device__ __forceinline__ void tree256x3(double* red, double& a, double& b, double& c) {
    const int t = threadIdx.x;
    auto op = [](double x, double y) { return MAX ? nanmax(x, y) : x + y; };
    if (t < kTreeLanes) {
        red[t] = a;
        red[kTreeLanes + t] = b;
        red[2 * kTreeLanes + t] = c;
    }
    __syncthreads();
    if (t < 128) {
#pragma unroll
        for (int q = 0; q < 3; ++q) red[q * kTreeLanes + t] = op(red[q * kTreeLanes + t], red[q * kTreeLanes + t + 128]);
    }
    __syncthreads();
    if (t < 64) {
        double r[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) r[q] = op(red[q * kTreeLanes + t], red[q * kTreeLanes + t + 64]);
#pragma unroll
        for (int h = 32; h > 0; h >>= 1)
#pragma unroll
            for (int q = 0; q < 3; ++q) r[q] = op(r[q], __shfl_down(r[q], h, 64));
        if (t == 0) {
            red[3 * kTreeLanes] = r[0];
            red[3 * kTreeLanes + 1] = r[1];
            red[3 * kTreeLanes + 2] = r[2];
        }
    }
    __syncthreads();
    a = red[3 * kTreeLanes];
    b = red[3 * kTreeLanes + 1];
    c = red[3 * kTreeLanes + 2];
    __syncthreads();
}

// Canonical chunk sums of v^2 and v/P over one 4096-particle chunk, one pass.
template <typename FV>
__device__ __forceinline__ void chunk_sums_v(double* red, int64_t n, double Pd, FV&& vel, double& s2,
                                             double& s1) {
    double a = 0.0, b = 0.0, z = 0.0;
    if (threadIdx.x < kTreeLanes) {
        double v[kTreeChunk / kTreeLanes];
#pragma unroll
        for (int i = 0; i < kTreeChunk / kTreeLanes; ++i) {
            const int64_t e = threadIdx.x + (int64_t)i * kTreeLanes;
            v[i] = e < n ? vel(e) : 0.0;
        }
#pragma unroll
        for (int i = 0; i < kTreeChunk / kTreeLanes; ++i) {
            const int64_t e = threadIdx.x + (int64_t)i * kTreeLanes;
            if (e < n) {
                a = a + v[i] * v[i];
                b = b + v[i] / Pd;
            }
        }
    }
    tree256x3<false>(red, a, b, z);
    s2 = a;
    s1 = b;
}

}  // namespace pic
