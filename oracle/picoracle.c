/*
 * picoracle.c -- CPU FP64 restatement of the electrostatic PIC time step of
 * jwscook/ParticleInCellCodeGolf.jl (Julia scripts under src/).
 *
 * TEST INFRASTRUCTURE ONLY (see picoracle.h).  Built with -ffp-contract=off so
 * every product and sum rounds exactly as written.
 *
 * How each function is pinned (checks live in tests/test_oracle_*.py):
 *   orc_jl_mod1f / mod1i / round  hand cases of Julia semantics
 *                                 (mod(-1e-20,1)==1.0, mod1(0,N)==N, round(2.5)==2)
 *   orc_exp, orc_erf              restatement of OpenLibm 0.8.1 (FreeBSD msun
 *                                 e_exp.c / s_erf.c); checked against glibc exp and
 *                                 erf (math.exp / math.erf) to <= 1 ulp over every
 *                                 branch, and against scipy.special.erf; odd symmetry,
 *                                 erf(0)=0, erf(>=6)=1.  Bit equality with OpenLibm
 *                                 itself is unpinned (OpenLibm is not in this image).
 *   orc_fft / orc_fieldsolve      O(N^2) brute-force DFT, numpy.fft, single-mode closed
 *                                 form E = sin(2 pi m j/N)/(2 pi m), k=0 and Nyquist
 *                                 zeroed.  FFTW 3.3.10's rounding is unpinned.
 *   orc_gauss_stencil             partition of unity, mirror symmetry, translation
 *                                 invariance, 16-vs-30-erf bit equality.
 *   orc_deposit                   brute-force sums on small P, exact vs sequential
 *                                 order bound, NGP = count*w exactly.
 *   orc_push_midpoint / kick      brute force on tiny P, free streaming with E = 0.
 *   orc_diagnostics               closed forms (initial D2 == 1).
 *   orc_run                       momentum conservation, plasma frequency sqrt(W),
 *                                 two-stream growth rate against the analytic gamma.
 */
#include "picoracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef unsigned __int128 u128;

static inline uint64_t d2u(double d) { uint64_t u; memcpy(&u, &d, 8); return u; }
static inline double u2d(uint64_t u) { double d; memcpy(&d, &u, 8); return d; }
static inline uint32_t hiword(double d) { return (uint32_t)(d2u(d) >> 32); }
static inline double with_hiword(uint32_t hi) { return u2d((uint64_t)hi << 32); }

/* ------------------------------------------------------------------------ */
/* Julia scalar semantics                                                    */
/* ------------------------------------------------------------------------ */

/* Julia 1.8 `mod(x::Float64, y::Float64)` with y = 1: r = rem(x, y); r == 0 ->
 * copysign(0, y); sign(r) != sign(y) -> r + y (rounded, so mod(-1e-20, 1) ==
 * 1.0); else r.  Used by u() at src/NGPFourier.jl:2 and the wrap at
 * src/GaussianFixedPointQuiet.jl:10. */
double orc_jl_mod1f(double x) {
    double r = fmod(x, 1.0);
    if (r == 0.0) return 0.0;
    if (r < 0.0) return r + 1.0;
    return r;
}

/* Julia `mod1(i, n)` for integers, result in 1..n (src/NGPFourier.jl:3,
 * src/GaussianFixedPointQuiet.jl:6). */
int64_t orc_jl_mod1i(int64_t i, int64_t n) {
    int64_t m = i % n;
    if (m < 0) m += n;
    return m == 0 ? n : m;
}

/* Julia `round(x)` = RoundNearest = ties to even; rint under the default
 * rounding mode is exactly that. */
double orc_jl_round(double x) { return rint(x); }

/* ------------------------------------------------------------------------ */
/* The exp and erf below restate FreeBSD msun e_exp.c and s_erf.c, whose      */
/* copyright notices are kept here as their licence asks:                    */
/*                                                                           */
/* ====================================================                      */
/* Copyright (C) 1993 by Sun Microsystems, Inc. All rights reserved.         */
/*                                                                           */
/* Developed at SunPro, a Sun Microsystems, Inc. business.                   */
/* Permission to use, copy, modify, and distribute this                      */
/* software is freely granted, provided that this notice                     */
/* is preserved.                                                             */
/* ====================================================                      */
/*                                                                           */
/* ====================================================                      */
/* Copyright (C) 2004 by Sun Microsystems, Inc. All rights reserved.         */
/*                                                                           */
/* Permission to use, copy, modify, and distribute this                      */
/* software is freely granted, provided that this notice                     */
/* is preserved.                                                             */
/* ====================================================                      */
/* ------------------------------------------------------------------------ */
/* exp: restatement of FreeBSD msun e_exp.c (the __ieee754_exp that OpenLibm  */
/* 0.8.1 ships and its erf calls).  Argument reduction x = k ln2 + r with     */
/* ln2 split hi/lo, a degree-5 Remez polynomial in r^2, reconstruction by     */
/* exponent insertion.                                                       */
/* ------------------------------------------------------------------------ */
static const double EXP_ln2HI = 6.93147180369123816490e-01;  /* 0x3fe62e42 fee00000 */
static const double EXP_ln2LO = 1.90821492927058770002e-10;  /* 0x3dea39ef 35793c76 */
static const double EXP_invln2 = 1.44269504088896338700e+00; /* 0x3ff71547 652b82fe */
static const double EXP_P1 = 1.66666666666666019037e-01;
static const double EXP_P2 = -2.77777777770155933842e-03;
static const double EXP_P3 = 6.61375632143793436117e-05;
static const double EXP_P4 = -1.65339022054652515390e-06;
static const double EXP_P5 = 4.13813679705723846039e-08;
static const double EXP_othr = 7.09782712893383973096e+02;
static const double EXP_uthr = -7.45133219101941108420e+02;

double orc_exp(double x) {
    double y, hi = 0.0, lo = 0.0, c, t, twopk;
    int32_t k = 0;
    uint32_t hx = hiword(x);
    int xsb = (int)((hx >> 31) & 1);
    hx &= 0x7fffffffu;
    if (hx >= 0x40862E42u) {
        if (hx >= 0x7ff00000u) {
            if (((hx & 0xfffffu) | (uint32_t)d2u(x)) != 0) return x + x;
            return xsb == 0 ? x : 0.0;
        }
        if (x > EXP_othr) return INFINITY;
        if (x < EXP_uthr) return 0.0;
    }
    if (hx > 0x3fd62e42u) {
        if (hx < 0x3FF0A2B2u) {
            hi = xsb ? x + EXP_ln2HI : x - EXP_ln2HI;
            lo = xsb ? -EXP_ln2LO : EXP_ln2LO;
            k = 1 - xsb - xsb;
        } else {
            k = (int32_t)(EXP_invln2 * x + (xsb ? -0.5 : 0.5));
            t = (double)k;
            hi = x - t * EXP_ln2HI;
            lo = t * EXP_ln2LO;
        }
        x = hi - lo;
    } else if (hx < 0x3e300000u) {
        return 1.0 + x;
    } else {
        k = 0;
    }
    t = x * x;
    if (k >= -1021)
        twopk = with_hiword(0x3ff00000u + ((uint32_t)k << 20));
    else
        twopk = with_hiword(0x3ff00000u + ((uint32_t)(k + 1000) << 20));
    c = x - t * (EXP_P1 + t * (EXP_P2 + t * (EXP_P3 + t * (EXP_P4 + t * EXP_P5))));
    if (k == 0) return 1.0 - ((x * c) / (c - 2.0) - x);
    y = 1.0 - ((lo - (x * c) / (2.0 - c)) - hi);
    if (k >= -1021) {
        if (k == 1024) return y * 2.0 * 0x1p1023;
        return y * twopk;
    }
    return y * twopk * 0x1p-1000;
}

/* ------------------------------------------------------------------------ */
/* erf: restatement of FreeBSD msun s_erf.c (OpenLibm 0.8.1, the erf that     */
/* SpecialFunctions 2.2.0 calls for the Gaussian shape at                    */
/* src/GaussianFixedPointQuiet.jl:5).  Four ranges: rational P/Q in x^2 on    */
/* [0,0.84375); erx + P/Q in (|x|-1) on [0.84375,1.25); 1 - exp(-z^2-0.5625)  */
/* exp((z-x)(z+x)+R/S)/x with R/S rational in 1/x^2 on [1.25,6); +-1 beyond. */
/* ------------------------------------------------------------------------ */
static const double ERF_erx = 8.45062911510467529297e-01;
static const double ERF_efx = 1.28379167095512586316e-01;
static const double ERF_efx8 = 1.02703333676410069053e+00;
static const double ERF_pp0 = 1.28379167095512558561e-01, ERF_pp1 = -3.25042107247001499370e-01,
                    ERF_pp2 = -2.84817495755985104766e-02, ERF_pp3 = -5.77027029648944159157e-03,
                    ERF_pp4 = -2.37630166566501626084e-05;
static const double ERF_qq1 = 3.97917223959155352819e-01, ERF_qq2 = 6.50222499887672944485e-02,
                    ERF_qq3 = 5.08130628187576562776e-03, ERF_qq4 = 1.32494738004321644526e-04,
                    ERF_qq5 = -3.96022827877536812320e-06;
static const double ERF_pa0 = -2.36211856075265944077e-03, ERF_pa1 = 4.14856118683748331666e-01,
                    ERF_pa2 = -3.72207876035701323847e-01, ERF_pa3 = 3.18346619901161753674e-01,
                    ERF_pa4 = -1.10894694282396677476e-01, ERF_pa5 = 3.54783043256182359371e-02,
                    ERF_pa6 = -2.16637559486879084300e-03;
static const double ERF_qa1 = 1.06420880400844228286e-01, ERF_qa2 = 5.40397917702171048937e-01,
                    ERF_qa3 = 7.18286544141962662868e-02, ERF_qa4 = 1.26171219808761642112e-01,
                    ERF_qa5 = 1.36370839120290507362e-02, ERF_qa6 = 1.19844998467991074170e-02;
static const double ERF_ra0 = -9.86494403484714822705e-03, ERF_ra1 = -6.93858572707181764372e-01,
                    ERF_ra2 = -1.05586262253232909814e+01, ERF_ra3 = -6.23753324503260060396e+01,
                    ERF_ra4 = -1.62396669462573470355e+02, ERF_ra5 = -1.84605092906711035994e+02,
                    ERF_ra6 = -8.12874355063065934246e+01, ERF_ra7 = -9.81432934416914548592e+00;
static const double ERF_sa1 = 1.96512716674392571292e+01, ERF_sa2 = 1.37657754143519042600e+02,
                    ERF_sa3 = 4.34565877475229228821e+02, ERF_sa4 = 6.45387271733267880336e+02,
                    ERF_sa5 = 4.29008140027567833386e+02, ERF_sa6 = 1.08635005541779435134e+02,
                    ERF_sa7 = 6.57024977031928170135e+00, ERF_sa8 = -6.04244152148580987438e-02;
static const double ERF_rb0 = -9.86494292470009928597e-03, ERF_rb1 = -7.99283237680523006574e-01,
                    ERF_rb2 = -1.77579549177547519889e+01, ERF_rb3 = -1.60636384855821916062e+02,
                    ERF_rb4 = -6.37566443368389627722e+02, ERF_rb5 = -1.02509513161107724954e+03,
                    ERF_rb6 = -4.83519191608651397019e+02;
static const double ERF_sb1 = 3.03380607434824582924e+01, ERF_sb2 = 3.25792512996573918826e+02,
                    ERF_sb3 = 1.53672958608443695994e+03, ERF_sb4 = 3.19985821950859553908e+03,
                    ERF_sb5 = 2.55305040643316442583e+03, ERF_sb6 = 4.74528541206955367215e+02,
                    ERF_sb7 = -2.24409524465858183362e+01;

double orc_erf(double x) {
    uint32_t hx = hiword(x);
    uint32_t ix = hx & 0x7fffffffu;
    int neg = (hx >> 31) != 0;
    double R, S, P, Q, s, y, z, r;
    if (ix >= 0x7ff00000u) { /* erf(nan) = nan, erf(+-inf) = +-1 */
        return (neg ? -1.0 : 1.0) + 1.0 / x;
    }
    if (ix < 0x3feb0000u) { /* |x| < 0.84375 */
        if (ix < 0x3e300000u) { /* |x| < 2^-28 */
            if (ix < 0x00800000u) return 0.125 * (8.0 * x + ERF_efx8 * x);
            return x + ERF_efx * x;
        }
        z = x * x;
        r = ERF_pp0 + z * (ERF_pp1 + z * (ERF_pp2 + z * (ERF_pp3 + z * ERF_pp4)));
        s = 1.0 + z * (ERF_qq1 + z * (ERF_qq2 + z * (ERF_qq3 + z * (ERF_qq4 + z * ERF_qq5))));
        y = r / s;
        return x + x * y;
    }
    if (ix < 0x3ff40000u) { /* 0.84375 <= |x| < 1.25 */
        s = fabs(x) - 1.0;
        P = ERF_pa0 + s * (ERF_pa1 + s * (ERF_pa2 + s * (ERF_pa3 + s * (ERF_pa4 + s * (ERF_pa5 + s * ERF_pa6)))));
        Q = 1.0 + s * (ERF_qa1 + s * (ERF_qa2 + s * (ERF_qa3 + s * (ERF_qa4 + s * (ERF_qa5 + s * ERF_qa6)))));
        return neg ? -ERF_erx - P / Q : ERF_erx + P / Q;
    }
    if (ix >= 0x40180000u) { /* |x| >= 6: one - tiny rounds to one */
        return neg ? -1.0 : 1.0;
    }
    x = fabs(x);
    s = 1.0 / (x * x);
    if (ix < 0x4006DB6Eu) { /* |x| < 1/0.35 */
        R = ERF_ra0 + s * (ERF_ra1 + s * (ERF_ra2 + s * (ERF_ra3 + s * (ERF_ra4 + s * (ERF_ra5 + s * (ERF_ra6 + s * ERF_ra7))))));
        S = 1.0 + s * (ERF_sa1 + s * (ERF_sa2 + s * (ERF_sa3 + s * (ERF_sa4 + s * (ERF_sa5 + s * (ERF_sa6 + s * (ERF_sa7 + s * ERF_sa8)))))));
    } else {
        R = ERF_rb0 + s * (ERF_rb1 + s * (ERF_rb2 + s * (ERF_rb3 + s * (ERF_rb4 + s * (ERF_rb5 + s * ERF_rb6)))));
        S = 1.0 + s * (ERF_sb1 + s * (ERF_sb2 + s * (ERF_sb3 + s * (ERF_sb4 + s * (ERF_sb5 + s * (ERF_sb6 + s * ERF_sb7))))));
    }
    z = u2d(d2u(x) & 0xffffffff00000000ull); /* SET_LOW_WORD(z, 0) */
    r = orc_exp(-z * z - 0.5625) * orc_exp((z - x) * (z + x) + R / S);
    return neg ? r / x - 1.0 : 1.0 - r / x;
}

/* ------------------------------------------------------------------------ */
/* Canonical reduction tree.  The scripts' `sum` is Julia's pairwise SIMD sum */
/* whose order is unpinnable; the oracle fixes one tree that a GPU reproduces */
/* exactly: chunks of 4096 elements; inside a chunk 256 lanes each sum the    */
/* elements lane, lane+256, ... in order from +0.0, then a halving tree over  */
/* the 256 lane sums (a[t] += a[t+h], h = 128..1); chunk sums are reduced by  */
/* the same rule recursively.                                                */
/* ------------------------------------------------------------------------ */
#define TREE_CHUNK 4096
#define TREE_LANES 256

static double chunk_sum(const double* x, int64_t n) {
    double a[TREE_LANES];
    for (int t = 0; t < TREE_LANES; ++t) {
        double s = 0.0;
        for (int64_t i = t; i < n; i += TREE_LANES) s = s + x[i];
        a[t] = s;
    }
    for (int h = TREE_LANES / 2; h > 0; h >>= 1)
        for (int t = 0; t < h; ++t) a[t] = a[t] + a[t + h];
    return a[0];
}

double orc_tree_sum(const double* x, int64_t n) {
    if (n <= TREE_CHUNK) return chunk_sum(x, n);
    int64_t C = (n + TREE_CHUNK - 1) / TREE_CHUNK;
    double* p = (double*)malloc(sizeof(double) * (size_t)C);
    for (int64_t c = 0; c < C; ++c) {
        int64_t m = n - c * TREE_CHUNK;
        p[c] = chunk_sum(x + c * TREE_CHUNK, m < TREE_CHUNK ? m : TREE_CHUNK);
    }
    double r = orc_tree_sum(p, C);
    free(p);
    return r;
}

/* NaN-propagating max of |x| (Julia maximum(abs, x)). */
static double maxabs(const double* x, int64_t n) {
    double m = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        double a = fabs(x[i]);
        if (a != a) return a;
        if (a > m) m = a;
    }
    return m;
}

/* LinearAlgebra.norm(x) (generic_norm2, julia/stdlib LinearAlgebra generic.jl):
 * maxabs == 0 or Inf -> maxabs; if n*maxabs^2 finite and maxabs^2 != 0 ->
 * sqrt(sum x^2); else maxabs*sqrt(sum (|x|/maxabs)^2).  The sum uses the
 * canonical tree above. */
double orc_norm2(const double* x, int64_t n) {
    double m = maxabs(x, n);
    if (m == 0.0 || isinf(m)) return m;
    double* sq = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    double r;
    if (isfinite((double)n * m * m) && m * m != 0.0) {
        for (int64_t i = 0; i < n; ++i) sq[i] = x[i] * x[i];
        r = sqrt(orc_tree_sum(sq, n));
    } else {
        for (int64_t i = 0; i < n; ++i) { double q = fabs(x[i]) / m; sq[i] = q * q; }
        r = m * sqrt(orc_tree_sum(sq, n));
    }
    free(sq);
    return r;
}

/* isapprox(F, E; rtol, atol) on arrays (src/GaussianFixedPointQuiet.jl:8):
 * d = norm(F - E); finite d -> d <= max(atol, rtol*max(norm(F), norm(E)));
 * otherwise elementwise scalar isapprox (so F filled with NaN is never close). */
int orc_isapprox(const double* F, const double* E, int64_t n, double rtol, double atol) {
    double* d = (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    for (int64_t i = 0; i < n; ++i) d[i] = F[i] - E[i];
    double dn = orc_norm2(d, n);
    free(d);
    if (isfinite(dn)) {
        double nf = orc_norm2(F, n), ne = orc_norm2(E, n);
        double mx = nf > ne ? nf : ne;
        if (nf != nf || ne != ne) mx = NAN;
        double tol = rtol * mx;
        double lim = atol > tol ? atol : tol;
        if (tol != tol) lim = tol;
        return dn <= lim;
    }
    for (int64_t i = 0; i < n; ++i) {
        double a = F[i], b = E[i];
        if (a == b) continue;
        if (!(isfinite(a) && isfinite(b))) return 0;
        double ma = fabs(a) > fabs(b) ? fabs(a) : fabs(b);
        double t = rtol * ma;
        if (!(fabs(a - b) <= (atol > t ? atol : t))) return 0;
    }
    return 1;
}

/* ------------------------------------------------------------------------ */
/* Exact per-cell accumulation.  Each contribution a >= 0 (a Float64 product  */
/* wt*w, as in `r[k[1]]+=k[2]*w`, src/GaussianFixedPointQuiet.jl:7) is        */
/* rounded to the integer grid 2^-q (ties to even), summed exactly in 128-bit */
/* integers and rounded once to Float64 (ties to even).  q puts P*w below     */
/* 2^125, so the sum can never wrap.  The result is independent of summation  */
/* order; the scripts' sequential order differs from it by at most the usual  */
/* (n-1)*eps*sum|a| bound.                                                    */
/* ------------------------------------------------------------------------ */
int orc_fxp_frac_bits(int64_t P, double w) {
    double total = (double)P * fabs(w);
    if (!(total > 0.0) || !isfinite(total)) return 64;
    int e = ilogb(total);
    int q = 124 - e;
    if (q > 1100) q = 1100;
    return q;
}

static u128 fxp_from(double a, int q) {
    uint64_t b = d2u(a);
    int ex = (int)((b >> 52) & 0x7ff);
    uint64_t m = b & 0xFFFFFFFFFFFFFull;
    int e;
    if (ex == 0x7ff) return 0;  /* Inf / NaN add nothing (a defined result; no grid holds them) */
    if (ex == 0) {
        if (m == 0) return 0;
        e = -1074;
    } else {
        m |= 1ull << 52;
        e = ex - 1075;
    }
    int s = e + q;
    if (s >= 0) return s <= 74 ? (u128)m << s : ~(u128)0;  /* beyond 2^128 of the grid: saturate */
    int r = -s;
    if (r >= 54) return 0;
    uint64_t qv = m >> r;
    uint64_t rem = m & ((1ull << r) - 1);
    uint64_t half = 1ull << (r - 1);
    if (rem > half || (rem == half && (qv & 1))) qv++;
    return qv;
}

static double fxp_to(u128 S, int q) {
    if (S == 0) return 0.0;
    uint64_t hi = (uint64_t)(S >> 64), lo = (uint64_t)S;
    int L = hi ? 128 - __builtin_clzll(hi) : 64 - __builtin_clzll(lo);
    uint64_t m;
    int e = 0;
    if (L <= 53) {
        m = lo;
    } else {
        int r = L - 53;
        m = (uint64_t)(S >> r);
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

void orc_fxp_from_double(double a, int q, uint64_t* lo, uint64_t* hi) {
    u128 v = fxp_from(a, q);
    *lo = (uint64_t)v;
    *hi = (uint64_t)(v >> 64);
}

double orc_fxp_to_double(uint64_t lo, uint64_t hi, int q) {
    return fxp_to(((u128)hi << 64) | lo, q);
}

/* ------------------------------------------------------------------------ */
/* FFT.  The scripts call FFTW's complex fft on the real charge vector (a     */
/* real array promoted to complex) and ifft with 1/N scaling                  */
/* (src/GaussianFixedPointQuiet.jl:9).  FFTW's codelet order is unpinnable;   */
/* the oracle fixes an iterative radix-2 decimation-in-time transform:        */
/* bit-reversed input order, stages len = 2, 4, ..., N, butterfly            */
/*   t = a[j+h]*w (tr = ar*wr - ai*wi, ti = ar*wi + ai*wr),                   */
/*   a[j+h] = a[j] - t, a[j] = a[j] + t,                                      */
/* twiddle w = W[j*(N/len)] with W[m] = exp(-2 pi i m/N) (conjugated for the  */
/* inverse).  The table is built from cos/sin on [0, pi/4] and exact          */
/* quarter-wave symmetries.                                                   */
/* ------------------------------------------------------------------------ */
void orc_twiddles(int32_t N, double* re, double* im) {
    const double twopi = 6.283185307179586; /* Julia 2pi = 2*Float64(pi) */
    int32_t h = N / 2, q = N / 4, e = N / 8;
    for (int32_t m = 0; m < h; ++m) {
        double c, s;
        if (N < 8) {
            if (m == 0) { c = 1.0; s = 0.0; }
            else { c = 0.0; s = 1.0; } /* N = 4, m = 1 */
        } else if (m <= e) {
            double th = twopi * (double)m / (double)N;
            c = cos(th); s = sin(th);
        } else if (m <= q) {
            double th = twopi * (double)(q - m) / (double)N;
            c = sin(th); s = cos(th);
        } else {
            int32_t mm = h - m;
            if (mm <= e) {
                double th = twopi * (double)mm / (double)N;
                c = -cos(th); s = sin(th);
            } else {
                double th = twopi * (double)(q - mm) / (double)N;
                c = -sin(th); s = cos(th);
            }
        }
        re[m] = c;
        im[m] = -s;
    }
}

static int32_t ilog2i(int32_t N) { int32_t l = 0; while ((1 << l) < N) ++l; return l; }

static uint32_t bitrev(uint32_t i, int32_t bits) {
    uint32_t r = 0;
    for (int32_t b = 0; b < bits; ++b) { r = (r << 1) | (i & 1); i >>= 1; }
    return r;
}

static void fft_with(int32_t N, double* re, double* im, int32_t inverse,
                     const double* twr, const double* twi) {
    int32_t bits = ilog2i(N);
    for (int32_t i = 0; i < N; ++i) {
        int32_t j = (int32_t)bitrev((uint32_t)i, bits);
        if (j > i) {
            double t = re[i]; re[i] = re[j]; re[j] = t;
            t = im[i]; im[i] = im[j]; im[j] = t;
        }
    }
    for (int32_t len = 2; len <= N; len <<= 1) {
        int32_t h = len >> 1, step = N / len;
        for (int32_t base = 0; base < N; base += len) {
            for (int32_t j = 0; j < h; ++j) {
                double wr = twr[j * step];
                double wi = inverse ? -twi[j * step] : twi[j * step];
                double ar = re[base + j + h], ai = im[base + j + h];
                double tr = ar * wr - ai * wi;
                double ti = ar * wi + ai * wr;
                double ur = re[base + j], ui = im[base + j];
                re[base + j] = ur + tr;
                im[base + j] = ui + ti;
                re[base + j + h] = ur - tr;
                im[base + j + h] = ui - ti;
            }
        }
    }
}

void orc_fft(int32_t N, double* re, double* im, int32_t inverse) {
    double* twr = (double*)malloc(sizeof(double) * (size_t)(N / 2 + 1));
    double* twi = (double*)malloc(sizeof(double) * (size_t)(N / 2 + 1));
    orc_twiddles(N, twr, twi);
    fft_with(N, re, im, inverse, twr, twi);
    free(twr);
    free(twi);
}

/* Textbook O(N^2) DFT in long double, unnormalised in both directions. */
void orc_dft_bruteforce(int32_t N, const double* re, const double* im,
                        double* out_re, double* out_im, int32_t inverse) {
    const long double tp = 6.283185307179586476925286766559L;
    for (int32_t k = 0; k < N; ++k) {
        long double sr = 0, si = 0;
        for (int32_t j = 0; j < N; ++j) {
            long double th = (inverse ? tp : -tp) * (long double)((int64_t)k * j % N) / N;
            long double c = cosl(th), s = sinl(th);
            sr += re[j] * c - im[j] * s;
            si += re[j] * s + im[j] * c;
        }
        out_re[k] = (double)sr;
        out_im[k] = (double)si;
    }
}

/* E = real(ifft((xi = fft(rho) ./ ik; xi[1] *= 0; xi)))
 * (src/GaussianFixedPointQuiet.jl:4,9; src/NGPFourier.jl:3,5).
 * ik = 2 pi i * [1, 1:N/2, -N/2+1:-1]; the first entry is a placeholder and
 * the k = 0 mode is then zeroed.  Julia's ComplexF64 division by (0, k)
 * (Baudin-Smith robust_cdiv, base/complex.jl) reduces to t = 1/k,
 * xi = (Im(rho^) * t, -(Re(rho^) * t)).  ifft scales by 1/N (exact for
 * power-of-two N) and real() keeps the real part. */
void orc_fieldsolve(int32_t N, const double* rho, double* E) {
    double* re = (double*)malloc(sizeof(double) * (size_t)N);
    double* im = (double*)malloc(sizeof(double) * (size_t)N);
    double* twr = (double*)malloc(sizeof(double) * (size_t)(N / 2 + 1));
    double* twi = (double*)malloc(sizeof(double) * (size_t)(N / 2 + 1));
    orc_twiddles(N, twr, twi);
    for (int32_t j = 0; j < N; ++j) { re[j] = rho[j]; im[j] = 0.0; }
    fft_with(N, re, im, 0, twr, twi);
    const double twopi = 6.283185307179586;
    for (int32_t m = 0; m < N; ++m) {
        int32_t mp = m == 0 ? 1 : (m <= N / 2 ? m : m - N);
        double k = twopi * (double)mp;
        double t = 1.0 / k;
        double xr = im[m] * t;
        double xi = -(re[m] * t);
        if (m == 0) { xr = xr * 0.0; xi = xi * 0.0; }
        re[m] = xr;
        im[m] = xi;
    }
    fft_with(N, re, im, 1, twr, twi);
    const double scale = 1.0 / (double)N;
    for (int32_t j = 0; j < N; ++j) E[j] = re[j] * scale;
    free(re); free(im); free(twr); free(twi);
}

/* ------------------------------------------------------------------------ */
/* Stencils                                                                  */
/* ------------------------------------------------------------------------ */

/* f(x) = Int(mod1(round(x*N), N)) (src/NGPFourier.jl:3), returned 0-based. */
int64_t orc_ngp_cell(double x, int32_t N) {
    double r = rint(x * (double)N);
    return orc_jl_mod1i((int64_t)r, N) - 1;
}

/* d(c) = ((mod1(i,N), ff(i,c)) for i in (-R:R) .+ Int(round(c*N))) with
 * f(g,c) = erf((g-c)*N)/2, ff(i,c) = f((i+0.5)/N,c) - f((i-0.5)/N,c)
 * (src/GaussianFixedPointQuiet.jl:5-6, R = 7; src/GaussianFixedPoint.jl:4-5,
 * R = 6).  All 2(2R+1) erf calls are made literally. */
void orc_gauss_stencil(int32_t N, int32_t R, double c, int64_t* cells, double* wt) {
    const double Nd = (double)N;
    int64_t i0 = (int64_t)rint(c * Nd);
    for (int32_t j = -R; j <= R; ++j) {
        int64_t i = i0 + j;
        double g1 = ((double)i + 0.5) / Nd;
        double g0 = ((double)i - 0.5) / Nd;
        double f1 = orc_erf((g1 - c) * Nd) / 2.0;
        double f0 = orc_erf((g0 - c) * Nd) / 2.0;
        wt[j + R] = f1 - f0;
        cells[j + R] = orc_jl_mod1i(i, N) - 1;
    }
}

/* d(y) = (i = Int(mod1(ceil(y*N), N)); o = ceil(y*N) - y*N;
 *         ((i, 1-o), (mod1(i-1, N), o)))
 * (src/AreaFixedPointQuietSimpson13.jl:5): linear weights onto the grid points
 * i/N and (i-1)/N around y.  ceil(y*N) is an integer-valued Float64, so the
 * Float64 mod1 of the script is the integer mod1 here.  Cells 0-based. */
void orc_area_stencil(int32_t N, double c, int64_t* cells, double* wt) {
    const double yN = c * (double)N;
    const double cy = ceil(yN);
    const int64_t i = orc_jl_mod1i((int64_t)cy, N);
    const double o = cy - yN;
    cells[0] = i - 1;
    wt[0] = 1.0 - o;
    cells[1] = orc_jl_mod1i(i - 1, N) - 1;
    wt[1] = o;
}

/* The stencil of a shape (6 / 7: Gaussian half-width, ORC_SHAPE_AREA: linear
 * area weights); returns the number of cells. */
static int32_t stencil_of(int32_t N, int32_t shape, double c, int64_t* cells, double* wt) {
    if (shape == ORC_SHAPE_AREA) {
        orc_area_stencil(N, c, cells, wt);
        return 2;
    }
    orc_gauss_stencil(N, shape, c, cells, wt);
    return 2 * shape + 1;
}

/* sum(k -> E[k[1]]*k[2], d(c)): a left fold over the stencil in order
 * (src/GaussianFixedPointQuiet.jl:10, src/AreaFixedPointQuietSimpson13.jl:11). */
double orc_gather_at(int32_t N, int32_t shape, const double* E, double c) {
    int64_t cells[2 * 16 + 1];
    double wt[2 * 16 + 1];
    const int32_t n = stencil_of(N, shape, c, cells, wt);
    double s = E[cells[0]] * wt[0];
    for (int32_t k = 1; k < n; ++k) s = s + E[cells[k]] * wt[k];
    return s;
}

/* ------------------------------------------------------------------------ */
/* Deposits                                                                  */
/* ------------------------------------------------------------------------ */
#define MAXR 16

/* NGP:  n .*= 0; for j in x; n[f(j)] += w; end   (src/NGPFourier.jl:5)
 * Gauss: r .*= 0; for l in d.((x.+y)./2); for k in l; r[k[1]] += k[2]*w
 *        (src/GaussianFixedPointQuiet.jl:7; y = x gives centre (x+x)/2 = x).
 * The leapfrog Gaussian deposit of src/Gaussian.jl:7 (k[2]*w/dx) is this
 * deposit with w/dx in place of w (dx a power of two). */
void orc_deposit(int32_t N, int32_t shape, int64_t P, const double* x,
                 const double* y, double w, int32_t mode, double* rho) {
    if (mode == 1) {
        for (int32_t i = 0; i < N; ++i) rho[i] = 0.0;
        if (shape == 0) {
            for (int64_t j = 0; j < P; ++j) rho[orc_ngp_cell(x[j], N)] += w;
            return;
        }
        int64_t cells[2 * MAXR + 1];
        double wt[2 * MAXR + 1];
        for (int64_t j = 0; j < P; ++j) {
            double c = y ? (x[j] + y[j]) / 2.0 : (x[j] + x[j]) / 2.0;
            const int32_t n = stencil_of(N, shape, c, cells, wt);
            for (int32_t k = 0; k < n; ++k) rho[cells[k]] += wt[k] * w;
        }
        return;
    }
    int q = orc_fxp_frac_bits(P, w);
    u128* acc = (u128*)calloc((size_t)N, sizeof(u128));
    if (shape == 0) {
        u128 one = fxp_from(w, q);
        for (int64_t j = 0; j < P; ++j) acc[orc_ngp_cell(x[j], N)] += one;
    } else {
        int64_t cells[2 * MAXR + 1];
        double wt[2 * MAXR + 1];
        for (int64_t j = 0; j < P; ++j) {
            double c = y ? (x[j] + y[j]) / 2.0 : (x[j] + x[j]) / 2.0;
            const int32_t n = stencil_of(N, shape, c, cells, wt);
            for (int32_t k = 0; k < n; ++k) acc[cells[k]] += fxp_from(wt[k] * w, q);
        }
    }
    for (int32_t i = 0; i < N; ++i) rho[i] = fxp_to(acc[i], q);
    free(acc);
}

/* ------------------------------------------------------------------------ */
/* Particle updates                                                          */
/* ------------------------------------------------------------------------ */

/* u() = (x .= mod.(x .+ v/2*dt, 1))   (src/NGPFourier.jl:2) */
void orc_drift_half(int64_t P, double* x, const double* v, double dt) {
    for (int64_t j = 0; j < P; ++j) x[j] = orc_jl_mod1f(x[j] + (v[j] / 2.0) * dt);
}

/* NGP:   v += E[f.(x)]*dt                                (src/NGPFourier.jl:6)
 * Gauss: v[j] += sum(k->E[k[1]]*k[2], d(x[j]))*dt        (src/Gaussian.jl:10)
 * `sum` over the stencil generator is a left fold (mapfoldl). */
void orc_kick(int32_t N, int32_t shape, int64_t P, double* v, const double* x,
              const double* E, double dt) {
    if (shape == 0) {
        for (int64_t j = 0; j < P; ++j) v[j] = v[j] + E[orc_ngp_cell(x[j], N)] * dt;
        return;
    }
    int64_t cells[2 * MAXR + 1];
    double wt[2 * MAXR + 1];
    for (int64_t j = 0; j < P; ++j) {
        orc_gauss_stencil(N, shape, x[j], cells, wt);
        double s = E[cells[0]] * wt[0];
        for (int32_t k = 1; k < 2 * shape + 1; ++k) s = s + E[cells[k]] * wt[k];
        v[j] = v[j] + s * dt;
    }
}

/* x .= X .+ (v .+ V)/2*dt   (src/GaussianFixedPointQuiet.jl:9) */
void orc_midpoint_position(int64_t P, const double* X, const double* V,
                           const double* v, double* x, double dt) {
    for (int64_t j = 0; j < P; ++j) x[j] = X[j] + ((v[j] + V[j]) / 2.0) * dt;
}

/* for j in 1:P; v[j] = V[j] + sum(k->E[k[1]]*k[2], d((x[j]+X[j])/2))*dt; end
 * (src/GaussianFixedPointQuiet.jl:10) */
void orc_push_midpoint(int32_t N, int32_t R, int64_t P, const double* X,
                       const double* V, const double* x, double* v,
                       const double* E, double dt) {
    int64_t cells[2 * MAXR + 1];
    double wt[2 * MAXR + 1];
    for (int64_t j = 0; j < P; ++j) {
        orc_gauss_stencil(N, R, (x[j] + X[j]) / 2.0, cells, wt);
        double s = E[cells[0]] * wt[0];
        for (int32_t k = 1; k < 2 * R + 1; ++k) s = s + E[cells[k]] * wt[k];
        v[j] = V[j] + s * dt;
    }
}

/* x .= mod.(x, 1)   (src/GaussianFixedPointQuiet.jl:10) */
void orc_wrap(int64_t P, double* x) {
    for (int64_t j = 0; j < P; ++j) x[j] = orc_jl_mod1f(x[j]);
}

/* ------------------------------------------------------------------------ */
/* Diagnostics (src/GaussianFixedPointQuiet.jl:11,13):                        */
/*   D[1:2] = (sum(E.^2)/N, sum(v.^2)*W/P) ./ 2                               */
/*   D[3:4] = (D1 + D2, sum(v/P));  D[1:3] .*= 2/W                            */
/* ------------------------------------------------------------------------ */
void orc_diagnostics(int32_t N, int64_t P, const double* E, const double* v,
                     double W, double* out4) {
    double* t = (double*)malloc(sizeof(double) * (size_t)(P > N ? P : N));
    for (int32_t i = 0; i < N; ++i) t[i] = E[i] * E[i];
    double se = orc_tree_sum(t, N);
    for (int64_t j = 0; j < P; ++j) t[j] = v[j] * v[j];
    double sv = orc_tree_sum(t, P);
    for (int64_t j = 0; j < P; ++j) t[j] = v[j] / (double)P;
    double sp = orc_tree_sum(t, P);
    free(t);
    double d1 = (se / (double)N) / 2.0;
    double d2 = ((sv * W) / (double)P) / 2.0;
    double d3 = d1 + d2;
    double s = 2.0 / W;
    out4[0] = d1 * s;
    out4[1] = d2 * s;
    out4[2] = d3 * s;
    out4[3] = sp;
}

/* x = (bitreverse.(0:P-1) .+ 2.0^63) / 2.0^64 over signed Int64;
 * v = collect(1:P .> P/2) .* 2 .- 1.   (src/GaussianFixedPointQuiet.jl:2-3) */
void orc_quiet_start(int64_t P, double* x, double* v) {
    for (int64_t i = 0; i < P; ++i) {
        uint64_t u = (uint64_t)i, r = 0;
        for (int b = 0; b < 64; ++b) { r = (r << 1) | (u & 1); u >>= 1; }
        double f = (double)(int64_t)r;
        x[i] = (f + 9223372036854775808.0) / 18446744073709551616.0;
        v[i] = (double)(i + 1) > (double)P / 2.0 ? 1.0 : -1.0;
    }
}

/* ------------------------------------------------------------------------ */
/* Whole runs                                                                */
/* ------------------------------------------------------------------------ */

/* Parallel exact deposit: per-thread u128 grids summed afterwards (integer
 * sums, so bitwise equal to the serial deposit). */
void orc_deposit_par(int32_t N, int32_t shape, int64_t P, const double* x,
                     const double* y, double w, double* rho, int32_t nthreads) {
#ifdef _OPENMP
    if (nthreads > 1) {
        int q = orc_fxp_frac_bits(P, w);
        u128* acc = (u128*)calloc((size_t)N * (size_t)nthreads, sizeof(u128));
#pragma omp parallel num_threads(nthreads)
        {
            int tid = omp_get_thread_num();
            u128* my = acc + (size_t)tid * (size_t)N;
            int64_t cells[2 * MAXR + 1];
            double wt[2 * MAXR + 1];
#pragma omp for schedule(static)
            for (int64_t j = 0; j < P; ++j) {
                if (shape == 0) {
                    my[orc_ngp_cell(x[j], N)] += fxp_from(w, q);
                } else {
                    double c = y ? (x[j] + y[j]) / 2.0 : (x[j] + x[j]) / 2.0;
                    const int32_t n = stencil_of(N, shape, c, cells, wt);
                    for (int32_t k = 0; k < n; ++k) my[cells[k]] += fxp_from(wt[k] * w, q);
                }
            }
        }
        for (int32_t i = 0; i < N; ++i) {
            u128 s = 0;
            for (int t = 0; t < nthreads; ++t) s += acc[(size_t)t * (size_t)N + (size_t)i];
            rho[i] = fxp_to(s, q);
        }
        free(acc);
        return;
    }
#endif
    (void)nthreads;
    orc_deposit(N, shape, P, x, y, w, 0, rho);
}

static void push_midpoint_par(int32_t N, int32_t R, int64_t P, const double* X,
                              const double* V, const double* x, double* v,
                              const double* E, double dt, int nthreads) {
#ifdef _OPENMP
    if (nthreads > 1) {
#pragma omp parallel for num_threads(nthreads) schedule(static)
        for (int64_t j = 0; j < P; ++j) {
            int64_t cells[2 * MAXR + 1];
            double wt[2 * MAXR + 1];
            orc_gauss_stencil(N, R, (x[j] + X[j]) / 2.0, cells, wt);
            double s = E[cells[0]] * wt[0];
            for (int32_t k = 1; k < 2 * R + 1; ++k) s = s + E[cells[k]] * wt[k];
            v[j] = V[j] + s * dt;
        }
        return;
    }
#endif
    (void)nthreads;
    orc_push_midpoint(N, R, P, X, V, x, v, E, dt);
}

int orc_run(int32_t N, int32_t shape, int32_t scheme, int64_t P, double W,
            double w, double dt, double rtol, int32_t max_sweeps,
            double* x, double* v, double* E, int64_t nsteps,
            double* diag, int32_t* sweeps, int32_t deposit_mode,
            int32_t nthreads) {
    if (N < 4 || (N & (N - 1)) != 0 || P < 1 || shape < 0 || shape > MAXR) return -1;
    if (scheme == 1 && shape == 0) return -1;
    double* rho = (double*)malloc(sizeof(double) * (size_t)N);
    if (scheme == 0) {
        for (int64_t t = 0; t < nsteps; ++t) {
            orc_drift_half(P, x, v, dt);
            if (deposit_mode == 0) orc_deposit_par(N, shape, P, x, NULL, w, rho, nthreads);
            else orc_deposit(N, shape, P, x, NULL, w, 1, rho);
            orc_fieldsolve(N, rho, E);
            orc_drift_half(P, x, v, dt);
            orc_kick(N, shape, P, v, x, E, dt);
            if (diag) orc_diagnostics(N, P, E, v, W, diag + 4 * t);
            if (sweeps) sweeps[t] = 1;
        }
        free(rho);
        return 0;
    }
    double* X = (double*)malloc(sizeof(double) * (size_t)P);
    double* V = (double*)malloc(sizeof(double) * (size_t)P);
    double* F = (double*)malloc(sizeof(double) * (size_t)N);
    for (int64_t t = 0; t < nsteps; ++t) {
        memcpy(X, x, sizeof(double) * (size_t)P);
        memcpy(V, v, sizeof(double) * (size_t)P);
        for (int32_t i = 0; i < N; ++i) F[i] = NAN; /* F .*= NaN */
        int32_t it;
        for (it = 0; it < max_sweeps; ++it) {
            if (orc_isapprox(F, E, N, rtol, 0.0)) break;
            memcpy(F, E, sizeof(double) * (size_t)N);
            orc_midpoint_position(P, X, V, v, x, dt);
            if (deposit_mode == 0) orc_deposit_par(N, shape, P, x, X, w, rho, nthreads);
            else orc_deposit(N, shape, P, x, X, w, 1, rho);
            orc_fieldsolve(N, rho, E);
            push_midpoint_par(N, shape, P, X, V, x, v, E, dt, nthreads);
        }
        orc_wrap(P, x);
        if (diag) orc_diagnostics(N, P, E, v, W, diag + 4 * t);
        if (sweeps) sweeps[t] = it;
    }
    free(X); free(V); free(F); free(rho);
    return 0;
}
