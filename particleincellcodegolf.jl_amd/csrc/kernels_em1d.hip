// kernels_em1d.hip -- FP64 gfx950 kernels of the 1D electromagnetic
// fixed-point step (src/GaussianEMFixedPoint.jl).  Compiled with
// -ffp-contract=off: results are bitwise equal to the CPU restatement.
//
// A sweep is four launches, each predicated on the sweep's flag so a step
// graph holds every sweep: move (x and the midpoints, the current maxima),
// deposit (charge and both currents into exact LDS limbs), solve (one block:
// Gauss, Ampere, Faraday and the test for the next sweep) and push (the
// velocity update).  The step's end wraps x, sorts the particles by x
// (bucket by position, then each particle's rank in its bucket by its key and
// index), deposits the seven moments, and writes the step's twelve fields.
//
// The explicit scheme (src/GaussianEMExplicit.jl) reuses the stencil, the
// exact limbs, the sort and the moments: begin (X, the charge, the move and the
// sort keys), a one-block Gauss and Faraday solve, the sort (X stays
// unsorted), the split-kick push, the current at (X + x)/2, a one-block Ampere
// solve, the moments and the step's fields.
#include <climits>

#include "fft_lines.h"
#include "pic_device.h"
#include "picem1d_internal.h"
#include "solve1d.h"

#pragma clang fp contract(off)

namespace pic {

constexpr u64 kEMLimbMask = (1ull << kLimbEM) - 1;

__device__ __forceinline__ int em_stride(int N) { return N + (N >> 4); }

// irn(c) = Int(mod1(round(c*N), N)) (:8)
__device__ __forceinline__ int64_t em_irn(double c, int N) {
    const double r = rint(c * (double)N);
    if (!(r > -4.0e18 && r < 4.0e18)) return N;
    const int64_t m = (int64_t)r & (int64_t)(N - 1);
    return m == 0 ? N : m;
}

// s(c) (:9-10): cells mod1(i, N) (0-based) and ff(i, c) for i = irn-6 .. irn+6.
// Adjacent cells share an edge: ((i+1) - 0.5) equals (i + 0.5) exactly.
__device__ __forceinline__ void em_stencil(double c, int N, double Nd, int32_t* cell, double* ff) {
    const int64_t i0 = em_irn(c, N);
    double lo = ol_erf((((double)(i0 - 6) - 0.5) / Nd - c) * Nd) / 2.0;
#pragma unroll
    for (int k = 0; k < 13; ++k) {
        const int64_t i = i0 - 6 + k;
        const double hi = ol_erf((((double)i + 0.5) / Nd - c) * Nd) / 2.0;
        ff[k] = hi - lo;
        lo = hi;
        cell[k] = (int32_t)((i - 1) & (int64_t)(N - 1));
    }
}

// P*|w|*M*2^q < 2^111
__device__ __forceinline__ int em_frac_bits(int64_t P, double w, double M) {
    const double total = (double)P * (fabs(w) * M);
    if (!(total > 0.0) || isinf(total) || total != total) return 64;
    const int q = 110 - ilogb(total);
    return q > 1100 ? 1100 : q;
}

__device__ __forceinline__ void em_add(u64* g, int N, int cell, i128 v) {
    const u128 u = (u128)v;
    const u64 l0 = (u64)u & kEMLimbMask;
    const u64 l1 = (u64)(u >> kLimbEM) & kEMLimbMask;
    const u64 l2 = (u64)(u >> (2 * kLimbEM)) & kEMLimbMask;
    if (l0) atomicAdd(g + cell, l0);
    if (l1) atomicAdd(g + N + cell, l1);
    if (l2) atomicAdd(g + 2 * N + cell, l2);
}

// the exact sum of a cell (three limb counters, modulo 2^114) rounded once
__device__ __forceinline__ double em_sum(u64 a, u64 b, u64 c, int q) {
    u128 v = (u128)a + ((u128)b << kLimbEM) + ((u128)c << (2 * kLimbEM));
    v &= ((u128)1 << (3 * kLimbEM)) - 1;
    const i128 s = (v >> (3 * kLimbEM - 1)) ? (i128)v - ((i128)1 << (3 * kLimbEM)) : (i128)v;
    return i128_to_double(s, q);
}

__device__ __forceinline__ double bits_d(u64 b) { return __longlong_as_double((long long)b); }
__device__ __forceinline__ u64 d_bits(double d) { return (u64)__double_as_longlong(d); }

// block-wide max of non-negative doubles, atomically into *dst (as bits)
__device__ __forceinline__ void block_max_into(double v, u64* dst, u64* red) {
    u64 b = d_bits(v);
#pragma unroll
    for (int h = 32; h > 0; h >>= 1) {
        const u64 o = __shfl_xor(b, h, 64);
        b = o > b ? o : b;
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = b;
    __syncthreads();
    if (threadIdx.x == 0) {
        u64 m = red[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = red[w] > m ? red[w] : m;
        atomicMax(dst, m);
    }
    __syncthreads();
}

__device__ __forceinline__ double fmax_num(double a, double b) { return b > a ? b : a; }

// ---------------------------------------------------------------------------
// sweep kernels
// ---------------------------------------------------------------------------
// x = X + ((V1 + v1)/2)*dt, c = (X + x)/2 (:15, :20); the first sweep also
// takes X, V from the particles (:13).  Maxima of |uy|, |uz| for the currents'
// fixed-point grids.
template <bool FIRST>
__global__ void __launch_bounds__(kEMThreads) k_em_move(EM1D s, int sweep) {
    __shared__ u64 red[kEMThreads / 64];
    if (s.ctl[sweep] == 0) return;
    double my = 0.0, mz = 0.0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < s.P; p += stride) {
        double X, V1, V2, V3;
        const double v1 = s.v1[p], v2 = s.v2[p], v3 = s.v3[p];
        if (FIRST) {
            X = s.x[p];
            V1 = v1;
            V2 = v2;
            V3 = v3;
            s.X[p] = X;
            s.V1[p] = V1;
            s.V2[p] = V2;
            s.V3[p] = V3;
        } else {
            X = s.X[p];
            V1 = s.V1[p];
            V2 = s.V2[p];
            V3 = s.V3[p];
        }
        const double x = X + (V1 + v1) / 2.0 * s.dt;
        s.x[p] = x;
        s.c[p] = (X + x) / 2.0;
        my = fmax_num(my, fabs((V2 + v2) / 2.0));
        mz = fmax_num(mz, fabs((V3 + v3) / 2.0));
    }
    block_max_into(my, s.umax + 0, red);
    block_max_into(mz, s.umax + 1, red);
}

// d((X+x)/2), d((X+x)/2, (V2+v2)/2), d((X+x)/2, (V3+v3)/2) (:20, :22, :23)
__global__ void __launch_bounds__(kEMDepThreads) k_em_deposit3(EM1D s, int sweep) {
    extern __shared__ u64 ldsu[];
    if (s.ctl[sweep] == 0) return;
    const int N = s.N;
    for (int i = threadIdx.x; i < 9 * N; i += blockDim.x) ldsu[i] = 0;
    __syncthreads();
    const double Nd = (double)N;
    const int q0 = em_frac_bits(s.P, s.w, 1.0);
    const int q1 = em_frac_bits(s.P, s.w, bits_d(s.umax[0]));
    const int q2 = em_frac_bits(s.P, s.w, bits_d(s.umax[1]));
    const int64_t p0 = (int64_t)blockIdx.x * s.tile;
    const int64_t p1 = p0 + s.tile < s.P ? p0 + s.tile : s.P;
    for (int64_t p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
        const double c = s.c[p];
        const double uy = (s.V2[p] + s.v2[p]) / 2.0, uz = (s.V3[p] + s.v3[p]) / 2.0;
        int32_t cell[13];
        double ff[13];
        em_stencil(c, N, Nd, cell, ff);
#pragma unroll
        for (int k = 0; k < 13; ++k) {
            const double a0 = ff[k] * s.w, a1 = uy * ff[k] * s.w, a2 = uz * ff[k] * s.w;
            if (a0 != 0.0) em_add(ldsu, N, cell[k], fix_round128(a0, q0));
            if (a1 != 0.0) em_add(ldsu + 3 * N, N, cell[k], fix_round128(a1, q1));
            if (a2 != 0.0) em_add(ldsu + 6 * N, N, cell[k], fix_round128(a2, q2));
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 9 * N; i += blockDim.x) {
        const u64 v = ldsu[i];
        if (v) atomicAdd(s.racc + i, v);
    }
}

// Complex products of the script: Julia's (ac - bd, ad + bc).
__device__ __forceinline__ void cmul(double ar, double ai, double br, double bi, double& cr, double& ci) {
    cr = ar * br - ai * bi;
    ci = ar * bi + ai * br;
}

__device__ __forceinline__ void ik_of(int N, int j, double& kr, double& ki) {
    const double m = (double)(j <= N / 2 ? j : j - N);
    kr = 0.0 * m;
    ki = kTwoPi * m;
}

// One sweep's field updates (:20-27) by one block of any size with
// N <= JPT*blockDim.x.  SRC 0: rho, Jy, Jz from the exact limbs g (rho, Jy, Jz
// at 0, 3N, 6N, three limbs each; cleared here) with fraction bits q0..q2; 1:
// from the arrays rin, jyin, jzin.  RUN: the step's saves at the first sweep,
// Fx .= Ex, and the return value is the test of the next sweep (:14).  sm:
// em_solve_lds_bytes(N) of LDS.
template <int SRC, bool RUN, int JPT>
__device__ bool em_solve_body(const EM1D& s, int sweep, u64* g, const double* __restrict__ rin,
                              const double* __restrict__ jyin, const double* __restrict__ jzin, int q0, int q1, int q2,
                              double* sm) {
    const int N = s.N, lN = s.lN, st = em_stride(N);
    const int t = threadIdx.x, nth = blockDim.x;
    double* re = sm;
    double* im = sm + 5 * st;
    double* red = sm + 10 * st;
    const double Nd = (double)N, sc = 1.0 / Nd, dt = s.dt, e2 = s.eps * s.eps;
    if (RUN) {
        if (sweep == 0)
            for (int j = t; j < N; j += blockDim.x) {  // Fy .= Ey; Fz .*= Ez; Cy .= By; Cz .= Bz (:13)
                s.Fy[j] = s.Ey[j];
                s.Fz[j] = s.Fz[j] * s.Ez[j];
                s.Cy[j] = s.By[j];
                s.Cz[j] = s.Bz[j];
            }
        for (int j = t; j < N; j += blockDim.x) s.Fx[j] = s.Ex[j];  // Fx .= Ex (:14)
    }
    __syncthreads();
    // lines: 0 rho, 1 bz, 2 Jy, 3 by, 4 Jz (bit-reversed, imaginary zero)
    for (int j = t; j < N; j += blockDim.x) {
        double r, jy, jz;
        if (SRC == 0) {
            r = em_sum(g[j], g[N + j], g[2 * N + j], q0);
            jy = em_sum(g[3 * N + j], g[4 * N + j], g[5 * N + j], q1);
            jz = em_sum(g[6 * N + j], g[7 * N + j], g[8 * N + j], q2);
            for (int l = 0; l < 9; ++l) g[l * N + j] = 0;
        } else {
            r = rin[j];
            jy = jyin[j];
            jz = jzin[j];
        }
        const double bzv = (s.Bz[j] + s.Cz[j]) / 2.0, byv = (s.By[j] + s.Cy[j]) / 2.0;
        const int pos = padi((int)brev((uint32_t)j, lN));
        const double v[5] = {r, bzv, jy, byv, jz};
#pragma unroll
        for (int l = 0; l < 5; ++l) {
            re[l * st + pos] = v[l];
            im[l * st + pos] = 0.0;
        }
    }
    __syncthreads();
    fft_lines(re, im, s.tw, lN, 5, st, false);
    // Gauss, Ampere y, Ampere z in the spectrum
    double o[JPT][6];
#pragma unroll
    for (int q = 0; q < JPT; ++q) {
        const int j = t + q * nth;
        if (j >= N) break;
        const int pj = padi(j);
        double kr, ki;
        ik_of(N, j, kr, ki);
        if (j == 0) {  // (0,0)/(0,0) is NaN; `*= false` keeps its (x86: negative) sign
            o[q][0] = -0.0;
            o[q][1] = -0.0;
        } else {
            const double tt = 1.0 / ki;
            o[q][0] = im[pj] * tt;
            o[q][1] = -(re[pj] * tt);
        }
        double pr, pi;
        cmul(-kr, -ki, re[st + pj], im[st + pj], pr, pi);
        o[q][2] = pr / e2 - re[2 * st + pj];
        o[q][3] = pi / e2 - im[2 * st + pj];
        cmul(kr, ki, re[3 * st + pj], im[3 * st + pj], pr, pi);
        o[q][4] = pr / e2 - re[4 * st + pj];
        o[q][5] = pi / e2 - im[4 * st + pj];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < JPT; ++q) {
        const int j = t + q * nth;
        if (j >= N) break;
        const int pos = padi((int)brev((uint32_t)j, lN));
#pragma unroll
        for (int l = 0; l < 3; ++l) {
            re[l * st + pos] = o[q][2 * l];
            im[l * st + pos] = o[q][2 * l + 1];
        }
    }
    __syncthreads();
    fft_lines(re, im, s.tw, lN, 3, st, true);
    double ezr[JPT], eyr[JPT];
#pragma unroll
    for (int q = 0; q < JPT; ++q) {
        const int j = t + q * nth;
        if (j >= N) break;
        const int pj = padi(j);
        const double Ex = re[pj] * sc;
        const double Ey = s.Fy[j] + dt * (re[st + pj] * sc);
        const double Ez = s.Fz[j] + dt * (re[2 * st + pj] * sc);
        s.Ex[j] = Ex;
        s.Ey[j] = Ey;
        s.Ez[j] = Ez;
        s.ex[j] = (Ex + s.Fx[j]) / 2.0;
        eyr[q] = (Ey + s.Fy[j]) / 2.0;
        ezr[q] = (Ez + s.Fz[j]) / 2.0;
        s.ey[j] = eyr[q];
        s.ez[j] = ezr[q];
    }
    __syncthreads();
    // Faraday: lines 0 ez, 1 ey
#pragma unroll
    for (int q = 0; q < JPT; ++q) {
        const int j = t + q * nth;
        if (j >= N) break;
        const int pos = padi((int)brev((uint32_t)j, lN));
        re[pos] = ezr[q];
        im[pos] = 0.0;
        re[st + pos] = eyr[q];
        im[st + pos] = 0.0;
    }
    __syncthreads();
    fft_lines(re, im, s.tw, lN, 2, st, false);
#pragma unroll
    for (int q = 0; q < JPT; ++q) {
        const int j = t + q * nth;
        if (j >= N) break;
        const int pj = padi(j);
        double kr, ki;
        ik_of(N, j, kr, ki);
        cmul(-kr, -ki, re[pj], im[pj], o[q][0], o[q][1]);
        cmul(kr, ki, re[st + pj], im[st + pj], o[q][2], o[q][3]);
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < JPT; ++q) {
        const int j = t + q * nth;
        if (j >= N) break;
        const int pos = padi((int)brev((uint32_t)j, lN));
        re[pos] = o[q][0];
        im[pos] = o[q][1];
        re[st + pos] = o[q][2];
        im[st + pos] = o[q][3];
    }
    __syncthreads();
    fft_lines(re, im, s.tw, lN, 2, st, true);
    for (int j = t; j < N; j += blockDim.x) {
        const int pj = padi(j);
        const double By = s.Cy[j] - dt * (re[pj] * sc);
        const double Bz = s.Cz[j] - dt * (re[st + pj] * sc);
        s.By[j] = By;
        s.Bz[j] = Bz;
        s.by[j] = (By + s.Cy[j]) / 2.0;
        s.bz[j] = (Bz + s.Cz[j]) / 2.0;
    }
    if (!RUN) return false;
    __syncthreads();
    // isapprox(Fx, Ex): the test at the top of the next sweep (:14)
    double me;
    return block_isapprox(red, [&](int i) { return s.Fx[i]; }, [&](int i) { return s.Ex[i]; }, N, s.rtol, 0.0, me);
}

template <int SRC, bool RUN>
__global__ void __launch_bounds__(kEMSolveThreads)
k_em_solve(EM1D s, int sweep, const double* __restrict__ rin, const double* __restrict__ jyin,
           const double* __restrict__ jzin) {
    extern __shared__ double sm[];
    if (RUN && s.ctl[sweep] == 0) return;
    int q0 = 0, q1 = 0, q2 = 0;
    if (SRC == 0) {
        q0 = em_frac_bits(s.P, s.w, 1.0);
        q1 = em_frac_bits(s.P, s.w, bits_d(s.umax[0]));
        q2 = em_frac_bits(s.P, s.w, bits_d(s.umax[1]));
    }
    const bool conv =
        em_solve_body<SRC, RUN, kMaxNEM / kEMSolveThreads>(s, sweep, s.racc, rin, jyin, jzin, q0, q1, q2, sm);
    if (RUN && threadIdx.x == 0) {
        s.ctl[kEMMaxSweeps] += 1;
        if (sweep + 1 < s.max_sweeps) s.ctl[sweep + 1] = conv ? 0 : 1;
        s.umax[0] = 0;
        s.umax[1] = 0;
    }
}


// v = u + (a + cross(u + a*dt/2, b))*dt (:28-29) with a, b gathered at c.
__device__ __forceinline__ void em_velocity(const double* fe, int N, double Nd, double c, double dt, double u1,
                                            double u2, double u3, double& v1, double& v2, double& v3) {
    int32_t cell[13];
    double ff[13];
    em_stencil(c, N, Nd, cell, ff);
    double a1 = 0, a2 = 0, a3 = 0, b1 = 0, b2 = 0, b3 = 0;
#pragma unroll
    for (int k = 0; k < 13; ++k) {
        const int m = cell[k];
        const double t1 = fe[m] * ff[k], t2 = fe[N + m] * ff[k], t3 = fe[2 * N + m] * ff[k];
        const double s1 = 0.0 * ff[k], s2 = fe[3 * N + m] * ff[k], s3 = fe[4 * N + m] * ff[k];
        if (k == 0) {
            a1 = t1;
            a2 = t2;
            a3 = t3;
            b1 = s1;
            b2 = s2;
            b3 = s3;
        } else {
            a1 = a1 + t1;
            a2 = a2 + t2;
            a3 = a3 + t3;
            b1 = b1 + s1;
            b2 = b2 + s2;
            b3 = b3 + s3;
        }
    }
    const double w1 = u1 + a1 * dt / 2.0, w2 = u2 + a2 * dt / 2.0, w3 = u3 + a3 * dt / 2.0;
    const double c1 = w2 * b3 - w3 * b2, c2 = w3 * b1 - w1 * b3, c3 = w1 * b2 - w2 * b1;
    v1 = u1 + (a1 + c1) * dt;
    v2 = u2 + (a2 + c2) * dt;
    v3 = u3 + (a3 + c3) * dt;
}

// stage ex, ey, ez, by, bz (5N doubles) in LDS
__device__ __forceinline__ void stage_fields(double* fe, const EM1D& s) {
    const int N = s.N;
    for (int j = threadIdx.x; j < N; j += blockDim.x) {
        fe[j] = s.ex[j];
        fe[N + j] = s.ey[j];
        fe[2 * N + j] = s.ez[j];
        fe[3 * N + j] = s.by[j];
        fe[4 * N + j] = s.bz[j];
    }
    __syncthreads();
}

__global__ void __launch_bounds__(kEMThreads) k_em_push(EM1D s, int sweep) {
    extern __shared__ double fe[];
    if (s.ctl[sweep] == 0) return;
    stage_fields(fe, s);
    const double Nd = (double)s.N;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < s.P; p += stride) {
        double v1, v2, v3;
        em_velocity(fe, s.N, Nd, s.c[p], s.dt, s.V1[p], s.V2[p], s.V3[p], v1, v2, v3);
        s.v1[p] = v1;
        s.v2[p] = v2;
        s.v3[p] = v3;
    }
}

// ---------------------------------------------------------------------------
// the step's end: x .= mod.(x, 1); p = sortperm(x) (:41); the moments and the
// fields of the step (:45-49)
// ---------------------------------------------------------------------------
__device__ __forceinline__ u64 em_order_key(double v) {
    if (v != v) return ~0ull;
    const u64 b = d_bits(v);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

__device__ __forceinline__ uint32_t em_bucket(double x, int nb) {
    if (!(x >= 0.0)) return x != x ? (uint32_t)(nb - 1) : 0u;  // NaN last; mod(x, 1) is never negative
    const double f = floor(x * (double)nb);
    return f >= (double)(nb - 1) ? (uint32_t)(nb - 1) : (uint32_t)f;
}

__global__ void __launch_bounds__(256) k_em_sort_zero(EM1D s) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < s.nb; i += gridDim.x * blockDim.x) {
        s.count[i] = 0;
        s.cursor[i] = 0;
    }
}

// x = mod(x, 1); SORT: the particle's bucket and key; else the moment maxima.
template <bool SORT>
__global__ void __launch_bounds__(kEMThreads) k_em_wrap(EM1D s) {
    __shared__ u64 red[kEMThreads / 64];
    double m1 = 0.0, m2 = 0.0, m3 = 0.0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < s.P; p += stride) {
        const double x = jl_mod1f(s.x[p]);
        s.x[p] = x;
        if (SORT) {
            const uint32_t b = em_bucket(x, s.nb);
            s.bucket[p] = b;
            s.vkey[p] = em_order_key(x);
            atomicAdd(s.count + b, 1u);
        } else {
            m1 = fmax_num(m1, fabs(s.v1[p]));
            m2 = fmax_num(m2, fabs(s.v2[p]));
            m3 = fmax_num(m3, fabs(s.v3[p]));
        }
    }
    if (!SORT) {
        block_max_into(m1, s.umax + 2, red);
        block_max_into(m2, s.umax + 3, red);
        block_max_into(m3, s.umax + 4, red);
    }
}

// exclusive scan of the bucket counts: per 4096-bucket tile, then the tile sums
__global__ void __launch_bounds__(1024) k_em_scan_tiles(EM1D s) {
    __shared__ uint32_t sums[1024];
    const int t = threadIdx.x;
    const int64_t base = (int64_t)blockIdx.x * kEMScanTile;
    uint32_t v[4], loc = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int64_t i = base + 4 * t + k;
        v[k] = i < s.nb ? s.count[i] : 0u;
        loc += v[k];
    }
    sums[t] = loc;
    __syncthreads();
    for (int h = 1; h < 1024; h <<= 1) {
        const uint32_t a = t >= h ? sums[t - h] : 0u;
        __syncthreads();
        sums[t] += a;
        __syncthreads();
    }
    uint32_t run = t > 0 ? sums[t - 1] : 0u;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int64_t i = base + 4 * t + k;
        if (i < s.nb) s.start[i] = run;
        run += v[k];
    }
    if (t == 1023) s.bsum[blockIdx.x] = sums[1023];
}

__global__ void __launch_bounds__(256) k_em_scan_top(EM1D s, int ntiles) {
    __shared__ uint32_t sums[256];
    const int t = threadIdx.x;
    sums[t] = t < ntiles ? s.bsum[t] : 0u;
    __syncthreads();
    for (int h = 1; h < 256; h <<= 1) {
        const uint32_t a = t >= h ? sums[t - h] : 0u;
        __syncthreads();
        sums[t] += a;
        __syncthreads();
    }
    if (t < ntiles) s.bsum[t] = t > 0 ? sums[t - 1] : 0u;  // exclusive tile offsets
}

__global__ void __launch_bounds__(256) k_em_scan_add(EM1D s) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < s.nb; i += (int64_t)gridDim.x * blockDim.x)
        s.start[i] += s.bsum[i / kEMScanTile];
    if (blockIdx.x == 0 && threadIdx.x == 0) s.start[s.nb] = (uint32_t)s.P;
}

__global__ void __launch_bounds__(256) k_em_sort_scatter(EM1D s) {
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < s.P; p += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t b = s.bucket[p];
        const uint32_t slot = s.start[b] + atomicAdd(s.cursor + b, 1u);
        s.skey[slot] = s.vkey[p];
        s.sidx[slot] = (uint32_t)p;
        s.sbucket[slot] = b;
    }
}

__global__ void __launch_bounds__(256) k_em_sort_rank(EM1D s) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < s.P; e += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t b = s.sbucket[e];
        const uint32_t s0 = s.start[b], s1 = s.start[b + 1];
        const u64 k = s.skey[e];
        const uint32_t i = s.sidx[e];
        uint32_t rank = 0;
        for (uint32_t f = s0; f < s1; ++f) {
            const u64 kf = s.skey[f];
            const uint32_t jf = s.sidx[f];
            rank += (kf < k || (kf == k && jf < i)) ? 1u : 0u;
        }
        s.perm[s0 + rank] = i;
    }
}

__global__ void __launch_bounds__(256) k_em_sort_apply(EM1D s) {
    const int64_t P = s.P;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < P; k += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t p = s.perm[k];
        s.tmp[k] = s.x[p];
        s.tmp[P + k] = s.v1[p];
        s.tmp[2 * P + k] = s.v2[p];
        s.tmp[3 * P + k] = s.v3[p];
    }
}

// the sorted particles back in place, and (MAX) the moment maxima
template <bool MAX>
__global__ void __launch_bounds__(kEMThreads) k_em_sort_back(EM1D s) {
    __shared__ u64 red[kEMThreads / 64];
    const int64_t P = s.P;
    double m1 = 0.0, m2 = 0.0, m3 = 0.0;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < P; k += (int64_t)gridDim.x * blockDim.x) {
        const double a = s.tmp[P + k], b = s.tmp[2 * P + k], c = s.tmp[3 * P + k];
        s.x[k] = s.tmp[k];
        s.v1[k] = a;
        s.v2[k] = b;
        s.v3[k] = c;
        m1 = fmax_num(m1, fabs(a));
        m2 = fmax_num(m2, fabs(b));
        m3 = fmax_num(m3, fabs(c));
    }
    if (!MAX) return;
    block_max_into(m1, s.umax + 2, red);
    block_max_into(m2, s.umax + 3, red);
    block_max_into(m3, s.umax + 4, red);
}

// sortperm(x) (:41) of at most kEMSmallSort particles in one block: a bitonic
// sort of the (key, index) pairs in LDS, a total order that equals the stable
// sort, then the particles permuted in place.  It replaces the nine launches
// of the bucket sort at the scripts' sizes.  WRAP: x = mod(x, 1) first; MAX:
// the moment maxima of the sorted velocities.
template <bool WRAP, bool MAX>
__global__ void __launch_bounds__(1024) k_em_sort_small(EM1D s) {
    __shared__ u64 key[kEMSmallSort];
    __shared__ uint32_t idx[kEMSmallSort];
    __shared__ u64 red[1024 / 64];
    constexpr int R = kEMSmallSort / 1024;
    const int t = threadIdx.x;
    const int P = (int)s.P;
    int n = 1;
    while (n < P) n <<= 1;
    for (int i = t; i < n; i += 1024) {
        if (i < P) {
            double x = s.x[i];
            if (WRAP) {
                x = jl_mod1f(x);
                s.x[i] = x;
            }
            key[i] = em_order_key(x);
            idx[i] = (uint32_t)i;
        } else {
            key[i] = ~0ull;  // after every particle, NaN keys included
            idx[i] = 0xFFFFFFFFu;
        }
    }
    __syncthreads();
    for (int k = 2; k <= n; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = t; i < n; i += 1024) {
                const int l = i ^ j;
                if (l > i) {
                    const u64 ka = key[i], kb = key[l];
                    const uint32_t ia = idx[i], ib = idx[l];
                    const bool gt = ka > kb || (ka == kb && ia > ib);
                    if (gt == ((i & k) == 0)) {
                        key[i] = kb;
                        key[l] = ka;
                        idx[i] = ib;
                        idx[l] = ia;
                    }
                }
            }
            __syncthreads();
        }
    double xv[R], a[R], b[R], c[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int k = t + r * 1024;
        if (k < P) {
            const uint32_t p = idx[k];
            s.perm[k] = p;
            xv[r] = s.x[p];
            a[r] = s.v1[p];
            b[r] = s.v2[p];
            c[r] = s.v3[p];
        }
    }
    __syncthreads();  // every read of the old order before the first write
    double m1 = 0.0, m2 = 0.0, m3 = 0.0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int k = t + r * 1024;
        if (k < P) {
            s.x[k] = xv[r];
            s.v1[k] = a[r];
            s.v2[k] = b[r];
            s.v3[k] = c[r];
            m1 = fmax_num(m1, fabs(a[r]));
            m2 = fmax_num(m2, fabs(b[r]));
            m3 = fmax_num(m3, fabs(c[r]));
        }
    }
    if (!MAX) return;
    block_max_into(m1, s.umax + 2, red);
    block_max_into(m2, s.umax + 3, red);
    block_max_into(m3, s.umax + 4, red);
}

// moment grid g (0: d(x); 1-3: d(x, v_g); 4-6: d(x, v_{g-3}.^2)) of the
// particle's velocity, and the fraction bits of its grid
__device__ __forceinline__ double moment_u(int g, double v1, double v2, double v3) {
    const double v = g == 1 || g == 4 ? v1 : (g == 2 || g == 5 ? v2 : v3);
    return g <= 3 ? v : v * v;
}

__device__ __forceinline__ int moment_q(const EM1D& s, int g) {
    if (g == 0) return em_frac_bits(s.P, s.w, 1.0);
    const double m = bits_d(s.umax[2 + (g - 1) % 3]);
    return em_frac_bits(s.P, s.w, g <= 3 ? m : m * m);
}

// the seven moment deposits at the final x; block (tile, group): grids 2*group
// and 2*group + 1
__global__ void __launch_bounds__(kEMThreads) k_em_moments(EM1D s) {
    extern __shared__ u64 ldsu[];
    const int N = s.N;
    const int g0 = 2 * blockIdx.y, ng = g0 + 1 < 7 ? 2 : 1;
    for (int i = threadIdx.x; i < 3 * N * ng; i += blockDim.x) ldsu[i] = 0;
    __syncthreads();
    const double Nd = (double)N;
    int q[2];
    for (int k = 0; k < ng; ++k) q[k] = moment_q(s, g0 + k);
    const int64_t p0 = (int64_t)blockIdx.x * s.tile;
    const int64_t p1 = p0 + s.tile < s.P ? p0 + s.tile : s.P;
    for (int64_t p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
        int32_t cell[13];
        double ff[13];
        em_stencil(s.x[p], N, Nd, cell, ff);
        const double v1 = s.v1[p], v2 = s.v2[p], v3 = s.v3[p];
        for (int k = 0; k < ng; ++k) {
            const int g = g0 + k;
            const double u = g == 0 ? 1.0 : moment_u(g, v1, v2, v3);
#pragma unroll
            for (int j = 0; j < 13; ++j) {
                const double a = (g == 0 ? ff[j] : u * ff[j]) * s.w;
                if (a != 0.0) em_add(ldsu + 3 * N * k, N, cell[j], fix_round128(a, q[k]));
            }
        }
    }
    __syncthreads();
    u64* dst = s.racc + 9 * N + 3 * N * g0;
    for (int i = threadIdx.x; i < 3 * N * ng; i += blockDim.x) {
        const u64 v = ldsu[i];
        if (v) atomicAdd(dst + i, v);
    }
}

// The step's output slot (fields[:, t, 1:12]) and the next step's flags.
template <bool MOMENTS>
__global__ void __launch_bounds__(256) k_em_finish(EM1D s) {
    const int N = s.N;
    const int64_t slot = s.t[0] - s.t[1];
    const bool keep = slot >= 0 && slot < s.out_cap;
    double* o = s.out + (keep ? slot : 0) * 12 * N;
    int q[7];
    if (MOMENTS)
        for (int g = 0; g < 7; ++g) q[g] = moment_q(s, g);
    for (int j = threadIdx.x; j < N; j += blockDim.x) {
        if (keep) {
            o[j] = s.Ex[j];
            o[N + j] = s.Ey[j];
            o[2 * N + j] = s.Ez[j];
            o[3 * N + j] = s.By[j];
            o[4 * N + j] = s.Bz[j];
        }
        if (MOMENTS)
            for (int g = 0; g < 7; ++g) {
                u64* l = s.racc + 9 * N + 3 * N * g;
                const double v = em_sum(l[j], l[N + j], l[2 * N + j], q[g]);
                l[j] = 0;
                l[N + j] = 0;
                l[2 * N + j] = 0;
                if (keep) o[(5 + g) * N + j] = v;
            }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (keep) s.sweeps_out[slot] = s.ctl[kEMMaxSweeps];
        s.t[0] += 1;
        s.ctl[0] = 1;
        for (int k = 1; k <= kEMMaxSweeps; ++k) s.ctl[k] = 0;
        s.umax[2] = 0;
        s.umax[3] = 0;
        s.umax[4] = 0;
    }
}

// ---------------------------------------------------------------------------
// the explicit scheme (src/GaussianEMExplicit.jl:12-30)
// ---------------------------------------------------------------------------
// X .= x; d(X) into the exact limbs (:13); x .= mod.(X .+ v[:,1]*dt, 1)
// (:14); SORT: the particle's bucket and key for sortperm(x).
template <bool SORT>
__global__ void __launch_bounds__(kEMThreads) k_emx_begin(EM1D s) {
    extern __shared__ u64 ldsu[];
    const int N = s.N;
    for (int i = threadIdx.x; i < 3 * N; i += blockDim.x) ldsu[i] = 0;
    __syncthreads();
    const double Nd = (double)N;
    const int q0 = em_frac_bits(s.P, s.w, 1.0);
    const int64_t p0 = (int64_t)blockIdx.x * s.tile;
    const int64_t p1 = p0 + s.tile < s.P ? p0 + s.tile : s.P;
    for (int64_t p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
        const double X = s.x[p];
        s.X[p] = X;
        int32_t cell[13];
        double ff[13];
        em_stencil(X, N, Nd, cell, ff);
#pragma unroll
        for (int k = 0; k < 13; ++k) {
            const double a = ff[k] * s.w;
            if (a != 0.0) em_add(ldsu, N, cell[k], fix_round128(a, q0));
        }
        const double x = jl_mod1f(X + s.v1[p] * s.dt);
        s.x[p] = x;
        if (SORT) {
            const uint32_t b = em_bucket(x, s.nb);
            s.bucket[p] = b;
            s.vkey[p] = em_order_key(x);
            atomicAdd(s.count + b, 1u);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 3 * N; i += blockDim.x) {
        const u64 v = ldsu[i];
        if (v) atomicAdd(s.racc + i, v);
    }
}

// One block of any size with N <= JPT*blockDim.x.  PHASE 0 (:13, :15-16):
// Ex = Gauss(rho); By .-= dt*rif(-ik.*fft(Ez)); Bz .-= dt*rif(ik.*fft(Ey));
// by, bz = (B + C)/2 with C the old B.  PHASE 1 (:23-24): Ey .+= dt*rif(-ik.*
// fft(Bz)/eps^2 .- fft(Jy)); Ez .+= dt*rif(+ik.*fft(By)/eps^2 .- fft(Jz)).
// SRC 0: rho, Jy, Jz from the exact limbs g (rho at 0, Jy at 3N, Jz at 6N;
// cleared here) with fraction bits q0..q2; 1: from the arrays rin, jyin, jzin.
template <int PHASE, int SRC, int JPT>
__device__ void emx_solve_body(const EM1D& s, u64* g, const double* __restrict__ rin, const double* __restrict__ jyin,
                               const double* __restrict__ jzin, int q0, int q1, int q2, double* sm) {
    const int N = s.N, lN = s.lN, st = em_stride(N);
    const int t = threadIdx.x, nth = blockDim.x;
    double* re = sm;
    double* im = sm + 5 * st;
    const double Nd = (double)N, sc = 1.0 / Nd, dt = s.dt, e2 = s.eps * s.eps;
    constexpr int NL = PHASE == 0 ? 3 : 4;
    // PHASE 0 lines: rho, Ez, Ey; PHASE 1: Bz, Jy, By, Jz
    for (int j = t; j < N; j += blockDim.x) {
        double v[NL];
        if (PHASE == 0) {
            if (SRC == 0) {
                v[0] = em_sum(g[j], g[N + j], g[2 * N + j], q0);
                for (int l = 0; l < 3; ++l) g[l * N + j] = 0;
            } else {
                v[0] = rin[j];
            }
            v[1] = s.Ez[j];
            v[2] = s.Ey[j];
        } else {
            if (SRC == 0) {
                v[1] = em_sum(g[3 * N + j], g[4 * N + j], g[5 * N + j], q1);
                v[3] = em_sum(g[6 * N + j], g[7 * N + j], g[8 * N + j], q2);
                for (int l = 3; l < 9; ++l) g[l * N + j] = 0;
            } else {
                v[1] = jyin[j];
                v[3] = jzin[j];
            }
            v[0] = s.Bz[j];
            v[2] = s.By[j];
        }
        const int pos = padi((int)brev((uint32_t)j, lN));
#pragma unroll
        for (int l = 0; l < NL; ++l) {
            re[l * st + pos] = v[l];
            im[l * st + pos] = 0.0;
        }
    }
    __syncthreads();
    fft_lines(re, im, s.tw, lN, NL, st, false);
    constexpr int NO = PHASE == 0 ? 3 : 2;
    double o[JPT][2 * NO];
#pragma unroll
    for (int q = 0; q < JPT; ++q) {
        const int j = t + q * nth;
        if (j >= N) break;
        const int pj = padi(j);
        double kr, ki, pr, pi;
        ik_of(N, j, kr, ki);
        if (PHASE == 0) {
            if (j == 0) {  // (0,0)/(0,0) is NaN; `*= false` keeps its (x86: negative) sign
                o[q][0] = -0.0;
                o[q][1] = -0.0;
            } else {
                const double tt = 1.0 / ki;
                o[q][0] = im[pj] * tt;
                o[q][1] = -(re[pj] * tt);
            }
            cmul(-kr, -ki, re[st + pj], im[st + pj], o[q][2], o[q][3]);
            cmul(kr, ki, re[2 * st + pj], im[2 * st + pj], o[q][4], o[q][5]);
        } else {
            cmul(-kr, -ki, re[pj], im[pj], pr, pi);
            o[q][0] = pr / e2 - re[st + pj];
            o[q][1] = pi / e2 - im[st + pj];
            cmul(kr, ki, re[2 * st + pj], im[2 * st + pj], pr, pi);
            o[q][2] = pr / e2 - re[3 * st + pj];
            o[q][3] = pi / e2 - im[3 * st + pj];
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < JPT; ++q) {
        const int j = t + q * nth;
        if (j >= N) break;
        const int pos = padi((int)brev((uint32_t)j, lN));
#pragma unroll
        for (int l = 0; l < NO; ++l) {
            re[l * st + pos] = o[q][2 * l];
            im[l * st + pos] = o[q][2 * l + 1];
        }
    }
    __syncthreads();
    fft_lines(re, im, s.tw, lN, NO, st, true);
    for (int j = t; j < N; j += blockDim.x) {
        const int pj = padi(j);
        if (PHASE == 0) {
            const double cy = s.By[j], cz = s.Bz[j];
            const double By = cy - dt * (re[st + pj] * sc);
            const double Bz = cz - dt * (re[2 * st + pj] * sc);
            s.Ex[j] = re[pj] * sc;
            s.By[j] = By;
            s.Bz[j] = Bz;
            s.by[j] = (By + cy) / 2.0;
            s.bz[j] = (Bz + cz) / 2.0;
        } else {
            s.Ey[j] = s.Ey[j] + dt * (re[pj] * sc);
            s.Ez[j] = s.Ez[j] + dt * (re[st + pj] * sc);
        }
    }
}

template <int PHASE, int SRC>
__global__ void __launch_bounds__(kEMSolveThreads)
k_emx_solve(EM1D s, const double* __restrict__ rin, const double* __restrict__ jyin,
            const double* __restrict__ jzin) {
    extern __shared__ double sm[];
    int q0 = 0, q1 = 0, q2 = 0;
    if (SRC == 0) {
        q0 = em_frac_bits(s.P, s.w, 1.0);
        q1 = em_frac_bits(s.P, s.w, bits_d(s.umax[3]));
        q2 = em_frac_bits(s.P, s.w, bits_d(s.umax[4]));
    }
    emx_solve_body<PHASE, SRC, kMaxNEM / kEMSolveThreads>(s, s.racc, rin, jyin, jzin, q0, q1, q2, sm);
    if (PHASE == 1 && SRC == 0 && threadIdx.x == 0) s.ctl[kEMMaxSweeps] = 1;
}


// v += a*dt/2; v += cross(v, b)*dt; v += a*dt/2 with a, b gathered at x
// (:17-22; fe: Ex, Ey, Ez, by, bz).
__device__ __forceinline__ void emx_velocity(const double* fe, int N, double Nd, double x, double dt, double& v1,
                                             double& v2, double& v3) {
    int32_t cell[13];
    double ff[13];
    em_stencil(x, N, Nd, cell, ff);
    double a1 = 0, a2 = 0, a3 = 0, b1 = 0, b2 = 0, b3 = 0;
#pragma unroll
    for (int k = 0; k < 13; ++k) {
        const int m = cell[k];
        const double t1 = fe[m] * ff[k], t2 = fe[N + m] * ff[k], t3 = fe[2 * N + m] * ff[k];
        const double s1 = 0.0 * ff[k], s2 = fe[3 * N + m] * ff[k], s3 = fe[4 * N + m] * ff[k];
        if (k == 0) {
            a1 = t1;
            a2 = t2;
            a3 = t3;
            b1 = s1;
            b2 = s2;
            b3 = s3;
        } else {
            a1 = a1 + t1;
            a2 = a2 + t2;
            a3 = a3 + t3;
            b1 = b1 + s1;
            b2 = b2 + s2;
            b3 = b3 + s3;
        }
    }
    const double u1 = v1 + a1 * dt / 2.0, u2 = v2 + a2 * dt / 2.0, u3 = v3 + a3 * dt / 2.0;
    const double c1 = u2 * b3 - u3 * b2, c2 = u3 * b1 - u1 * b3, c3 = u1 * b2 - u2 * b1;
    const double w1 = u1 + c1 * dt, w2 = u2 + c2 * dt, w3 = u3 + c3 * dt;
    v1 = w1 + a1 * dt / 2.0;
    v2 = w2 + a2 * dt / 2.0;
    v3 = w3 + a3 * dt / 2.0;
}

__device__ __forceinline__ void stage_fields_x(double* fe, const EM1D& s) {
    const int N = s.N;
    for (int j = threadIdx.x; j < N; j += blockDim.x) {
        fe[j] = s.Ex[j];
        fe[N + j] = s.Ey[j];
        fe[2 * N + j] = s.Ez[j];
        fe[3 * N + j] = s.by[j];
        fe[4 * N + j] = s.bz[j];
    }
    __syncthreads();
}

// the velocity update of every particle at its (sorted) x, and the maxima of
// |v1|, |v2|, |v3| for the currents' and the moments' fixed-point grids
__global__ void __launch_bounds__(kEMThreads) k_emx_push(EM1D s) {
    extern __shared__ double fe[];
    __shared__ u64 red[kEMThreads / 64];
    stage_fields_x(fe, s);
    const double Nd = (double)s.N;
    double m1 = 0.0, m2 = 0.0, m3 = 0.0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < s.P; p += stride) {
        double v1 = s.v1[p], v2 = s.v2[p], v3 = s.v3[p];
        emx_velocity(fe, s.N, Nd, s.x[p], s.dt, v1, v2, v3);
        s.v1[p] = v1;
        s.v2[p] = v2;
        s.v3[p] = v3;
        m1 = fmax_num(m1, fabs(v1));
        m2 = fmax_num(m2, fabs(v2));
        m3 = fmax_num(m3, fabs(v3));
    }
    block_max_into(m1, s.umax + 2, red);
    block_max_into(m2, s.umax + 3, red);
    block_max_into(m3, s.umax + 4, red);
}

// d((X.+x)/2, v[:,2]) and d((X.+x)/2, v[:,3]) (:23-24) into limbs 3..8
__global__ void __launch_bounds__(kEMThreads) k_emx_current(EM1D s) {
    extern __shared__ u64 ldsu[];
    const int N = s.N;
    for (int i = threadIdx.x; i < 6 * N; i += blockDim.x) ldsu[i] = 0;
    __syncthreads();
    const double Nd = (double)N;
    const int q1 = em_frac_bits(s.P, s.w, bits_d(s.umax[3]));
    const int q2 = em_frac_bits(s.P, s.w, bits_d(s.umax[4]));
    const int64_t p0 = (int64_t)blockIdx.x * s.tile;
    const int64_t p1 = p0 + s.tile < s.P ? p0 + s.tile : s.P;
    for (int64_t p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
        const double c = (s.X[p] + s.x[p]) / 2.0;
        const double uy = s.v2[p], uz = s.v3[p];
        int32_t cell[13];
        double ff[13];
        em_stencil(c, N, Nd, cell, ff);
#pragma unroll
        for (int k = 0; k < 13; ++k) {
            const double a1 = uy * ff[k] * s.w, a2 = uz * ff[k] * s.w;
            if (a1 != 0.0) em_add(ldsu, N, cell[k], fix_round128(a1, q1));
            if (a2 != 0.0) em_add(ldsu + 3 * N, N, cell[k], fix_round128(a2, q2));
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 6 * N; i += blockDim.x) {
        const u64 v = ldsu[i];
        if (v) atomicAdd(s.racc + 3 * N + i, v);
    }
}

// ---------------------------------------------------------------------------
// operators
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kEMThreads) k_em_umax(const double* __restrict__ u, int64_t n, u64* dst) {
    __shared__ u64 red[kEMThreads / 64];
    double m = 0.0;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x)
        m = fmax_num(m, fabs(u[p]));
    block_max_into(m, dst, red);
}

__global__ void __launch_bounds__(kEMThreads)
k_em_deposit_op(int N, const double* __restrict__ y, const double* __restrict__ u, int64_t n, double w,
                const u64* __restrict__ umax, u64* __restrict__ racc, int tile) {
    extern __shared__ u64 ldsu[];
    for (int i = threadIdx.x; i < 3 * N; i += blockDim.x) ldsu[i] = 0;
    __syncthreads();
    const int q = em_frac_bits(n, w, u ? bits_d(*umax) : 1.0);
    const double Nd = (double)N;
    const int64_t p0 = (int64_t)blockIdx.x * tile;
    const int64_t p1 = p0 + tile < n ? p0 + tile : n;
    for (int64_t p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
        int32_t cell[13];
        double ff[13];
        em_stencil(y[p], N, Nd, cell, ff);
#pragma unroll
        for (int k = 0; k < 13; ++k) {
            const double a = (u ? u[p] * ff[k] : ff[k]) * w;
            if (a != 0.0) em_add(ldsu, N, cell[k], fix_round128(a, q));
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 3 * N; i += blockDim.x) {
        const u64 v = ldsu[i];
        if (v) atomicAdd(racc + i, v);
    }
}

__global__ void __launch_bounds__(256) k_em_deposit_out(int N, int64_t n, double w, const double* u,
                                                        const u64* __restrict__ umax, u64* __restrict__ racc,
                                                        double* __restrict__ rho) {
    const int q = em_frac_bits(n, w, u ? bits_d(*umax) : 1.0);
    for (int j = threadIdx.x; j < N; j += blockDim.x) {
        rho[j] = em_sum(racc[j], racc[N + j], racc[2 * N + j], q);
        racc[j] = 0;
        racc[N + j] = 0;
        racc[2 * N + j] = 0;
    }
}

__global__ void __launch_bounds__(kEMThreads)
k_em_push_op(EM1D s, const double* __restrict__ c, const double* __restrict__ u, int64_t n, double* __restrict__ v) {
    extern __shared__ double fe[];
    stage_fields(fe, s);
    const double Nd = (double)s.N;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
        double a, b, d;
        em_velocity(fe, s.N, Nd, c[p], s.dt, u[p], u[n + p], u[2 * n + p], a, b, d);
        v[p] = a;
        v[n + p] = b;
        v[2 * n + p] = d;
    }
}

__global__ void __launch_bounds__(kEMThreads)
k_emx_push_op(EM1D s, const double* __restrict__ xs, const double* __restrict__ u, int64_t n,
              double* __restrict__ v) {
    extern __shared__ double fe[];
    stage_fields_x(fe, s);
    const double Nd = (double)s.N;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
        double a = u[p], b = u[n + p], d = u[2 * n + p];
        emx_velocity(fe, s.N, Nd, xs[p], s.dt, a, b, d);
        v[p] = a;
        v[n + p] = b;
        v[2 * n + p] = d;
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static unsigned em_blocks(int64_t n) {
    const int64_t b = (n + kEMThreads - 1) / kEMThreads;
    return (unsigned)(b < 4096 ? (b > 0 ? b : 1) : 4096);
}

size_t em_solve_lds_bytes(int N) { return sizeof(double) * (10 * (size_t)(N + (N >> 4)) + kRedDoubles); }

void em_launch_sweep(const EM1D& s, int sweep, hipStream_t st) {
    const unsigned nb = em_blocks(s.P);
    const unsigned nt = (unsigned)((s.P + s.tile - 1) / s.tile);
    if (sweep == 0) k_em_move<true><<<nb, kEMThreads, 0, st>>>(s, sweep);
    else k_em_move<false><<<nb, kEMThreads, 0, st>>>(s, sweep);
    k_em_deposit3<<<nt, kEMDepThreads, sizeof(u64) * 9 * (size_t)s.N, st>>>(s, sweep);
    k_em_solve<0, true><<<1, kEMSolveThreads, em_solve_lds_bytes(s.N), st>>>(s, sweep, nullptr, nullptr, nullptr);
    k_em_push<<<nb, kEMThreads, sizeof(double) * 5 * (size_t)s.N, st>>>(s, sweep);
}

void em_launch_end(const EM1D& s, bool sort, bool moments, hipStream_t st) {
    const unsigned nb = em_blocks(s.P);
    if (sort && s.P <= kEMSmallSort) {
        k_em_sort_small<true, true><<<1, 1024, 0, st>>>(s);
    } else if (sort) {
        const int ntiles = (s.nb + kEMScanTile - 1) / kEMScanTile;
        k_em_sort_zero<<<(unsigned)((s.nb + 255) / 256 < 1024 ? (s.nb + 255) / 256 : 1024), 256, 0, st>>>(s);
        k_em_wrap<true><<<nb, kEMThreads, 0, st>>>(s);
        k_em_scan_tiles<<<(unsigned)ntiles, 1024, 0, st>>>(s);
        k_em_scan_top<<<1, 256, 0, st>>>(s, ntiles);
        k_em_scan_add<<<(unsigned)((s.nb + 255) / 256 < 1024 ? (s.nb + 255) / 256 : 1024), 256, 0, st>>>(s);
        k_em_sort_scatter<<<nb, 256, 0, st>>>(s);
        k_em_sort_rank<<<nb, 256, 0, st>>>(s);
        k_em_sort_apply<<<nb, 256, 0, st>>>(s);
        k_em_sort_back<true><<<nb, kEMThreads, 0, st>>>(s);
    } else {
        k_em_wrap<false><<<nb, kEMThreads, 0, st>>>(s);
    }
    if (moments) {
        const unsigned nt = (unsigned)((s.P + s.tile - 1) / s.tile);
        k_em_moments<<<dim3(nt, kEMMomentGroups), kEMThreads, sizeof(u64) * 6 * (size_t)s.N, st>>>(s);
        k_em_finish<true><<<1, 256, 0, st>>>(s);
    } else {
        k_em_finish<false><<<1, 256, 0, st>>>(s);
    }
}

void emx_launch_step(const EM1D& s, bool sort, bool moments, hipStream_t st) {
    const unsigned nb = em_blocks(s.P);
    const unsigned nt = (unsigned)((s.P + s.tile - 1) / s.tile);
    const size_t solve = em_solve_lds_bytes(s.N);
    if (sort && s.P <= kEMSmallSort) {
        k_emx_begin<false><<<nt, kEMThreads, sizeof(u64) * 3 * (size_t)s.N, st>>>(s);
        k_emx_solve<0, 0><<<1, kEMSolveThreads, solve, st>>>(s, nullptr, nullptr, nullptr);
        k_em_sort_small<false, false><<<1, 1024, 0, st>>>(s);
    } else if (sort) {
        const int ntiles = (s.nb + kEMScanTile - 1) / kEMScanTile;
        const unsigned nbk = (unsigned)((s.nb + 255) / 256 < 1024 ? (s.nb + 255) / 256 : 1024);
        k_em_sort_zero<<<nbk, 256, 0, st>>>(s);
        k_emx_begin<true><<<nt, kEMThreads, sizeof(u64) * 3 * (size_t)s.N, st>>>(s);
        k_emx_solve<0, 0><<<1, kEMSolveThreads, solve, st>>>(s, nullptr, nullptr, nullptr);
        k_em_scan_tiles<<<(unsigned)ntiles, 1024, 0, st>>>(s);
        k_em_scan_top<<<1, 256, 0, st>>>(s, ntiles);
        k_em_scan_add<<<nbk, 256, 0, st>>>(s);
        k_em_sort_scatter<<<nb, 256, 0, st>>>(s);
        k_em_sort_rank<<<nb, 256, 0, st>>>(s);
        k_em_sort_apply<<<nb, 256, 0, st>>>(s);
        k_em_sort_back<false><<<nb, kEMThreads, 0, st>>>(s);
    } else {
        k_emx_begin<false><<<nt, kEMThreads, sizeof(u64) * 3 * (size_t)s.N, st>>>(s);
        k_emx_solve<0, 0><<<1, kEMSolveThreads, solve, st>>>(s, nullptr, nullptr, nullptr);
    }
    k_emx_push<<<nb, kEMThreads, sizeof(double) * 5 * (size_t)s.N, st>>>(s);
    k_emx_current<<<nt, kEMThreads, sizeof(u64) * 6 * (size_t)s.N, st>>>(s);
    k_emx_solve<1, 0><<<1, kEMSolveThreads, solve, st>>>(s, nullptr, nullptr, nullptr);
    if (moments) {
        k_em_moments<<<dim3(nt, kEMMomentGroups), kEMThreads, sizeof(u64) * 6 * (size_t)s.N, st>>>(s);
        k_em_finish<true><<<1, 256, 0, st>>>(s);
    } else {
        k_em_finish<false><<<1, 256, 0, st>>>(s);
    }
}

void emx_launch_fields_op(const EM1D& s, const double* rho, const double* Jy, const double* Jz, hipStream_t st) {
    const size_t solve = em_solve_lds_bytes(s.N);
    k_emx_solve<0, 1><<<1, kEMSolveThreads, solve, st>>>(s, rho, Jy, Jz);
    k_emx_solve<1, 1><<<1, kEMSolveThreads, solve, st>>>(s, rho, Jy, Jz);
}

void emx_launch_push_op(const EM1D& s, const double* x, const double* u, int64_t n, double* v, hipStream_t st) {
    if (n > 0) k_emx_push_op<<<em_blocks(n), kEMThreads, sizeof(double) * 5 * (size_t)s.N, st>>>(s, x, u, n, v);
}

void em_launch_deposit_op(const EM1D& s, const double* y, const double* u, int64_t n, double w, double* rho,
                          unsigned long long* scratch_max, hipStream_t st) {
    if (u) k_em_umax<<<em_blocks(n), kEMThreads, 0, st>>>(u, n, scratch_max);
    const int tile = 1024;
    const unsigned nt = (unsigned)((n + tile - 1) / tile);
    if (n > 0)
        k_em_deposit_op<<<nt, kEMThreads, sizeof(u64) * 3 * (size_t)s.N, st>>>(s.N, y, u, n, w, scratch_max, s.racc,
                                                                               tile);
    k_em_deposit_out<<<1, 256, 0, st>>>(s.N, n, w, u, scratch_max, s.racc, rho);
}

void em_launch_fields_op(const EM1D& s, const double* rho, const double* Jy, const double* Jz, hipStream_t st) {
    k_em_solve<1, false><<<1, kEMSolveThreads, em_solve_lds_bytes(s.N), st>>>(s, 0, rho, Jy, Jz);
}

void em_launch_push_op(const EM1D& s, const double* c, const double* u, int64_t n, double* v, hipStream_t st) {
    if (n > 0) k_em_push_op<<<em_blocks(n), kEMThreads, sizeof(double) * 5 * (size_t)s.N, st>>>(s, c, u, n, v);
}

int configure_kernels_em1d() {
    const int bytes = (int)em_solve_lds_bytes(kMaxNEM);
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(k_em_solve<0, true>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(k_em_solve<1, false>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(k_emx_solve<0, 0>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(k_emx_solve<1, 0>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(k_emx_solve<0, 1>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(k_emx_solve<1, 1>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    const int dep = (int)(sizeof(u64) * 9 * kMaxNEM);
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(k_em_deposit3), hipFuncAttributeMaxDynamicSharedMemorySize,
                              dep);
    (void)hipGetLastError();
    return 0;
}

}  // namespace pic
