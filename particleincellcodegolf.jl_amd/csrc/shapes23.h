// shapes23.h -- device helpers of the structured 2D3V module (src/PIC2D3V.jl)
// shared by its electrostatic (kernels_2d3v.hip) and Lorenz-gauge
// (kernels_2d3v_lg.hip) kernels: the shapes, unimod, the Boris forms, the
// signed exact limbs and the push windows.
#pragma once
#include "pic_device.h"
#include "pic2d3v_internal.h"

#pragma clang fp contract(off)

namespace pic {

// ---------------------------------------------------------------------------
// shapes: depositindicesfractions + gridinteractiontuple (:1106-1187)
// SH = 0 NGP, 1 area, 2 + n B-spline of order n; SH_COUNT cells per axis.
// ---------------------------------------------------------------------------
template <int SH>
struct ShapeCount {
    static constexpr int value = SH == 0 ? 1 : (SH == 1 ? 2 : SH - 1);
};

__device__ __forceinline__ double q2(double t) { return t * t; }
__device__ __forceinline__ double q3(double t) { return t * (t * t); }
__device__ __forceinline__ double q4(double t) {
    const double s = t * t;
    return s * s;
}
__device__ __forceinline__ double q5(double t) {
    const double s = t * t;
    return t * (s * s);
}

// bspline(BSplineWeighting{n}(), x) (:1124-1162), written out left to right
// with integer powers as LLVM's powi expands them.
template <int n>
__device__ __forceinline__ void bspline_w(double x, double* w) {
    if constexpr (n == 0) {
        w[0] = 1.0;
    } else if constexpr (n == 1) {
        w[0] = x;
        w[1] = 1.0 - x;
    } else if constexpr (n == 2) {
        const double a = x - 1.5, b = x - 0.5, c = x + 0.5;
        w[0] = 9.0 / 8.0 + 3.0 / 2.0 * a + 1.0 / 2.0 * q2(a);
        w[1] = 3.0 / 4.0 - q2(b);
        w[2] = 9.0 / 8.0 - 3.0 / 2.0 * c + 1.0 / 2.0 * q2(c);
    } else if constexpr (n == 3) {
        const double a = x - 2.0, b = x - 1.0, c = x, d = x + 1.0;
        w[0] = 4.0 / 3.0 + 2.0 * a + q2(a) + 1.0 / 6.0 * q3(a);
        w[1] = 2.0 / 3.0 - q2(b) - 1.0 / 2.0 * q3(b);
        w[2] = 2.0 / 3.0 - q2(c) + 1.0 / 2.0 * q3(c);
        w[3] = 4.0 / 3.0 - 2.0 * d + q2(d) - 1.0 / 6.0 * q3(d);
    } else if constexpr (n == 4) {
        const double a = x - 2.5, b = x - 1.5, c = x - 0.5, d = x + 0.5, e = x + 1.5;
        w[0] = 625.0 / 384.0 + 125.0 / 48.0 * a + 25.0 / 16.0 * q2(a) + 5.0 / 12.0 * q3(a) + 1.0 / 24.0 * q4(a);
        w[1] = 55.0 / 96.0 - 5.0 / 24.0 * b - 5.0 / 4.0 * q2(b) - 5.0 / 6.0 * q3(b) - 1.0 / 6.0 * q4(b);
        w[2] = 115.0 / 192.0 - 5.0 / 8.0 * q2(c) + 1.0 / 4.0 * q4(c);
        w[3] = 55.0 / 96.0 + 5.0 / 24.0 * d - 5.0 / 4.0 * q2(d) + 5.0 / 6.0 * q3(d) - 1.0 / 6.0 * q4(d);
        w[4] = 625.0 / 384.0 - 125.0 / 48.0 * e + 25.0 / 16.0 * q2(e) - 5.0 / 12.0 * q3(e) + 1.0 / 24.0 * q4(e);
    } else {
        const double a = x - 3.0, b = x - 2.0, c = x - 1.0, d0 = x + 0.0, d1 = x - 0.0, e = x + 1.0, f = x + 2.0;
        w[0] = 243.0 / 120.0 + 81.0 / 24.0 * a + 9.0 / 4.0 * q2(a) + 3.0 / 4.0 * q3(a) + 1.0 / 8.0 * q4(a) +
               1.0 / 120.0 * q5(a);
        w[1] = 17.0 / 40.0 - 5.0 / 8.0 * b - 7.0 / 4.0 * q2(b) - 5.0 / 4.0 * q3(b) - 3.0 / 8.0 * q4(b) -
               1.0 / 24.0 * q5(b);
        w[2] = 22.0 / 40.0 - 1.0 / 2.0 * q2(c) + 1.0 / 4.0 * q4(c) + 1.0 / 12.0 * q5(c);
        w[3] = 22.0 / 40.0 - 1.0 / 2.0 * q2(d0) + 1.0 / 4.0 * q4(d1) - 1.0 / 12.0 * q5(d1);
        w[4] = 17.0 / 40.0 + 5.0 / 8.0 * e - 7.0 / 4.0 * q2(e) + 5.0 / 4.0 * q3(e) - 3.0 / 8.0 * q4(e) +
               1.0 / 24.0 * q5(e);
        w[5] = 243.0 / 120.0 - 81.0 / 24.0 * f + 9.0 / 4.0 * q2(f) - 3.0 / 4.0 * q3(f) + 1.0 / 8.0 * q4(f) -
               1.0 / 120.0 * q5(f);
    }
}

// Cells (0-based, wrapped: the ghost layers hold periodic images) and weights
// along one axis.  zN = z*NZ_Lz, i = ceil(zN), r = i - zN (:1108-1113).
template <int SH>
__device__ __forceinline__ void shape_w(double z, double NZ_Lz, int32_t N, int32_t* c, double* w) {
    const double zN = z * NZ_Lz;
    const double ci = ceil(zN);
    const int64_t i = (ci > -4.0e18 && ci < 4.0e18) ? (int64_t)ci : 0;
    const double r = (double)i - zN;
    const int64_t m = N - 1;
    if constexpr (SH == 0) {
        c[0] = (int32_t)((i - 1) & m);
        w[0] = 1.0;
    } else if constexpr (SH == 1) {
        c[0] = (int32_t)((i - 1) & m);
        c[1] = (int32_t)(i & m);
        w[0] = 1.0 - r;
        w[1] = r;
    } else {
        constexpr int n = SH - 2;
        int64_t j;
        double x;
        if constexpr (n & 1) {
            j = i;
            x = 1.0 - r;
        } else {
            const int q = r > 0.5;
            j = i + q;
            x = ((double)q + 0.5) - r;
        }
        bspline_w<n>(x, w);
        const int64_t lo = j - n / 2;
#pragma unroll
        for (int k = 0; k <= n; ++k) c[k] = (int32_t)((lo + k - 1) & m);
    }
}

// unimod(x, n) = x > n ? x - n : x > 0 ? x : x + n   (:9)
__device__ __forceinline__ double unimod23(double x, double n) { return x > n ? x - n : (x > 0.0 ? x : x + n); }

// Boris with E = (Ex, Ey, 0) and the uniform B.  BV 0: ElectrostaticBoris
// (:39-54); BV 1: ElectromagneticBoris with Ez = 0 (:56-69).
template <int BV>
__device__ __forceinline__ void boris23(const Grid23& g, double q_m, double Ex, double Ey, double& vx, double& vy,
                                        double& vz) {
    double t1, t2, t3, e1, e2, e3, sc = 0.0, den = 0.0, qm2 = 0.0;
    const double dt_2 = g.dt / 2.0;
    if constexpr (BV == 0) {
        t1 = g.B[0] * g.dt / 2.0;
        t2 = g.B[1] * g.dt / 2.0;
        t3 = g.B[2] * g.dt / 2.0;
        const double tt = t1 * t1 + t2 * t2 + t3 * t3;
        e1 = Ex * dt_2 * q_m;
        e2 = Ey * dt_2 * q_m;
        e3 = 0.0 * dt_2 * q_m;
        qm2 = q_m * q_m;
        den = 1.0 + qm2 * tt;
    } else {
        const double th = dt_2 * q_m;
        t1 = g.B[0] * th;
        t2 = g.B[1] * th;
        t3 = g.B[2] * th;
        sc = 2.0 / (1.0 + (t1 * t1 + t2 * t2 + t3 * t3));
        e1 = Ex * th;
        e2 = Ey * th;
        e3 = 0.0 * th;
    }
    const double m1 = vx + e1, m2 = vy + e2, m3 = vz + e3;
    const double c1 = m2 * t3 - m3 * t2, c2 = m3 * t1 - m1 * t3, c3 = m1 * t2 - m2 * t1;
    const double u1 = m1 + c1, u2 = m2 + c2, u3 = m3 + c3;
    const double d1 = u2 * t3 - u3 * t2, d2 = u3 * t1 - u1 * t3, d3 = u1 * t2 - u2 * t1;
    double p1, p2, p3;
    if constexpr (BV == 0) {
        p1 = m1 + d1 * qm2 * 2.0 / den;
        p2 = m2 + d2 * qm2 * 2.0 / den;
        p3 = m3 + d3 * qm2 * 2.0 / den;
    } else {
        p1 = m1 + d1 * sc;
        p2 = m2 + d2 * sc;
        p3 = m3 + d3 * sc;
    }
    vx = p1 + e1;
    vy = p2 + e2;
    vz = p3 + e3;
}

// Signed exact charge: v = rint(c*2^q) as three 38-bit two's-complement limbs,
// each added to its own u64 counter (no carries between counters).
__device__ __forceinline__ void add_limbs_s(u64* __restrict__ racc, int64_t G, int64_t k, i128 v) {
    const u128 u = (u128)v;
    constexpr u64 M = (1ull << kLimbS) - 1;
    const u64 l0 = (u64)u & M;
    const u64 l1 = (u64)(u >> kLimbS) & M;
    const u64 l2 = (u64)(u >> (2 * kLimbS)) & M;
    if (l0) atomicAdd(racc + k, l0);
    if (l1) atomicAdd(racc + G + k, l1);
    if (l2) atomicAdd(racc + 2 * G + k, l2);
}

template <int SH>
__device__ __forceinline__ void gather23(const Grid23& g, const double2* __restrict__ E, double x, double y, double& a,
                                         double& b) {
    constexpr int n = ShapeCount<SH>::value;
    int32_t cx[n], cy[n];
    double wx[n], wy[n];
    shape_w<SH>(x, g.NX_Lx, g.NX, cx, wx);
    shape_w<SH>(y, g.NY_Ly, g.NY, cy, wy);
    a = 0.0;
    b = 0.0;
#pragma unroll
    for (int jj = 0; jj < n; ++jj)
#pragma unroll
        for (int ii = 0; ii < n; ++ii) {
            const double wxy = wx[ii] * wy[jj];
            const double2 e = E[cx[ii] + (int64_t)g.NX * cy[jj]];
            a = fma(e.x, wxy, a);
            b = fma(e.y, wxy, b);
        }
}

// deposit!(rho, shape, x, y, NX_Lx, NY_Ly, w) (:1249-1256): j outer, i inner,
// contribution (wx*wy)*w.
template <int SH>
__device__ __forceinline__ void deposit23(const Grid23& g, u64* __restrict__ racc, double x, double y, double w) {
    constexpr int n = ShapeCount<SH>::value;
    const int64_t G = (int64_t)g.NX * g.NY;
    int32_t cx[n], cy[n];
    double wx[n], wy[n];
    shape_w<SH>(x, g.NX_Lx, g.NX, cx, wx);
    shape_w<SH>(y, g.NY_Ly, g.NY, cy, wy);
#pragma unroll
    for (int jj = 0; jj < n; ++jj)
#pragma unroll
        for (int ii = 0; ii < n; ++ii) {
            const double c = wx[ii] * wy[jj] * w;
            add_limbs_s(racc, G, cx[ii] + (int64_t)g.NX * cy[jj], fix_round128(c, g.q));
        }
}

// ---------------------------------------------------------------------------
// push windows: a block takes a chunk of kTile23Chunk consecutive particles
// and stages the cells about them in LDS (see the push kernels)
// ---------------------------------------------------------------------------
constexpr int kTile23Threads = 256;
constexpr int kTile23PPT = 4;
constexpr int kTile23Chunk = kTile23Threads * kTile23PPT;
// window cells: 16 B of field + 3 x 8 B of limbs each in the FP64 push
// (39.4 KB with its bookkeeping: four blocks share a CU's 160 KB)
constexpr int kTile23Cells = 1008;
constexpr int kTile23Halo = 1;

// stencil extent about the base cell i = ceil(zN): cells i - lo .. i + hi
template <int SH>
struct ShapeExtent {
    static constexpr int n = SH - 2;
    static constexpr int lo = SH <= 1 ? 0 : ((n & 1) ? (n - 1) / 2 : n / 2);
    static constexpr int hi = SH == 0 ? 0 : (SH == 1 ? 1 : ((n & 1) ? (n + 1) / 2 : n / 2 + 1));
};

__device__ __forceinline__ int32_t base_cell(double z, double NZ_Lz, int32_t N) {
    const double ci = ceil(z * NZ_Lz);
    const int64_t i = (ci > -4.0e18 && ci < 4.0e18) ? (int64_t)ci : 0;
    return (int32_t)((i - 1) & (int64_t)(N - 1));
}

__device__ __forceinline__ int wave_min_i(int v) {
#pragma unroll
    for (int h = 32; h > 0; h >>= 1) v = min(v, __shfl_xor(v, h, 64));
    return v;
}
__device__ __forceinline__ int wave_max_i(int v) {
#pragma unroll
    for (int h = 32; h > 0; h >>= 1) v = max(v, __shfl_xor(v, h, 64));
    return v;
}

// adds the wave's count of fallback particles to *fallback (every lane of
// the wave calls it)
__device__ __forceinline__ void count_fallback(u64* fallback, int fb) {
#pragma unroll
    for (int h = 32; h > 0; h >>= 1) fb += __shfl_xor(fb, h, 64);
    if ((threadIdx.x & 63) == 0 && fb && fallback) atomicAdd(fallback, (u64)fb);
}

__device__ __forceinline__ void add_limbs_lds(u64* Rw, int idx, i128 v) {
    const u128 u = (u128)v;
    constexpr u64 M = (1ull << kLimbS) - 1;
    const u64 l0 = (u64)u & M;
    const u64 l1 = (u64)(u >> kLimbS) & M;
    const u64 l2 = (u64)(u >> (2 * kLimbS)) & M;
    if (l0) atomicAdd(Rw + idx, l0);
    if (l1) atomicAdd(Rw + kTile23Cells + idx, l1);
    if (l2) atomicAdd(Rw + 2 * kTile23Cells + idx, l2);
}

}  // namespace pic
