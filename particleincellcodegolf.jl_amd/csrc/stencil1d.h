// stencil1d.h -- device helpers of the 1D FP64 kernels: the scripts' shape
// stencils (Gaussian and area weights) for gather and exact deposit, the LDS
// and device-scope charge accumulators, and the loop-state predicates of the
// step graph.  Compiled with -ffp-contract=off by every includer.
#pragma once
#include "pic_device.h"
#include "pic_internal.h"

#pragma clang fp contract(off)

namespace pic {
// ---------------------------------------------------------------------------
// Stencil helpers (R = 6 or 7)
// ---------------------------------------------------------------------------

// S = sum(k -> E[k[1]]*k[2], d(c)) as a left fold over the 2R+1 cells
// (src/GaussianFixedPointQuiet.jl:10); E is staged in LDS.
template <int R>
__device__ __forceinline__ double gather_gauss(const double* __restrict__ Eld, double c,
                                               double Nd, int32_t nmask) {
    const int64_t i0 = (int64_t)rint(c * Nd);
    double prev = edge_half_erf(i0 - R, c, Nd);
    double S = 0.0;
#pragma unroll 1
    for (int j = -R; j <= R; ++j) {
        double next = edge_half_erf(i0 + j + 1, c, Nd);
        double wt = next - prev;
        prev = next;
        double term = Eld[cell0(i0 + j, nmask)] * wt;
        S = (j == -R) ? term : S + term;
    }
    return S;
}

// r[k[1]] += k[2]*w for the 2R+1 cells of d(c) (src/GaussianFixedPointQuiet.jl:7),
// each product rounded to the 2^-q integer grid and added into LDS limbs.
template <int R>
__device__ __forceinline__ void deposit_gauss(u64* __restrict__ acc, int32_t N, double c,
                                              double Nd, int32_t nmask, double w, int q) {
    const int64_t i0 = (int64_t)rint(c * Nd);
    double prev = edge_half_erf(i0 - R, c, Nd);
#pragma unroll 1
    for (int j = -R; j <= R; ++j) {
        double next = edge_half_erf(i0 + j + 1, c, Nd);
        double wt = next - prev;
        prev = next;
        double a = wt * w;
        if (a != 0.0) lds_add_fxp(acc, N, cell0(i0 + j, nmask), fxp_from(a, q));
    }
}

__device__ __forceinline__ void zero_lds(u64* p, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

// Adds this block's LDS limbs into the instance's per-cell accumulator: the
// block total of a cell is normalised to 128 bits and re-split into 43-bit
// limbs, so device-scope u64 atomics from every block never carry.
__device__ __forceinline__ void flush_charge64(u128 v, u64* __restrict__ a) {
    const u64 l0 = (u64)v & kLimbMask, l1 = (u64)(v >> 43) & kLimbMask, l2 = (u64)(v >> 86);
    if (l0) atomicAdd(a, l0);
    if (l1) atomicAdd(a + 1, l1);
    if (l2) atomicAdd(a + 2, l2);
}

__device__ __forceinline__ void flush_limbs(const u64* __restrict__ acc, int32_t N, u64* __restrict__ racc) {
    for (int c = threadIdx.x; c < N; c += blockDim.x)
        flush_charge64(limbs_to_u128(acc[c], acc[N + c], acc[2 * N + c]), racc + 3 * c);
}

__device__ __forceinline__ bool sweep_over(const InstState& S, int k) {
    return S.last_sweep >= 0 && S.last_sweep < k;
}

// *halt: 0 = running, 1 = a solve in the graph's last sweep slot left an
// instance unconverged (that slot's push still runs), 2 = the graph is halted.
__device__ __forceinline__ bool halted(const int32_t* __restrict__ halt, bool last_slot) {
    const int32_t h = *halt;
    return last_slot ? h >= 2 : h >= 1;
}


// d(y) = (i = Int(mod1(ceil(y*N), N)); o = ceil(y*N) - y*N;
//         ((i, 1-o), (mod1(i-1, N), o)))  (src/AreaFixedPointQuietSimpson13.jl:5)
__device__ __forceinline__ void area_weights(double c, double Nd, int32_t nmask, int32_t& c0, double& w0,
                                             int32_t& c1, double& w1) {
    const double yN = c * Nd;
    const double cy = ceil(yN);
    const int64_t i = (int64_t)cy;
    const double o = cy - yN;
    c0 = cell0(i, nmask);
    w0 = 1.0 - o;
    c1 = cell0(i - 1, nmask);
    w1 = o;
}

template <int SH>
__device__ __forceinline__ double gather_shape(const double* __restrict__ Eld, double c, double Nd,
                                               int32_t nmask) {
    if constexpr (SH == 1) {
        int32_t c0, c1;
        double w0, w1;
        area_weights(c, Nd, nmask, c0, w0, c1, w1);
        const double s = Eld[c0] * w0;
        return s + Eld[c1] * w1;
    } else {
        return gather_gauss<SH>(Eld, c, Nd, nmask);
    }
}

template <int SH>
__device__ __forceinline__ void deposit_shape(u64* __restrict__ acc, int32_t N, double c, double Nd,
                                              int32_t nmask, double w, int q) {
    if constexpr (SH == 1) {
        int32_t c0, c1;
        double w0, w1;
        area_weights(c, Nd, nmask, c0, w0, c1, w1);
        const double a0 = w0 * w, a1 = w1 * w;
        if (a0 != 0.0) lds_add_fxp(acc, N, c0, fxp_from(a0, q));
        if (a1 != 0.0) lds_add_fxp(acc, N, c1, fxp_from(a1, q));
    } else {
        deposit_gauss<SH>(acc, N, c, Nd, nmask, w, q);
    }
}

// Reads and clears an instance's three-limb charge, into re/im in bit-reversed order.
__device__ __forceinline__ void take_charge64(u64* __restrict__ a, double* re, double* im, int N, int log2N, int q) {
    for (int j = threadIdx.x; j < N; j += blockDim.x) {
        u64* p = a + 3 * j;
        const u64 l0 = p[0], l1 = p[1], l2 = p[2];
        p[0] = 0;
        p[1] = 0;
        p[2] = 0;
        const int b = padi((int)brev((uint32_t)j, log2N));
        re[b] = fxp_to(limbs_to_u128(l0, l1, l2), q);
        im[b] = 0.0;
    }
}


}  // namespace pic
