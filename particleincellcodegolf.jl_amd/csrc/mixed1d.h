// mixed1d.h -- device helpers shared by the mixed-precision 1D kernels
// (kernels_mixed.hip: the midpoint push; kernels_mixed_aux.hip: the
// Simpson-1/3 and magnetised 1D2V forms): the +-5 stencil of a 32-bit
// fixed-point centre, its exact charge quanta and the LDS accumulators.
#pragma once
#include "pic_device.h"
#include "pic_internal.h"
#include "stencil1d.h"
#include "weights32.h"

namespace pic {

// The +-5 window of a centre in cell i0 covers array cells i0-6 .. i0+4 (Julia
// index i <-> array i-1, i0 in 0..N).  LDS copies of E and of the charge
// accumulator carry a halo: slot k holds cell (k - kHalo) mod N, so a window
// is kWin consecutive slots starting at i0 with no wrap-around arithmetic.
constexpr int kHalo = kStencil32 + 1;
constexpr int kWin = 2 * kStencil32 + 1;
__host__ __device__ constexpr int pad_cells(int N) { return N + 2 * kHalo + 4; }

// Stencil of a fixed-point centre: i0 = nearest cell centre (Julia index,
// N == 0) and the kWin weights, in 2^-30 units of a particle's charge, of the
// cells i0-6 .. i0+4 (array index = Julia index - 1).
struct Stencil32 {
    int32_t i0;
    float w[kWin];
};

__device__ __forceinline__ void stencil32(uint32_t cq, int shift, Stencil32& s) {
    const uint64_t t = (uint64_t)cq + (1ull << (shift - 1));
    const uint32_t i0 = (uint32_t)(t >> shift);
    const int32_t dint = (int32_t)(cq - (i0 << shift));
    const float d = ldexpf((float)dint, -shift);  // centre - cell centre, in cells
    const float u = d * d;
    s.i0 = (int32_t)i0;
    s.w[kStencil32] = qA0(u);
#define PIC_PAIR(J)                                  \
    {                                                \
        const float a = qA##J(u), b = d * qB##J(u);  \
        s.w[kStencil32 + J] = a + b;                 \
        s.w[kStencil32 - J] = a - b;                 \
    }
    PIC_PAIR(1)
    PIC_PAIR(2)
    PIC_PAIR(3)
    PIC_PAIR(4)
    PIC_PAIR(5)
#undef PIC_PAIR
}

// x + (v_a + v_b) * scale * dt on the fixed-point grid, velocities in steps
// of the grid: scale 1/2 gives the position, 1/4 the midpoint centre (exact).
__device__ __forceinline__ uint32_t advance(uint32_t xq, float v_sum, float scale) {
    return xq + (uint32_t)(int32_t)rintf(v_sum * scale);
}

// Deposit quanta of window slot J: the weight in units of 2^-30 of a
// particle's charge, rounded half up; the |j| >= 4 tails are clamped at zero
// (fit noise can dip below zero there; nearer cells are never below 1e4
// quanta).
template <int J>
__device__ __forceinline__ uint32_t quanta(float w30) {
    constexpr int d = J - kStencil32;
    const float r = w30 + 0.5f;
    return __float2uint_rz((d >= 4 || d <= -4) ? fmaxf(r, 0.0f) : r);
}

// Deposit of one particle's stencil: 11 exact integer adds into the block's
// LDS accumulator, a slot's count being hi * 2^32 + lo.  Each add goes to the
// 32-bit lo word and returns its old value; an add that wraps lo (old + q <
// old, q < 2^32) adds one to hi.  A 32-bit LDS atomic costs fewer cycles than
// a 64-bit one on random slots, and the rare carry adds run with few lanes.
// A slot takes at most 2^30 quanta per particle of the tile, so hi counts at
// most a quarter of the tile's particles.  Called by every lane of a wave (dep marks lanes with a
// stencil to deposit).
__device__ __forceinline__ void lane_deposit(uint32_t* __restrict__ lo, uint32_t* __restrict__ hi,
                                             const Stencil32& s, bool dep) {
    if (!dep) return;
    uint32_t q[kWin], old[kWin];
#define PIC_DEP(J)                       \
    q[J] = quanta<J>(s.w[J]);            \
    old[J] = atomicAdd(&lo[s.i0 + J], q[J]);
    PIC_DEP(0) PIC_DEP(1) PIC_DEP(2) PIC_DEP(3) PIC_DEP(4) PIC_DEP(5) PIC_DEP(6) PIC_DEP(7) PIC_DEP(8) PIC_DEP(9)
    PIC_DEP(10)
#undef PIC_DEP
#pragma unroll
    for (int j = 0; j < kWin; ++j)
        if (old[j] + q[j] < old[j]) atomicAdd(&hi[s.i0 + j], 1u);
}

// One particle's stencil quanta added into (sign +1) or taken out of (-1)
// the block's u64 LDS accumulator of charge changes (two's complement, so the
// sums stay exact modulo 2^64 and the instance's totals, which are never
// negative, come out exact).
template <int SIGN>
__device__ __forceinline__ void lane_change(u64* __restrict__ acc, const Stencil32& s) {
    u64* a = acc + s.i0;
#define PIC_CH(J) atomicAdd(&a[J], (u64)((int64_t)SIGN * (int64_t)quanta<J>(s.w[J])));
    PIC_CH(0) PIC_CH(1) PIC_CH(2) PIC_CH(3) PIC_CH(4) PIC_CH(5) PIC_CH(6) PIC_CH(7) PIC_CH(8) PIC_CH(9)
    PIC_CH(10)
#undef PIC_CH
}

// Adds a block's u64 LDS charge changes into the instance's accumulator
// (modulo 2^64; cell c collects every slot k == c + kHalo (mod N)).
__device__ __forceinline__ void flush_delta(const u64* acc, int N, int NPAD, u64* __restrict__ racc) {
    for (int c = threadIdx.x; c < N; c += blockDim.x) {
        u64 v = 0;
        for (int kk = (c + kHalo) % N; kk < NPAD; kk += N) v += acc[kk];
        if (v) atomicAdd(racc + c, v);
    }
}

// Adds this block's halo-padded LDS charge into the instance's per-cell
// accumulator (device-scope integer atomics; cell c collects every slot
// k == c + kHalo (mod N)).
__device__ __forceinline__ void flush_charge32(const uint32_t* lo, const uint32_t* hi, int N, int NPAD,
                                               u64* __restrict__ racc) {
    for (int c = threadIdx.x; c < N; c += blockDim.x) {
        u64 v = 0;
        for (int kk = (c + kHalo) % N; kk < NPAD; kk += N) v += ((u64)hi[kk] << 32) + lo[kk];
        if (v) atomicAdd(racc + c, v);
    }
}

// The step's kinetic diagnostics of the final sweep: the block's exact sums
// of v^2 and v (128-bit each) into the tile's partial pair of dpart (red:
// 4 words of LDS per wave).  Called by every thread of the block.
__device__ __forceinline__ void store_tile_kinetic(i128 q2, i128 q1, u64* red, double* __restrict__ dpart) {
    q2 = wave_sum_i128(q2);
    q1 = wave_sum_i128(q1);
    __syncthreads();
    const int wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[4 * wv] = (u64)q2;
        red[4 * wv + 1] = (u64)((u128)q2 >> 64);
        red[4 * wv + 2] = (u64)q1;
        red[4 * wv + 3] = (u64)((u128)q1 >> 64);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        i128 a = 0, b = 0;
        for (int w = 0; w < nw; ++w) {
            a += (i128)(((u128)red[4 * w + 1] << 64) | red[4 * w]);
            b += (i128)(((u128)red[4 * w + 3] << 64) | red[4 * w + 2]);
        }
        u64* o = reinterpret_cast<u64*>(dpart) + 4 * (int64_t)blockIdx.x;
        o[0] = (u64)a;
        o[1] = (u64)((u128)a >> 64);
        o[2] = (u64)b;
        o[3] = (u64)((u128)b >> 64);
    }
}

}  // namespace pic
