// kernels_mixed.hip -- mixed-precision kernels of the implicit-midpoint step
// (src/GaussianFixedPoint.jl:7-9, src/GaussianFixedPointQuiet.jl:8-10).
//
// Particle state: position as a 32-bit fixed-point fraction of the periodic
// box (2^-32 resolution, wrap-around for free, cell index = top log2 N bits,
// sub-cell offset = the rest), velocity as an FP32 hi + lo pair (the lo word
// is read and written once per step to keep the time integration compensated),
// Picard iterate in FP32.  Velocities are held in steps of the fixed-point
// position per time step (v * dt * 2^32), so the half and quarter steps of
// the midpoint rule are exact power-of-two scalings and no rounded dt factor
// biases every particle of a run the same way; the field arrives scaled the
// same way (E * dt^2 * 4, once per cell in FP64, by the solve).  Shape
// weights are FP32 polynomials in the sub-cell offset (weights32.h); the
// gather sums in FP32; the charge is accumulated exactly as integer multiples
// of 2^-30 of a particle's charge, so the per-cell total is exact and
// independent of scheduling.  The FFT, the convergence
// test and the diagnostics stay FP64 (kernels_1d.hip).
#include "pic1d2v_internal.h"
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
// LDS of the push: the charge accumulator (32-bit lo and hi words per padded
// cell) and FP32 E over the padded cells, then 32
// doubles for the block's diagnostics partials
int push32_lds_bytes(int N) { return 12 * pad_cells(N) + 8 * 4 + 512; }

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

template <int MODE>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(8)))
k_mid_push32(const InstDesc* __restrict__ insts, InstState* __restrict__ st,
             const TileDesc* __restrict__ tiles, uint32_t* Xq0, uint32_t* Xq1, float* Vh0,
             float* Vh1, float* Vl0, float* Vl1, float* vf0, float* vf1,
             const float* __restrict__ E32, u64* __restrict__ racc, int k, const int32_t* __restrict__ halt,
             int last_slot, double* __restrict__ dpart, int prime) {
    extern __shared__ u64 lds[];
    // the loads the early exits need go out before one test of them all:
    // halt and the tile together, then the instance and its state (each
    // scalar-cache round trip costs a launch part of a microsecond; I.N,
    // never 0, keeps the instance's load ahead of the test)
    const int32_t h = *halt;  // see RunBuffers::halt
    const TileDesc T = tiles[blockIdx.x];
    const InstDesc I = insts[T.inst];
    InstState S = st[T.inst];
    // (| rather than ||: every operand is evaluated, so no load waits
    // behind a branch of its own)
    const bool over = MODE == 1 && S.last_sweep >= 0 && S.last_sweep < k;
    if ((h >= ((MODE == 1 && last_slot) ? 2 : 1)) | (I.N <= 0) | over) return;
    if (MODE == 0) {
        if (T.local == 0 && threadIdx.x == 0) st[T.inst].last_sweep = -1;
        S.last_sweep = -1;
    }
    const bool last = (MODE == 1) && (S.last_sweep == k);
    const int32_t N = I.N, nmask = I.nmask;
    const int NPAD = pad_cells(N);
    u64* acc = lds;  // the lo words of the NPAD slots, then their hi words
    uint32_t* acc_lo = reinterpret_cast<uint32_t*>(acc);
    uint32_t* acc_hi = acc_lo + NPAD;
    float* Eld = reinterpret_cast<float*>(lds + NPAD);
    if (MODE == 1)
        for (int i = threadIdx.x; i < NPAD; i += blockDim.x)
            Eld[i] = E32[I.goff + ((i - kHalo) & nmask)];  // the kick per 2^-30 of weight
    // prime: the final sweep also deposits the next step's first charge
    // (centre x + V*dt/2 with v_0 = V), which a first-deposit launch would
    // otherwise compute from the state it writes
    const bool deposit = !last || prime;
    // a middle sweep (k >= 1, not the last) finds the charge of the centres
    // it gathers at still in the accumulator (the solve keeps it) and adds,
    // for each particle whose next centre differs, the exact quanta of its
    // new stencil minus those of its old one; once the Picard iterate settles
    // most centres repeat bit for bit and add nothing.  The result is the
    // charge a full deposit gives, bit for bit.
    const bool delta = (MODE == 1) && k >= 1 && !last;
    if (deposit)
        for (int i = threadIdx.x; i < NPAD; i += blockDim.x) acc[i] = 0;
    __syncthreads();

    const int par = S.step & 1;
    const uint32_t* Xq = par ? Xq1 : Xq0;
    const float* Vh = par ? Vh1 : Vh0;
    const float* Vl = par ? Vl1 : Vl0;
    uint32_t* Xo = par ? Xq0 : Xq1;
    float* Vho = par ? Vh0 : Vh1;
    float* Vlo = par ? Vl0 : Vl1;
    const float* vin = (k & 1) ? vf1 : vf0;
    float* vout = ((k + 1) & 1) ? vf1 : vf0;
    const int shift = 32 - I.log2N;
    constexpr float cscale = 0.25f;  // centre (x + X)/2 = X + (v + V)/4*dt
    constexpr float xscale = 0.5f;   // x = X + (v + V)/2*dt
    const int64_t base = I.poff + T.start;
    // final sweep: exact sums of v^2 (2^-kV2Bits grid) and v (2^-kV1Bits grid)
    // over this thread's particles
    i128 q2 = 0, q1 = 0;
    // every lane of a wave runs the same iterations (the deposit reduces
    // across the wave); lanes past the tile's end carry no particle.  The
    // loop is compiled once per kind of launch, chosen once per block:
    // 0 the first deposit, 1 sweep 0 (full deposit of the next centres),
    // 2 a middle sweep (charge changes only), 3 the final sweep.
    const int64_t nit = (T.count + blockDim.x - 1) / blockDim.x;
    auto sweep = [&](auto kind) {
        constexpr int K = decltype(kind)::value;
        for (int64_t it = 0; it < nit; ++it) {
            const int64_t p = threadIdx.x + it * blockDim.x;
            const bool valid = p < T.count;
            const int64_t g = base + (valid ? p : 0);
            Stencil32 s;
            s.i0 = 0;
            bool dep = false;
            if (valid) {
                const uint32_t xq = Xq[g];
                const float vh = Vh[g];
                uint32_t cq = 0;
                uint32_t cdep;  // centre of this particle's deposit
                dep = true;
                if constexpr (K == 0) {
                    cdep = advance(xq, vh + vh, cscale);
                } else {
                    const float vk = (K == 1) ? vh : (K == 2) ? vin[g] : ((k == 0) ? vh : vin[g]);
                    cq = advance(xq, vk + vh, cscale);
                    stencil32(cq, shift, s);
                    const float* e = Eld + s.i0;  // slot of array cell i0-6
                    float kick = 0.0f;
#pragma unroll
                    for (int j = 0; j < kWin; ++j) kick = fmaf(e[j], s.w[j], kick);
                    if constexpr (K == 3) {
                        // V_new = (V_hi + V_lo) + kick, renormalised (two-sum)
                        const float sum = vh + kick;
                        const float bb = sum - vh;
                        const float err = (vh - (sum - bb)) + (kick - bb);
                        const float t = Vl[g] + err;
                        const float hi = sum + t;
                        const float lo = t - (hi - sum);
                        const uint32_t xn = advance(xq, vk + vh, xscale);
                        Xo[g] = xn;
                        Vho[g] = hi;
                        Vlo[g] = lo;
                        const double vd = ((double)hi + (double)lo) * I.inv_vstep;  // the sum is exact
                        q2 += fix_round128(vd * vd, kV2Bits);
                        q1 += fix_round128(vd, kV1Bits);
                        dep = prime != 0;
                        cdep = advance(xn, hi + hi, cscale);  // next step's first centre, v_0 = V
                    } else {
                        const float vn = vh + kick;
                        vout[g] = vn;
                        cdep = advance(xq, vn + vh, cscale);
                        if constexpr (K == 2) {
                            if (cdep != cq) {
                                lane_change<-1>(acc, s);
                                stencil32(cdep, shift, s);
                                lane_change<1>(acc, s);
                            }
                            dep = false;
                        }
                    }
                }
                // once the Picard iterate settles the next centre repeats bit for bit
                if (dep && (K == 0 || cdep != cq)) stencil32(cdep, shift, s);
            }
            if constexpr (K != 2) {
                if (K != 3 || prime) lane_deposit(acc_lo, acc_hi, s, dep);
            }
        }
    };
    if (MODE == 0) sweep(std::integral_constant<int, 0>{});
    else if (last) sweep(std::integral_constant<int, 3>{});
    else if (delta) sweep(std::integral_constant<int, 2>{});
    else sweep(std::integral_constant<int, 1>{});
    if (last) {
        store_tile_kinetic(q2, q1, lds + NPAD + (NPAD + 1) / 2, dpart);
        if (!prime) return;
    }
    __syncthreads();
    if (delta) flush_delta(acc, N, NPAD, racc + I.red0);
    else flush_charge32(acc_lo, acc_hi, N, NPAD, racc + I.red0);
}

void launch_mid_push32(const RunBuffers& b, int mode, int k, bool last_slot, hipStream_t s) {
    const int ls = last_slot ? 1 : 0;
    const dim3 g(b.ntiles), t(b.block_push);
    u64* part = (u64*)b.racc;
    if (mode == 0)
        k_mid_push32<0><<<g, t, b.lds_push, s>>>(b.insts, b.st, b.tiles, b.Xq[0], b.Xq[1], b.Vh[0], b.Vh[1],
                                                 b.Vl[0], b.Vl[1], b.vf[0], b.vf[1], b.E32, part, 0, b.halt, 0,
                                                 b.dpart, 0);
    else
        k_mid_push32<1><<<g, t, b.lds_push, s>>>(b.insts, b.st, b.tiles, b.Xq[0], b.Xq[1], b.Vh[0], b.Vh[1],
                                                 b.Vl[0], b.Vl[1], b.vf[0], b.vf[1], b.E32, part, k, b.halt, ls,
                                                 b.dpart, b.prime_next);
}

// ---------------------------------------------------------------------------
// Mixed Simpson-1/3 (src/GaussianFixedPointQuietSimpson13.jl:8-17,
// src/AreaFixedPointQuietSimpson13.jl:5-17) on the mixed particle state:
// fixed-point positions, velocities as hi + lo pairs in steps of the position
// grid, FP32 Picard iterate.  The three stage fields arrive as FP32 copies in
// the same units as the midpoint's (E * dt^2 * 4, so that a gather with
// weights in 2^-30 quanta is the kick in steps of the grid); charges are exact
// integer quanta; the solves and the 3 x N convergence test stay FP64.
// SH: 1 = the area weights of AreaFixedPointQuietSimpson13.jl:5 (exact
// quanta from the fixed-point centre), 6 / 7 = the Gaussian shapes (FP32
// weights over +-5 cells, as the midpoint).
// ---------------------------------------------------------------------------

// Area weights of a fixed-point centre c: i = ceil(c N) (Julia index, 0 for
// c = 0), o = i - c N; 1 - o goes to array cell i - 1 and o to cell i - 2
// (mod N), here in 2^-30 quanta (exact: o has 32 - log2 N fraction bits) and
// as padded LDS slots (slot s0 holds cell i - 1, s0 - 1 cell i - 2).
struct Area32 {
    int32_t s0;
    uint32_t q0, q1;
};
__device__ __forceinline__ Area32 area32(uint32_t cq, int shift) {
    const uint32_t i = (uint32_t)(((uint64_t)cq + ((1ull << shift) - 1)) >> shift);
    const uint32_t o = (uint32_t)(((uint64_t)i << shift) - cq);
    const uint32_t o30 = o << (30 - shift);
    return {(int32_t)i + kHalo - 1, (1u << 30) - o30, o30};
}

// One exact add into the block's LDS charge (lo word, hi word on a wrap).
__device__ __forceinline__ void add_quanta(uint32_t* __restrict__ lo, uint32_t* __restrict__ hi, int slot,
                                           uint32_t q) {
    const uint32_t old = atomicAdd(&lo[slot], q);
    if (old + q < old) atomicAdd(&hi[slot], 1u);
}

// The kick of a padded FP32 field at centre cq, in steps of the grid.
template <int SH>
__device__ __forceinline__ float simp_gather32(const float* __restrict__ Eld, uint32_t cq, int shift) {
    if constexpr (SH == 1) {
        const Area32 a = area32(cq, shift);
        return fmaf(Eld[a.s0 - 1], (float)a.q1, Eld[a.s0] * (float)a.q0);
    } else {
        Stencil32 s;
        stencil32(cq, shift, s);
        const float* e = Eld + s.i0;
        float kick = 0.0f;
#pragma unroll
        for (int j = 0; j < kWin; ++j) kick = fmaf(e[j], s.w[j], kick);
        return kick;
    }
}

template <int SH>
__device__ __forceinline__ void simp_deposit32(uint32_t* __restrict__ lo, uint32_t* __restrict__ hi, uint32_t cq,
                                               int shift) {
    if constexpr (SH == 1) {
        const Area32 a = area32(cq, shift);
        add_quanta(lo, hi, a.s0, a.q0);
        if (a.q1) add_quanta(lo, hi, a.s0 - 1, a.q1);
    } else {
        Stencil32 s;
        stencil32(cq, shift, s);
        lane_deposit(lo, hi, s, true);
    }
}

// rho(X, X) of the step (centre X) into racc; the step's Picard loop starts.
template <int SH>
__global__ void __launch_bounds__(1024)
k_simp_first32(const InstDesc* __restrict__ insts, InstState* __restrict__ st, const TileDesc* __restrict__ tiles,
               const uint32_t* Xq0, const uint32_t* Xq1, u64* __restrict__ racc, const int32_t* __restrict__ halt) {
    extern __shared__ u64 lds[];
    if (halted(halt, false)) return;
    const TileDesc T = tiles[blockIdx.x];
    const InstDesc I = insts[T.inst];
    const InstState S = st[T.inst];
    if (T.local == 0 && threadIdx.x == 0) st[T.inst].last_sweep = -1;
    const int NPAD = pad_cells(I.N);
    uint32_t* lo = reinterpret_cast<uint32_t*>(lds);
    uint32_t* hi = lo + NPAD;
    zero_lds(lds, NPAD);
    __syncthreads();
    const uint32_t* Xq = (S.step & 1) ? Xq1 : Xq0;
    const int shift = 32 - I.log2N;
    const int64_t base = I.poff + T.start;
    for (int64_t p = threadIdx.x; p < T.count; p += blockDim.x) simp_deposit32<SH>(lo, hi, Xq[base + p], shift);
    __syncthreads();
    flush_charge32(lo, hi, I.N, NPAD, racc + I.red0);
}

// Per particle: the first stage's kick g1/6 (the script's Vg - V, in steps of
// the grid) and, with v_0 = V, the charges rho(X, x_0) and rho(x_0, x_0).
template <int SH>
__global__ void __launch_bounds__(1024)
k_simp_start32(const InstDesc* __restrict__ insts, const InstState* __restrict__ st,
               const TileDesc* __restrict__ tiles, const uint32_t* Xq0, const uint32_t* Xq1, const float* Vh0,
               const float* Vh1, float* __restrict__ G1s, const float* __restrict__ E32s1, u64* __restrict__ racc2,
               u64* __restrict__ racc3, const int32_t* __restrict__ halt) {
    extern __shared__ u64 lds[];
    if (halted(halt, false)) return;
    const TileDesc T = tiles[blockIdx.x];
    const InstDesc I = insts[T.inst];
    const InstState S = st[T.inst];
    const int NPAD = pad_cells(I.N);
    uint32_t* lo2 = reinterpret_cast<uint32_t*>(lds);
    uint32_t* hi2 = lo2 + NPAD;
    uint32_t* lo3 = hi2 + NPAD;
    uint32_t* hi3 = lo3 + NPAD;
    float* E1ld = reinterpret_cast<float*>(lds + 2 * NPAD);
    for (int i = threadIdx.x; i < NPAD; i += blockDim.x) E1ld[i] = E32s1[I.goff + ((i - kHalo) & I.nmask)];
    zero_lds(lds, 2 * NPAD);
    __syncthreads();
    const int par = S.step & 1;
    const uint32_t* Xq = par ? Xq1 : Xq0;
    const float* Vh = par ? Vh1 : Vh0;
    const int shift = 32 - I.log2N;
    const int64_t base = I.poff + T.start;
    for (int64_t p = threadIdx.x; p < T.count; p += blockDim.x) {
        const int64_t g = base + p;
        const uint32_t xq = Xq[g];
        const float vh = Vh[g];
        G1s[g] = simp_gather32<SH>(E1ld, xq, shift) / 6.0f;
        simp_deposit32<SH>(lo2, hi2, advance(xq, vh + vh, 0.25f), shift);  // (X + x_0)/2
        simp_deposit32<SH>(lo3, hi3, advance(xq, vh + vh, 0.5f), shift);   // x_0
    }
    __syncthreads();
    flush_charge32(lo2, hi2, I.N, NPAD, racc2 + I.red0);
    flush_charge32(lo3, hi3, I.N, NPAD, racc3 + I.red0);
}

// Sweep k: v_{k+1} = V + g1/6 + (4 g2)/6 + g3/6 with g2 at (X + x_k)/2 and g3 at
// x_k; the final sweep adds the change to the hi + lo pair, writes x_k and
// sums the exact kinetic terms; any other sweep writes the iterate and
// deposits rho(X, x_{k+1}) and rho(x_{k+1}, x_{k+1}).
template <int SH>
__global__ void __launch_bounds__(1024)
k_simp_push32(const InstDesc* __restrict__ insts, const InstState* __restrict__ st,
              const TileDesc* __restrict__ tiles, uint32_t* Xq0, uint32_t* Xq1, float* Vh0, float* Vh1, float* Vl0,
              float* Vl1, float* vf0, float* vf1, const float* __restrict__ G1s, const float* __restrict__ E32s2,
              const float* __restrict__ E32s3, u64* __restrict__ racc2, u64* __restrict__ racc3, int k,
              const int32_t* __restrict__ halt, int last_slot, double* __restrict__ dpart) {
    extern __shared__ u64 lds[];
    if (halted(halt, last_slot)) return;
    const TileDesc T = tiles[blockIdx.x];
    const InstDesc I = insts[T.inst];
    const InstState S = st[T.inst];
    if (sweep_over(S, k)) return;
    const bool last = S.last_sweep == k;
    const int NPAD = pad_cells(I.N);
    uint32_t* lo2 = reinterpret_cast<uint32_t*>(lds);
    uint32_t* hi2 = lo2 + NPAD;
    uint32_t* lo3 = hi2 + NPAD;
    uint32_t* hi3 = lo3 + NPAD;
    float* E2ld = reinterpret_cast<float*>(lds + 2 * NPAD);
    float* E3ld = E2ld + NPAD;
    u64* red = lds + 3 * NPAD;  // 4 words per wave
    for (int i = threadIdx.x; i < NPAD; i += blockDim.x) {
        const int c = I.goff + ((i - kHalo) & I.nmask);
        E2ld[i] = E32s2[c];
        E3ld[i] = E32s3[c];
    }
    if (!last) zero_lds(lds, 2 * NPAD);
    __syncthreads();
    const int par = S.step & 1;
    const uint32_t* Xq = par ? Xq1 : Xq0;
    const float* Vh = par ? Vh1 : Vh0;
    const float* Vl = par ? Vl1 : Vl0;
    uint32_t* Xo = par ? Xq0 : Xq1;
    float* Vho = par ? Vh0 : Vh1;
    float* Vlo = par ? Vl0 : Vl1;
    const float* vin = (k & 1) ? vf1 : vf0;
    float* vout = ((k + 1) & 1) ? vf1 : vf0;
    const int shift = 32 - I.log2N;
    const int64_t base = I.poff + T.start;
    i128 q2 = 0, q1 = 0;
    for (int64_t p = threadIdx.x; p < T.count; p += blockDim.x) {
        const int64_t g = base + p;
        const uint32_t xq = Xq[g];
        const float vh = Vh[g];
        const float vk = (k == 0) ? vh : vin[g];
        const uint32_t xk = advance(xq, vk + vh, 0.5f);
        const float g2 = simp_gather32<SH>(E2ld, advance(xq, vk + vh, 0.25f), shift);
        const float g3 = simp_gather32<SH>(E3ld, xk, shift);
        const float d = (G1s[g] + (g2 * 4.0f) / 6.0f) + g3 / 6.0f;
        if (last) {
            // V_new = (V_hi + V_lo) + d, renormalised (two-sum)
            const float sum = vh + d;
            const float bb = sum - vh;
            const float err = (vh - (sum - bb)) + (d - bb);
            const float t = Vl[g] + err;
            const float h = sum + t;
            const float l = t - (h - sum);
            Xo[g] = xk;
            Vho[g] = h;
            Vlo[g] = l;
            const double vd = ((double)h + (double)l) * I.inv_vstep;
            q2 += fix_round128(vd * vd, kV2Bits);
            q1 += fix_round128(vd, kV1Bits);
            continue;
        }
        const float vn = vh + d;
        vout[g] = vn;
        simp_deposit32<SH>(lo2, hi2, advance(xq, vn + vh, 0.25f), shift);  // (X + x_{k+1})/2
        simp_deposit32<SH>(lo3, hi3, advance(xq, vn + vh, 0.5f), shift);   // x_{k+1}
    }
    if (last) {
        store_tile_kinetic(q2, q1, red, dpart);
        return;
    }
    __syncthreads();
    flush_charge32(lo2, hi2, I.N, NPAD, racc2 + I.red0);
    flush_charge32(lo3, hi3, I.N, NPAD, racc3 + I.red0);
}

// LDS: two padded charges (8 B per slot each), two padded FP32 fields, and
// the final sweep's per-wave partials.
int simpson32_push_lds_bytes(int N) { return 24 * pad_cells(N) + 32 * 16; }

void launch_simp32_first(const RunBuffers& b, hipStream_t s) {
    const dim3 g(b.ntiles), t(b.block_push);
    const size_t lds = 8 * (size_t)pad_cells(b.maxN);
    u64* r = (u64*)b.racc;
    if (b.R == 1) k_simp_first32<1><<<g, t, lds, s>>>(b.insts, b.st, b.tiles, b.Xq[0], b.Xq[1], r, b.halt);
    else k_simp_first32<7><<<g, t, lds, s>>>(b.insts, b.st, b.tiles, b.Xq[0], b.Xq[1], r, b.halt);
}

void launch_simp32_start(const RunBuffers& b, hipStream_t s) {
    const dim3 g(b.ntiles), t(b.block_push);
    const size_t lds = 20 * (size_t)pad_cells(b.maxN);
    u64 *r2 = (u64*)b.racc2, *r3 = (u64*)b.racc3;
    if (b.R == 1)
        k_simp_start32<1><<<g, t, lds, s>>>(b.insts, b.st, b.tiles, b.Xq[0], b.Xq[1], b.Vh[0], b.Vh[1], b.G1s,
                                            b.E32s1, r2, r3, b.halt);
    else
        k_simp_start32<7><<<g, t, lds, s>>>(b.insts, b.st, b.tiles, b.Xq[0], b.Xq[1], b.Vh[0], b.Vh[1], b.G1s,
                                            b.E32s1, r2, r3, b.halt);
}

void launch_simp32_push(const RunBuffers& b, int k, bool last_slot, hipStream_t s) {
    const dim3 g(b.ntiles), t(b.block_push);
    const size_t lds = (size_t)simpson32_push_lds_bytes(b.maxN);
    const int ls = last_slot ? 1 : 0;
    u64 *r2 = (u64*)b.racc2, *r3 = (u64*)b.racc3;
    if (b.R == 1)
        k_simp_push32<1><<<g, t, lds, s>>>(b.insts, b.st, b.tiles, b.Xq[0], b.Xq[1], b.Vh[0], b.Vh[1], b.Vl[0],
                                           b.Vl[1], b.vf[0], b.vf[1], b.G1s, b.E32s2, b.E32, r2, r3, k, b.halt, ls,
                                           b.dpart);
    else
        k_simp_push32<7><<<g, t, lds, s>>>(b.insts, b.st, b.tiles, b.Xq[0], b.Xq[1], b.Vh[0], b.Vh[1], b.Vl[0],
                                           b.Vl[1], b.vf[0], b.vf[1], b.G1s, b.E32s2, b.E32, r2, r3, k, b.halt, ls,
                                           b.dpart);
}

// ---------------------------------------------------------------------------
// Mixed magnetised 1D2V (src/NGP1D2V.jl, src/NGP1D2V2S.jl): the Gaussian
// deposit and gather over +-5 cells with the FP32 weights, the Boris update
// (src/NGP1D2V2S.jl:5-11, B along z) computed in FP32 as velocity changes and
// added to hi + lo pairs, the move by the rounded velocity, exact charge
// quanta per species.  The solve, the binned field and the diagnostics
// stay FP64 (kernels_1d2v.hip).
// ---------------------------------------------------------------------------

// v += d on a hi + lo pair (two-sum, error folded into lo, renormalised).
__device__ __forceinline__ void add_pair(float& hi, float& lo, float d) {
    const float s = hi + d;
    const float bb = s - hi;
    const float err = (hi - (s - bb)) + (d - bb);
    const float t = lo + err;
    hi = s + t;
    lo = t - (hi - s);
}

// rint(hi + lo): the integer part of hi first, so the rounding sees both words.
__device__ __forceinline__ int32_t rint_pair(float hi, float lo) {
    const float hi_i = rintf(hi);
    return (int32_t)hi_i + (int32_t)rintf((hi - hi_i) + lo);
}

// The charge of every tile's particles at their current positions.
__global__ void __launch_bounds__(1024)
k12_deposit32(const Tile12* __restrict__ tiles, const uint32_t* __restrict__ xq, u64* __restrict__ racc32, int N,
              int log2N) {
    extern __shared__ u64 lds[];
    const Tile12 T = tiles[blockIdx.x];
    const int NPAD = pad_cells(N);
    uint32_t* lo = reinterpret_cast<uint32_t*>(lds);
    uint32_t* hi = lo + NPAD;
    zero_lds(lds, NPAD);
    __syncthreads();
    const int shift = 32 - log2N;
    for (int64_t p = threadIdx.x; p < T.count; p += blockDim.x) {
        Stencil32 s;
        stencil32(xq[T.start + p], shift, s);
        lane_deposit(lo, hi, s, true);
    }
    __syncthreads();
    flush_charge32(lo, hi, N, NPAD, racc32 + (int64_t)T.species * N);
}

// Ej at x, the Boris update, x += vx*dt (mod 1 by the fixed point), and the
// charge at the new x for the next step.
__global__ void __launch_bounds__(1024)
k12_push32(const Tile12* __restrict__ tiles, uint32_t* __restrict__ xq, float* __restrict__ vxh,
           float* __restrict__ vxl, float* __restrict__ vyh, float* __restrict__ vyl, const float* __restrict__ E32,
           u64* __restrict__ racc32, int N, int log2N, float kh0, float kh1, float t30, float t31) {
    extern __shared__ u64 lds[];
    const Tile12 T = tiles[blockIdx.x];
    const int NPAD = pad_cells(N);
    const int32_t nmask = N - 1;
    uint32_t* lo = reinterpret_cast<uint32_t*>(lds);
    uint32_t* hi = lo + NPAD;
    float* Eld = reinterpret_cast<float*>(lds + NPAD);
    for (int i = threadIdx.x; i < NPAD; i += blockDim.x) Eld[i] = E32[(i - kHalo) & nmask];
    zero_lds(lds, NPAD);
    __syncthreads();
    const float kh = T.species ? kh1 : kh0, t3 = T.species ? t31 : t30;
    const float s3 = 1.0f + t3 * t3;
    const int shift = 32 - log2N;
    for (int64_t p = threadIdx.x; p < T.count; p += blockDim.x) {
        const int64_t g = T.start + p;
        uint32_t x = xq[g];
        float ux = vxh[g], lx = vxl[g], uy = vyh[g], ly = vyl[g];
        Stencil32 s;
        stencil32(x, shift, s);
        const float* e = Eld + s.i0;
        float gsum = 0.0f;
#pragma unroll
        for (int j = 0; j < kWin; ++j) gsum = fmaf(e[j], s.w[j], gsum);
        const float ek = gsum * kh;
        // m = v + E h; c = m x t; a = m + c; d = a x t; v' = m + 2d/s + E h
        const float m1 = ux + ek, m2 = uy;
        const float a1 = m1 + m2 * t3, a2 = m2 - m1 * t3;
        add_pair(ux, lx, (ek + ek) + (2.0f * (a2 * t3)) / s3);
        add_pair(uy, ly, (2.0f * (-(a1 * t3))) / s3);
        x += (uint32_t)rint_pair(ux, lx);
        xq[g] = x;
        vxh[g] = ux;
        vxl[g] = lx;
        vyh[g] = uy;
        vyl[g] = ly;
        stencil32(x, shift, s);
        lane_deposit(lo, hi, s, true);
    }
    __syncthreads();
    flush_charge32(lo, hi, N, NPAD, racc32 + (int64_t)T.species * N);
}

// Exact diagnostics partials when t % bin == 0, per chunk of 4096 particle
// indices and per species: sums of vx^2 + vy^2 on a 2^-50 grid and of vx, vy
// on a 2^-40 grid (velocities in natural units), so the totals do not depend
// on the order of the particles or the chunks.
__global__ void __launch_bounds__(256)
k12_diag_chunks32(const float* __restrict__ vxh, const float* __restrict__ vxl, const float* __restrict__ vyh,
                  const float* __restrict__ vyl, int64_t P, int ns, double inv_vstep, int bin,
                  const State12* __restrict__ st, i128* __restrict__ part) {
    __shared__ i128 red[3 * 256];
    if ((st->step + 1) % bin != 0) return;
    const int64_t c0 = (int64_t)blockIdx.x * kTreeChunk;
    const int64_t n = P - c0 < kTreeChunk ? P - c0 : kTreeChunk;
    for (int s = 0; s < ns; ++s) {
        i128 a = 0, bx = 0, by = 0;
        for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
            const int64_t j = (int64_t)s * P + c0 + i;
            const double vx = ((double)vxh[j] + (double)vxl[j]) * inv_vstep;
            const double vy = ((double)vyh[j] + (double)vyl[j]) * inv_vstep;
            a += fix_round128(vx * vx + vy * vy, kV2Bits);
            bx += fix_round128(vx, kV1Bits);
            by += fix_round128(vy, kV1Bits);
        }
        red[threadIdx.x] = a;
        red[256 + threadIdx.x] = bx;
        red[512 + threadIdx.x] = by;
        __syncthreads();
        for (int h = 128; h > 0; h >>= 1) {
            if ((int)threadIdx.x < h) {
                red[threadIdx.x] += red[threadIdx.x + h];
                red[256 + threadIdx.x] += red[256 + threadIdx.x + h];
                red[512 + threadIdx.x] += red[512 + threadIdx.x + h];
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            i128* o = part + 6 * (int64_t)blockIdx.x + 3 * s;
            o[0] = red[0];
            o[1] = red[256];
            o[2] = red[512];
        }
        __syncthreads();
    }
}

// D[t/bin, 1:5] (src/NGP1D2V2S.jl:52-56) from the exact partials when
// t % bin == 0; then step += 1.
__global__ void __launch_bounds__(256)
k12_diag_final32(const i128* __restrict__ part, int64_t nchunks, int64_t P, int ns, double m0, double m1, int N,
                 double n0, int bin, State12* __restrict__ st, double* __restrict__ D, int64_t cap_bins) {
    __shared__ i128 red[6 * 256];
    const int64_t t = st->step + 1;
    if (t % bin == 0) {
        i128 v[6] = {0, 0, 0, 0, 0, 0};
        for (int64_t c = threadIdx.x; c < nchunks; c += blockDim.x)
            for (int q = 0; q < 3 * ns; ++q) v[q] += part[6 * c + q];
        for (int q = 0; q < 6; ++q) red[q * 256 + threadIdx.x] = v[q];
        __syncthreads();
        for (int h = 128; h > 0; h >>= 1) {
            if ((int)threadIdx.x < h)
                for (int q = 0; q < 6; ++q) red[q * 256 + threadIdx.x] += red[q * 256 + threadIdx.x + h];
            __syncthreads();
        }
        const int64_t bi = t / bin - 1;
        if (threadIdx.x == 0 && bi < cap_bins) {
            const double k2 = 0x1p-50, k1 = 0x1p-40, Pd = (double)P;
            const double ms[2] = {m0, m1};
            double sk = 0.0, sx = 0.0, sy = 0.0;
            for (int s = 0; s < ns; ++s) {
                sk = sk + ms[s] * ((double)red[(3 * s) * 256] * k2);
                sx = sx + (ms[s] * ((double)red[(3 * s + 1) * 256] * k1)) / Pd;
                sy = sy + (ms[s] * ((double)red[(3 * s + 2) * 256] * k1)) / Pd;
            }
            double* row = D + 5 * bi;
            row[0] = row[0] + (st->se / (double)N) / 2.0;
            row[1] = row[1] + ((sk * n0) / Pd) / 2.0;
            row[3] = row[3] + sx;
            row[4] = row[4] + sy;
            row[2] = row[2] + (row[0] + row[1]);
            for (int c = 0; c < 3; ++c) row[c] = row[c] * (2.0 / n0);
        }
    }
    if (threadIdx.x == 0) st->step = t;
}

void launch12_deposit32(const Buffers12& b, hipStream_t s) {
    const size_t lds = 8 * (size_t)pad_cells(b.N);
    k12_deposit32<<<b.ntiles, b.block, lds, s>>>(b.tiles, b.xq, b.racc32, b.N, b.log2N);
}

void launch12_push32(const Buffers12& b, hipStream_t s) {
    const size_t lds = 12 * (size_t)pad_cells(b.N);
    k12_push32<<<b.ntiles, b.block, lds, s>>>(b.tiles, b.xq, b.vxh, b.vxl, b.vyh, b.vyl, b.E32, b.racc32, b.N,
                                              b.log2N, b.kick_half[0], b.kick_half[1], b.t3[0], b.t3[1]);
}

void launch12_diag32(const Buffers12& b, hipStream_t s) {
    i128* part = reinterpret_cast<i128*>(b.dpart32);
    k12_diag_chunks32<<<(unsigned)b.nchunks, 256, 0, s>>>(b.vxh, b.vxl, b.vyh, b.vyl, b.P, b.ns, b.inv_vstep,
                                                         b.bin, b.st, part);
    k12_diag_final32<<<1, 256, 0, s>>>(part, b.nchunks, b.P, b.ns, b.mass[0], b.mass[1], b.N, b.n0, b.bin, b.st,
                                       b.D, b.cap_bins);
}

int configure_kernels_mixed() {
    const int bytes = push32_lds_bytes(kMaxN);
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(k_mid_push32<0>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(k_mid_push32<1>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    (void)hipGetLastError();
    return 0;
}

}  // namespace pic
