// kernels_mixed_aux.hip -- the mixed-precision Simpson-1/3 fixed point and
// magnetised 1D2V kernels (the particle state and charge quanta of the
// midpoint's, mixed1d.h).  A unit of its own so that its compiler options
// can differ from the midpoint push's.
#include "pic1d2v_internal.h"
#include "pic_device.h"
#include "pic_internal.h"
#include "stencil1d.h"
#include "weights32.h"
#include "mixed1d.h"

namespace pic {

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

}  // namespace pic
