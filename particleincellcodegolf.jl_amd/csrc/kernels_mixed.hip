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
#include "mixed1d.h"

namespace pic {

// LDS of the push: the charge accumulator (32-bit lo and hi words per padded
// cell) and FP32 E over the padded cells, then 32
// doubles for the block's diagnostics partials
int push32_lds_bytes(int N) { return 12 * pad_cells(N) + 8 * 4 + 512; }

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
