// kernels_1d.hip -- FP64 gfx950 kernels of the 1D1V electrostatic PIC step
// (src/NGPFourier.jl, src/Gaussian.jl, src/GaussianFixedPoint.jl,
// src/GaussianFixedPointQuiet.jl).  Compiled with -ffp-contract=off: every
// operation rounds as the CPU restatement does, so results are bitwise equal.
//
// Data layout in HBM: structure-of-arrays, one FP64 array per particle
// quantity (X, V, Picard iterate v), instances packed back to back; grids
// (rho, E) packed the same way.  Per sweep one push kernel streams the
// particles once: it gathers the field at the sweep's midpoint, updates v,
// and deposits the charge of the next midpoint into an LDS grid of exact
// integer accumulators (no float atomics, so results never depend on
// scheduling).  Per-tile partials are summed per cell by the reduce kernel;
// one workgroup per instance runs the FFT pair and the convergence test in
// LDS.
#include "pic_device.h"
#include "pic_internal.h"
#include "solve1d.h"
#include "stencil1d.h"

#pragma clang fp contract(off)

namespace pic {

// ---------------------------------------------------------------------------
// Implicit-midpoint push (src/GaussianFixedPointQuiet.jl:8-10).
// MODE 0: the first deposit of a step, v_0 = V, so x_0 = X + ((V+V)/2)*dt.
// MODE 1: sweep k: gather E_k at c_k = (x_k+X)/2, v_{k+1} = V + S*dt; if
//         sweep k ends the loop write the step's result (mod(x_k,1), v_{k+1}),
//         else deposit the charge of sweep k+1 at c_{k+1}.
// ---------------------------------------------------------------------------
template <int R, int MODE>
__global__ void __launch_bounds__(1024)
k_mid_push64(const InstDesc* __restrict__ insts, InstState* __restrict__ st,
             const TileDesc* __restrict__ tiles, double* X0, double* X1, double* V0,
             double* V1, double* vit0, double* vit1, const double* __restrict__ E0,
             const double* __restrict__ E1, u64* __restrict__ racc, int k, const int32_t* __restrict__ halt,
             int last_slot, int prime) {
    extern __shared__ u64 lds[];
    if (halted(halt, MODE == 1 && last_slot)) return;
    const TileDesc T = tiles[blockIdx.x];
    const InstDesc I = insts[T.inst];
    InstState S = st[T.inst];
    if (MODE == 0) {
        if (T.local == 0 && threadIdx.x == 0) st[T.inst].last_sweep = -1;
        S.last_sweep = -1;
    }
    if (MODE == 1 && sweep_over(S, k)) return;
    const bool last = (MODE == 1) && (S.last_sweep == k);
    const int32_t N = I.N;
    double* Eld = reinterpret_cast<double*>(lds);
    u64* acc = lds + N;
    if (MODE == 1) {
        const double* Ein = (S.e_cur ? E1 : E0) + I.goff;
        for (int i = threadIdx.x; i < N; i += blockDim.x) Eld[i] = Ein[i];
    }
    // prime: the final sweep also deposits the next step's first charge (the
    // deposit of a MODE 0 launch on the state it writes)
    const bool deposit = !last || prime;
    if (deposit) zero_lds(acc, 3 * N);
    __syncthreads();

    const int par = S.step & 1;
    const double* X = par ? X1 : X0;
    const double* V = par ? V1 : V0;
    double* Xo = par ? X0 : X1;
    double* Vo = par ? V0 : V1;
    const double* vin = (k & 1) ? vit1 : vit0;
    double* vout = ((k + 1) & 1) ? vit1 : vit0;
    const double Nd = (double)N, dt = I.dt, w = I.w;
    const int32_t nmask = I.nmask;
    const int q = I.q;
    const int64_t base = I.poff + T.start;
    // the particle loop is compiled once for the final sweep and once for
    // the others, and the block picks one
    auto sweep = [&](auto final_sweep) {
        constexpr bool LAST = decltype(final_sweep)::value;
        for (int64_t p = threadIdx.x; p < T.count; p += blockDim.x) {
            const int64_t g = base + p;
            const double Xp = X[g], Vp = V[g];
            double vn = Vp;
            if (MODE == 1) {
                const double vk = (k == 0) ? Vp : vin[g];
                const double x = Xp + ((vk + Vp) / 2.0) * dt;
                const double c = (x + Xp) / 2.0;
                const double Sg = gather_gauss<R>(Eld, c, Nd, nmask);
                vn = Vp + Sg * dt;
                if constexpr (LAST) {
                    const double Xn = jl_mod1f(x);
                    Xo[g] = Xn;
                    Vo[g] = vn;
                    if (prime) {
                        // next step: X = Xn, V = vn, x_0 = X + ((V + V)/2)*dt
                        const double x0 = Xn + ((vn + vn) / 2.0) * dt;
                        deposit_gauss<R>(acc, N, (x0 + Xn) / 2.0, Nd, nmask, w, q);
                    }
                    continue;
                }
                vout[g] = vn;
            }
            const double x2 = Xp + ((vn + Vp) / 2.0) * dt;
            const double c2 = (x2 + Xp) / 2.0;
            deposit_gauss<R>(acc, N, c2, Nd, nmask, w, q);
        }
    };
    if (last) sweep(std::true_type{});
    else sweep(std::false_type{});
    if (!deposit) return;
    __syncthreads();
    flush_limbs(acc, N, racc + 3 * I.red0);
}

// ---------------------------------------------------------------------------
// Leapfrog (src/NGPFourier.jl:4-6, src/Gaussian.jl:8-10).
// first:  u(); deposit at x.        second: u(); v += gather(E at x)*dt.
// ---------------------------------------------------------------------------
template <int R>
__global__ void __launch_bounds__(1024)
k_lf_first64(const InstDesc* __restrict__ insts, const InstState* __restrict__ st,
             const TileDesc* __restrict__ tiles, double* X0, double* X1,
             const double* V0, const double* V1, u64* __restrict__ racc) {
    extern __shared__ u64 lds[];
    const TileDesc T = tiles[blockIdx.x];
    const InstDesc I = insts[T.inst];
    const InstState S = st[T.inst];
    const int32_t N = I.N;
    u64* acc = lds;
    zero_lds(acc, R == 0 ? N : 3 * N);
    __syncthreads();
    const int par = S.step & 1;
    const double* X = par ? X1 : X0;
    const double* V = par ? V1 : V0;
    double* Xo = par ? X0 : X1;
    const double Nd = (double)N, dt = I.dt;
    const int64_t base = I.poff + T.start;
    for (int64_t p = threadIdx.x; p < T.count; p += blockDim.x) {
        const int64_t g = base + p;
        const double x = jl_mod1f(X[g] + (V[g] / 2.0) * dt);
        Xo[g] = x;
        if (R == 0) {
            atomicAdd(&acc[ngp_cell(x, Nd, I.nmask)], 1ull);
        } else {
            deposit_gauss<R>(acc, N, (x + x) / 2.0, Nd, I.nmask, I.w, I.q);
        }
    }
    __syncthreads();
    u64* out = racc + 3 * I.red0;
    if (R == 0) {
        const u128 one = fxp_from(I.w, I.q);
        for (int c = threadIdx.x; c < N; c += blockDim.x)
            if (acc[c]) flush_charge64((u128)acc[c] * one, out + 3 * c);
    } else {
        flush_limbs(acc, N, out);
    }
}

template <int R>
__global__ void __launch_bounds__(1024)
k_lf_second64(const InstDesc* __restrict__ insts, const InstState* __restrict__ st,
              const TileDesc* __restrict__ tiles, double* X0, double* X1, const double* V0,
              const double* V1, double* Vo0, double* Vo1, const double* __restrict__ E0) {
    extern __shared__ u64 lds[];
    const TileDesc T = tiles[blockIdx.x];
    const InstDesc I = insts[T.inst];
    const InstState S = st[T.inst];
    const int32_t N = I.N;
    double* Eld = reinterpret_cast<double*>(lds);
    for (int i = threadIdx.x; i < N; i += blockDim.x) Eld[i] = E0[I.goff + i];
    __syncthreads();
    const int par = S.step & 1;
    const double* V = par ? V1 : V0;
    double* Xn = par ? X0 : X1;  // written by the first half of this step
    double* Vn = par ? Vo0 : Vo1;
    const double Nd = (double)N, dt = I.dt;
    const int64_t base = I.poff + T.start;
    for (int64_t p = threadIdx.x; p < T.count; p += blockDim.x) {
        const int64_t g = base + p;
        const double v = V[g];
        const double x = jl_mod1f(Xn[g] + (v / 2.0) * dt);
        Xn[g] = x;
        double Sg;
        if (R == 0) Sg = Eld[ngp_cell(x, Nd, I.nmask)];
        else Sg = gather_gauss<R>(Eld, x, Nd, I.nmask);
        Vn[g] = v + Sg * dt;
    }
}

template <bool TEST>
__global__ void __launch_bounds__(kSolveThreads)
k_solve(const InstDesc* __restrict__ insts, InstState* __restrict__ st,
        void* __restrict__ racc, int precision, double* E0, double* E1, float* E32,
        const double2* __restrict__ twp, int k, int32_t* __restrict__ halt, int stall) {
    extern __shared__ double sm[];
    // halt, the instance and its state are loaded together before one test
    // of every early exit (| evaluates each operand, so no load waits behind
    // a branch of its own; I.N is never 0)
    const int32_t h = *halt;
    const InstDesc I = insts[blockIdx.x];
    InstState S = st[blockIdx.x];
    if ((h >= 1) | (I.N <= 0) | (TEST && sweep_over(S, k))) return;
    const int N = I.N, log2N = I.log2N;
    const int NP = solve_np(N);
    double* re = sm;
    double* im = sm + NP;
    double* Fl = sm + 2 * NP;
    double* red = Fl + N;
    const int cur = TEST ? S.e_cur : 0;
    const int nxt = TEST ? 1 - cur : 0;
    const double* F = (cur ? E1 : E0) + I.goff;
    constexpr int kPer = kMaxN / kSolveThreads;
    // batched loads: all of this thread's cells are in flight together, ahead
    // of the transform's table loads (which are only needed later)
    u64 a0[kPer], a1[kPer], a2[kPer];
    double fv[kPer];
#pragma unroll
    for (int c = 0; c < kPer; ++c) {
        const int j = threadIdx.x + c * kSolveThreads;
        if (j < N) {
            if (precision == 0) {
                const u64* a = (const u64*)racc + 3 * (I.red0 + j);
                a0[c] = a[0];
                a1[c] = a[1];
                a2[c] = a[2];
            } else {
                a0[c] = ((const u64*)racc)[I.red0 + j];
            }
            if (TEST) fv[c] = F[j];
        }
    }
    SolveRegs R;
    solve_prefetch(R, twp, N, log2N);
#pragma unroll
    for (int c = 0; c < kPer; ++c) {
        const int j = threadIdx.x + c * kSolveThreads;
        if (j < N) {
            const double r = precision == 0 ? fxp_to(limbs_to_u128(a0[c], a1[c], a2[c]), I.q) : (double)a0[c] * I.wq;
            const int b = padi((int)brev((uint32_t)j, log2N));
            re[b] = r;
            im[b] = 0.0;
            if (TEST) Fl[j] = fv[c];
        }
    }
    __syncthreads();
    solve_lds(re, im, R, twp, N, log2N);
    bool conv = false;
    double emax;
    if (TEST) {
        conv = block_isapprox(red, [&](int i) { return Fl[i]; }, [&](int i) { return solve_at(re, i); }, N, I.rtol,
                              0.0, emax);
    } else {
        emax = block_maxabs(red, N, [&](int64_t i) { return solve_at(re, (int)i); });
    }
    // mixed: when the next push is a middle sweep (k >= 1, the loop goes on)
    // it adds only the change of the charge, so the charge stays
    const bool keep = TEST && precision != 0 && k >= 1 && !(conv || k == I.max_sweeps - 1);
    // global stores last: every barrier above waits for a thread's outstanding
    // stores, and nothing in this kernel reads these buffers back
    double* Eo = (nxt ? E1 : E0) + I.goff;
    for (int j = threadIdx.x; j < N; j += blockDim.x) {
        const double e = solve_at(re, j);
        Eo[j] = e;
        if (precision != 0) E32[I.goff + j] = (float)(e * I.escale2);  // E*dt in steps of the position, x 2^-30
        if (precision == 0) {
            u64* a = (u64*)racc + 3 * (I.red0 + j);
            a[0] = 0;
            a[1] = 0;
            a[2] = 0;
        } else if (!keep) {
            ((u64*)racc)[I.red0 + j] = 0;
        }
    }
    if (threadIdx.x == 0) {
        if (!isfinite(emax)) S.err |= 1;
        if (TEST) {
            S.e_cur = nxt;
            if (conv || k == I.max_sweeps - 1) S.last_sweep = k;
            else if (stall) *halt = 1;
        } else {
            S.e_cur = 0;
            S.last_sweep = 0;
        }
        st[blockIdx.x] = S;
    }
}

// ---------------------------------------------------------------------------
// Diagnostics (src/GaussianFixedPointQuiet.jl:11,13):
// D[1:2] = (sum(E.^2)/N, sum(v.^2)*W/P) ./ 2; D[3:4] = (D1 + D2, sum(v/P));
// D[1:3] .*= 2/W.  Sums use the canonical tree.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_diag_chunks64(const InstDesc* __restrict__ insts, const InstState* __restrict__ st,
                const ChunkDesc* __restrict__ chunks, const double* V0, const double* V1,
                double* __restrict__ dpart, const int32_t* __restrict__ halt) {
    __shared__ double red[4 * kTreeLanes + 8];
    if (halted(halt, false)) return;
    const ChunkDesc C = chunks[blockIdx.x];
    const InstDesc I = insts[C.inst];
    const InstState S = st[C.inst];
    // the step's result sits in the buffer of the next parity
    const double* v = (((S.step + 1) & 1) ? V1 : V0) + I.poff + C.start;
    double s2, s1;
    chunk_sums_v(red, C.count, (double)I.P, [&](int64_t i) { return v[i]; }, s2, s1);
    if (threadIdx.x == 0) {
        dpart[2 * blockIdx.x] = s2;
        dpart[2 * blockIdx.x + 1] = s1;
    }
}

__global__ void __launch_bounds__(256)
k_diag_final(const InstDesc* __restrict__ insts, InstState* __restrict__ st,
             double* __restrict__ dpart, const double* E0, const double* E1,
             double* __restrict__ diag, int32_t* __restrict__ sweeps, int64_t cap,
             int scheme, int32_t* __restrict__ halt, int tile_partials) {
    __shared__ double red[4 * kTreeLanes + 8];
    // halt, the instance and its state load together ahead of the test
    // (I.N is never 0 and e_cur is 0 or 1: those operands only keep the
    // loads ahead of the branch)
    const int32_t h = *halt;
    const InstDesc I = insts[blockIdx.x];
    InstState S = st[blockIdx.x];
    if ((h != 0) | (I.N <= 0) | (S.e_cur < 0)) {
        // the step did not finish: no diagnostics; from here on the graph is halted
        if (h != 0 && blockIdx.x == 0 && threadIdx.x == 0) *halt = 2;
        return;
    }
    const double* E = (S.e_cur ? E1 : E0) + I.goff;
    double se, sv, sp;
    if (tile_partials) {
        // mixed: exact sums of v^2 and v, one 128-bit pair per tile written by
        // the final sweep of the push; the total does not depend on the tiles
        se = block_chunk_sum(red, I.N, [&](int64_t i) { return E[i] * E[i]; });
        const u64* tp = reinterpret_cast<const u64*>(dpart) + 4 * (int64_t)I.tile0;
        i128 a = 0, b = 0;
        for (int t = threadIdx.x; t < I.ntiles; t += blockDim.x) {
            a += (i128)(((u128)tp[4 * t + 1] << 64) | tp[4 * t]);
            b += (i128)(((u128)tp[4 * t + 3] << 64) | tp[4 * t + 2]);
        }
        a = wave_sum_i128(a);
        b = wave_sum_i128(b);
        u64* w = reinterpret_cast<u64*>(red);
        const int wv = threadIdx.x >> 6;
        if ((threadIdx.x & 63) == 0) {
            w[4 * wv] = (u64)a;
            w[4 * wv + 1] = (u64)((u128)a >> 64);
            w[4 * wv + 2] = (u64)b;
            w[4 * wv + 3] = (u64)((u128)b >> 64);
        }
        __syncthreads();
        a = 0;
        b = 0;
        for (int q = 0; q < (int)(blockDim.x >> 6); ++q) {
            a += (i128)(((u128)w[4 * q + 1] << 64) | w[4 * q]);
            b += (i128)(((u128)w[4 * q + 3] << 64) | w[4 * q + 2]);
        }
        sv = i128_to_double(a, kV2Bits);
        sp = i128_to_double(b, kV1Bits) / (double)I.P;
    } else {
        // canonical partials per 4096-particle chunk
        diag_sums(red, E, I.N, dpart + 2 * (int64_t)I.chunk0, I.nchunks, se, sv, sp);
    }
    if (threadIdx.x == 0) {
        const double d1 = (se / (double)I.N) / 2.0;
        const double d2 = ((sv * I.W) / (double)I.P) / 2.0;
        const double d3 = d1 + d2;
        const double s = 2.0 / I.W;
        const int64_t row = (int64_t)(S.step - S.run_base);
        if (row >= 0 && row < cap) {
            double* o = diag + ((int64_t)blockIdx.x * cap + row) * 4;
            o[0] = d1 * s;
            o[1] = d2 * s;
            o[2] = d3 * s;
            o[3] = sp;
            sweeps[(int64_t)blockIdx.x * cap + row] = scheme != 0 ? S.last_sweep + 1 : 1;
        }
        // the step is complete: the next one starts a fresh Picard loop (no
        // first-deposit launch does this where the final sweep primes the
        // next step's charge)
        S.last_sweep = -1;
        S.step += 1;
        st[blockIdx.x] = S;
    }
}

// ---------------------------------------------------------------------------
// Simpson-1/3 fixed point (src/GaussianFixedPointQuietSimpson13.jl:8-17,
// src/AreaFixedPointQuietSimpson13.jl:7-16).  Per step E[1,:] =
// solve(rho(X, X)); per sweep k, with x_k = X + ((v_k + V)/2)*dt,
//   v_{k+1} = ((V + (g1*dt)/6) + (g2*(4*dt))/6) + (g3*dt)/6,
// g1 = gather(E[1,:], (X+X)/2), g2 = gather(E[2,:], (X+x_k)/2) with
// E[2,:] = solve(rho(X, x_k)), g3 = gather(E[3,:], (x_k+x_k)/2) with
// E[3,:] = solve(rho(x_k, x_k)); the convergence test compares the 3 x N
// field matrix of consecutive sweeps.  SH is the shape: 1 = the linear area
// weights of src/AreaFixedPointQuietSimpson13.jl:5, 6 / 7 = Gaussian.
// The push of sweep k computes v_{k+1} and, unless the loop ends there,
// deposits both charges of sweep k+1, like the midpoint push.
// ---------------------------------------------------------------------------

// rho(X, X) of the step into racc (the new loop starts: last_sweep = -1).
template <int SH>
__global__ void __launch_bounds__(1024)
k_simp_first(const InstDesc* __restrict__ insts, InstState* __restrict__ st, const TileDesc* __restrict__ tiles,
             const double* X0, const double* X1, u64* __restrict__ racc, const int32_t* __restrict__ halt) {
    extern __shared__ u64 lds[];
    if (halted(halt, false)) return;
    const TileDesc T = tiles[blockIdx.x];
    const InstDesc I = insts[T.inst];
    const InstState S = st[T.inst];
    if (T.local == 0 && threadIdx.x == 0) st[T.inst].last_sweep = -1;
    const int32_t N = I.N;
    u64* acc = lds;
    zero_lds(acc, 3 * N);
    __syncthreads();
    const double* X = (S.step & 1) ? X1 : X0;
    const double Nd = (double)N;
    const int64_t base = I.poff + T.start;
    for (int64_t p = threadIdx.x; p < T.count; p += blockDim.x) {
        const double Xp = X[base + p];
        deposit_shape<SH>(acc, N, (Xp + Xp) / 2.0, Nd, I.nmask, I.w, I.q);
    }
    __syncthreads();
    flush_limbs(acc, N, racc + 3 * I.red0);
}

// Reads and clears an instance's charge into re/im in bit-reversed order:
// three limbs (FP64) or u64 quanta of wq each (mixed).
__device__ __forceinline__ void take_charge(u64* __restrict__ racc, const InstDesc& I, int precision, double* re,
                                            double* im) {
    if (precision == 0) return take_charge64(racc + 3 * I.red0, re, im, I.N, I.log2N, I.q);
    for (int j = threadIdx.x; j < I.N; j += blockDim.x) {
        u64* p = racc + I.red0 + j;
        const u64 a = *p;
        *p = 0;
        const int b = padi((int)brev((uint32_t)j, I.log2N));
        re[b] = (double)a * I.wq;
        im[b] = 0.0;
    }
}

// E[1,:] = solve(rho(X, X)): one block per instance; mixed: also the FP32
// copy for the gathers (E32s1, E * dt^2 * 4).
__global__ void __launch_bounds__(kSolveThreads)
k_simp_solve1(const InstDesc* __restrict__ insts, InstState* __restrict__ st, u64* __restrict__ racc,
              int precision, double* __restrict__ Es1, float* __restrict__ E32s1, const double2* __restrict__ twp,
              const int32_t* __restrict__ halt) {
    extern __shared__ double sm[];
    if (halted(halt, false)) return;
    const InstDesc I = insts[blockIdx.x];
    const int N = I.N, log2N = I.log2N, NP = solve_np(N);
    double* re = sm;
    double* im = sm + NP;
    double* red = sm + 2 * NP + N;
    SolveRegs R;
    solve_prefetch(R, twp, N, log2N);
    take_charge(racc, I, precision, re, im);
    __syncthreads();
    solve_lds(re, im, R, twp, N, log2N);
    for (int j = threadIdx.x; j < N; j += blockDim.x) {
        const double e = solve_at(re, j);
        Es1[I.goff + j] = e;
        if (precision != 0) E32s1[I.goff + j] = (float)(e * I.escale2);
    }
    const double emax = block_maxabs(red, N, [&](int64_t i) { return solve_at(re, (int)i); });
    if (threadIdx.x == 0 && !isfinite(emax)) st[blockIdx.x].err |= 1;
}

// Per particle: Vg = V + (g1*dt)/6 and, with v_0 = V, the charges rho(X, x_0)
// and rho(x_0, x_0) of sweep 0.
template <int SH>
__global__ void __launch_bounds__(1024)
k_simp_start(const InstDesc* __restrict__ insts, const InstState* __restrict__ st,
             const TileDesc* __restrict__ tiles, const double* X0, const double* X1, const double* V0,
             const double* V1, double* __restrict__ Vg, const double* __restrict__ Es1, u64* __restrict__ racc2,
             u64* __restrict__ racc3, const int32_t* __restrict__ halt) {
    extern __shared__ u64 lds[];
    if (halted(halt, false)) return;
    const TileDesc T = tiles[blockIdx.x];
    const InstDesc I = insts[T.inst];
    const InstState S = st[T.inst];
    const int32_t N = I.N;
    double* E1ld = reinterpret_cast<double*>(lds);
    u64* acc2 = lds + N;
    u64* acc3 = lds + 4 * N;
    for (int i = threadIdx.x; i < N; i += blockDim.x) E1ld[i] = Es1[I.goff + i];
    zero_lds(acc2, 6 * N);
    __syncthreads();
    const int par = S.step & 1;
    const double* X = par ? X1 : X0;
    const double* V = par ? V1 : V0;
    const double Nd = (double)N, dt = I.dt;
    const int64_t base = I.poff + T.start;
    for (int64_t p = threadIdx.x; p < T.count; p += blockDim.x) {
        const int64_t g = base + p;
        const double Xp = X[g], Vp = V[g];
        const double g1 = gather_shape<SH>(E1ld, (Xp + Xp) / 2.0, Nd, I.nmask);
        Vg[g] = Vp + (g1 * dt) / 6.0;
        const double x0 = Xp + ((Vp + Vp) / 2.0) * dt;
        deposit_shape<SH>(acc2, N, (Xp + x0) / 2.0, Nd, I.nmask, I.w, I.q);
        deposit_shape<SH>(acc3, N, (x0 + x0) / 2.0, Nd, I.nmask, I.w, I.q);
    }
    __syncthreads();
    flush_limbs(acc2, N, racc2 + 3 * I.red0);
    flush_limbs(acc3, N, racc3 + 3 * I.red0);
}

// Sweep k: E[2,:] and E[3,:] from the two charges, then the test the script
// runs at the top of sweep k+1, isapprox(F, E) over the 3 x N matrices in
// column-major order (F: the matrix at the start of sweep k).  One block per
// instance; requires 3N <= 4096 (one canonical chunk).
__global__ void __launch_bounds__(kSolveThreads)
k_simp_solve(const InstDesc* __restrict__ insts, InstState* __restrict__ st, u64* __restrict__ racc2,
             u64* __restrict__ racc3, int precision, const double* __restrict__ Es1, double* Es2a, double* Es2b,
             double* E3a, double* E3b, float* __restrict__ E32s2, float* __restrict__ E32s3,
             const double2* __restrict__ twp, int k, int32_t* __restrict__ halt, int stall) {
    extern __shared__ double sm[];
    if (halted(halt, false)) return;
    const InstDesc I = insts[blockIdx.x];
    InstState S = st[blockIdx.x];
    if (sweep_over(S, k)) return;
    const int N = I.N, log2N = I.log2N, NP = solve_np(N);
    double* re = sm;
    double* im = sm + NP;
    double* e2n = sm + 2 * NP;
    double* red = e2n + N;
    const int cur = S.e_cur, nxt = 1 - cur;
    const double* E1 = Es1 + I.goff;
    const double* E2c = (cur ? Es2b : Es2a) + I.goff;
    const double* E3c = (cur ? E3b : E3a) + I.goff;
    double* E2o = (nxt ? Es2b : Es2a) + I.goff;
    double* E3o = (nxt ? E3b : E3a) + I.goff;
    SolveRegs R;
    solve_prefetch(R, twp, N, log2N);
    take_charge(racc2, I, precision, re, im);
    __syncthreads();
    solve_lds(re, im, R, twp, N, log2N);
    for (int j = threadIdx.x; j < N; j += blockDim.x) {
        const double e = solve_at(re, j);
        E2o[j] = e;
        e2n[j] = e;
        if (precision != 0) E32s2[I.goff + j] = (float)(e * I.escale2);
    }
    __syncthreads();
    solve_prefetch(R, twp, N, log2N);  // the first solve reloaded inverse twiddles
    take_charge(racc3, I, precision, re, im);
    __syncthreads();
    solve_lds(re, im, R, twp, N, log2N);
    for (int j = threadIdx.x; j < N; j += blockDim.x) {
        const double e = solve_at(re, j);
        E3o[j] = e;
        if (precision != 0) E32s3[I.goff + j] = (float)(e * I.escale2);
    }
    double emax;
    const bool conv = block_isapprox(
        red,
        [&](int e) {
            const int r = e % 3, i = e / 3;
            return r == 0 ? E1[i] : (r == 1 ? E2c[i] : E3c[i]);
        },
        [&](int e) {
            const int r = e % 3, i = e / 3;
            return r == 0 ? E1[i] : (r == 1 ? e2n[i] : solve_at(re, i));
        },
        3 * N, I.rtol, 0.0, emax);
    if (threadIdx.x == 0) {
        if (!isfinite(emax)) S.err |= 1;
        S.e_cur = nxt;
        if (conv || k == I.max_sweeps - 1) S.last_sweep = k;
        else if (stall) *halt = 1;
        st[blockIdx.x] = S;
    }
}

// Sweep k: v_{k+1} from g2 and g3 at x_k; the final sweep writes
// (mod(x_k, 1), v_{k+1}), any other deposits rho(X, x_{k+1}) and rho(x_{k+1}, x_{k+1}).
template <int SH>
__global__ void __launch_bounds__(1024)
k_simp_push(const InstDesc* __restrict__ insts, const InstState* __restrict__ st,
            const TileDesc* __restrict__ tiles, double* X0, double* X1, double* V0, double* V1, double* vit0,
            double* vit1, const double* __restrict__ Vg, const double* Es2a, const double* Es2b,
            const double* E3a, const double* E3b, u64* __restrict__ racc2, u64* __restrict__ racc3, int k,
            const int32_t* __restrict__ halt, int last_slot) {
    extern __shared__ u64 lds[];
    if (halted(halt, last_slot)) return;
    const TileDesc T = tiles[blockIdx.x];
    const InstDesc I = insts[T.inst];
    const InstState S = st[T.inst];
    if (sweep_over(S, k)) return;
    const bool last = S.last_sweep == k;
    const int32_t N = I.N;
    double* E2ld = reinterpret_cast<double*>(lds);
    double* E3ld = E2ld + N;
    u64* acc2 = lds + 2 * N;
    u64* acc3 = lds + 5 * N;
    const double* E2in = (S.e_cur ? Es2b : Es2a) + I.goff;
    const double* E3in = (S.e_cur ? E3b : E3a) + I.goff;
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
        E2ld[i] = E2in[i];
        E3ld[i] = E3in[i];
    }
    if (!last) zero_lds(acc2, 6 * N);
    __syncthreads();
    const int par = S.step & 1;
    const double* X = par ? X1 : X0;
    const double* V = par ? V1 : V0;
    double* Xo = par ? X0 : X1;
    double* Vo = par ? V0 : V1;
    const double* vin = (k & 1) ? vit1 : vit0;
    double* vout = ((k + 1) & 1) ? vit1 : vit0;
    const double Nd = (double)N, dt = I.dt;
    const int32_t nmask = I.nmask;
    const int64_t base = I.poff + T.start;
    for (int64_t p = threadIdx.x; p < T.count; p += blockDim.x) {
        const int64_t g = base + p;
        const double Xp = X[g], Vp = V[g];
        const double vk = (k == 0) ? Vp : vin[g];
        const double x = Xp + ((vk + Vp) / 2.0) * dt;
        const double g2 = gather_shape<SH>(E2ld, (Xp + x) / 2.0, Nd, nmask);
        const double g3 = gather_shape<SH>(E3ld, (x + x) / 2.0, Nd, nmask);
        double vn = Vg[g] + (g2 * (4.0 * dt)) / 6.0;
        vn = vn + (g3 * dt) / 6.0;
        if (last) {
            Xo[g] = jl_mod1f(x);
            Vo[g] = vn;
            continue;
        }
        vout[g] = vn;
        const double x2 = Xp + ((vn + Vp) / 2.0) * dt;
        deposit_shape<SH>(acc2, N, (Xp + x2) / 2.0, Nd, nmask, I.w, I.q);
        deposit_shape<SH>(acc3, N, (x2 + x2) / 2.0, Nd, nmask, I.w, I.q);
    }
    if (last) return;
    __syncthreads();
    flush_limbs(acc2, N, racc2 + 3 * I.red0);
    flush_limbs(acc3, N, racc3 + 3 * I.red0);
}

// ---------------------------------------------------------------------------
// Host launchers for the run
// ---------------------------------------------------------------------------
int push_lds_bytes(int N, int precision) {
    if (precision == 0) return N * 8 * 4;  // E + three limb arrays
    return push32_lds_bytes(N);             // FP32 E + u64 accumulators with halos, partials
}

void launch_mid_push0(const RunBuffers& b, hipStream_t s) {
    if (b.precision == 1) return launch_mid_push32(b, 0, 0, false, s);
    const dim3 g(b.ntiles), t(b.block_push);
    u64* part = (u64*)b.racc;
    if (b.R == 6)
        k_mid_push64<6, 0><<<g, t, b.lds_push, s>>>(b.insts, b.st, b.tiles, b.X[0], b.X[1], b.V[0], b.V[1],
                                                   b.vit[0], b.vit[1], b.E[0], b.E[1], part, 0, b.halt, 0, 0);
    else
        k_mid_push64<7, 0><<<g, t, b.lds_push, s>>>(b.insts, b.st, b.tiles, b.X[0], b.X[1], b.V[0], b.V[1],
                                                   b.vit[0], b.vit[1], b.E[0], b.E[1], part, 0, b.halt, 0, 0);
}

void launch_mid_push(const RunBuffers& b, int k, bool last_slot, hipStream_t s) {
    if (b.precision == 1) return launch_mid_push32(b, 1, k, last_slot, s);
    const int ls = last_slot ? 1 : 0;
    const dim3 g(b.ntiles), t(b.block_push);
    u64* part = (u64*)b.racc;
    if (b.R == 6)
        k_mid_push64<6, 1><<<g, t, b.lds_push, s>>>(b.insts, b.st, b.tiles, b.X[0], b.X[1], b.V[0], b.V[1],
                                                   b.vit[0], b.vit[1], b.E[0], b.E[1], part, k, b.halt, ls,
                                                   b.prime_next);
    else
        k_mid_push64<7, 1><<<g, t, b.lds_push, s>>>(b.insts, b.st, b.tiles, b.X[0], b.X[1], b.V[0], b.V[1],
                                                   b.vit[0], b.vit[1], b.E[0], b.E[1], part, k, b.halt, ls,
                                                   b.prime_next);
}

static int solve_lds_bytes(int N) { return solve_lds_doubles(N) * 8; }

void launch_solve(const RunBuffers& b, int k, bool stall, hipStream_t s) {
    const int lds = solve_lds_bytes(b.maxN);
    if (b.scheme == 1)
        k_solve<true><<<b.ninst, kSolveThreads, lds, s>>>(b.insts, b.st, b.racc, b.precision, b.E[0], b.E[1], b.E32,
                                                (const double2*)b.tw, k, b.halt, stall ? 1 : 0);
    else
        k_solve<false><<<b.ninst, kSolveThreads, lds, s>>>(b.insts, b.st, b.racc, b.precision, b.E[0], b.E[1], b.E32,
                                                 (const double2*)b.tw, k, b.halt, 0);
}

void launch_lf_first(const RunBuffers& b, hipStream_t s) {
    const dim3 g(b.ntiles), t(b.block_push);
    u64* part = (u64*)b.racc;
    if (b.R == 0)
        k_lf_first64<0><<<g, t, b.lds_push, s>>>(b.insts, b.st, b.tiles, b.X[0], b.X[1], b.V[0], b.V[1], part);
    else if (b.R == 6)
        k_lf_first64<6><<<g, t, b.lds_push, s>>>(b.insts, b.st, b.tiles, b.X[0], b.X[1], b.V[0], b.V[1], part);
    else
        k_lf_first64<7><<<g, t, b.lds_push, s>>>(b.insts, b.st, b.tiles, b.X[0], b.X[1], b.V[0], b.V[1], part);
}

void launch_lf_second(const RunBuffers& b, hipStream_t s) {
    const dim3 g(b.ntiles), t(b.block_push);
    if (b.R == 0)
        k_lf_second64<0><<<g, t, b.lds_push, s>>>(b.insts, b.st, b.tiles, b.X[0], b.X[1], b.V[0], b.V[1],
                                                 b.V[0], b.V[1], b.E[0]);
    else if (b.R == 6)
        k_lf_second64<6><<<g, t, b.lds_push, s>>>(b.insts, b.st, b.tiles, b.X[0], b.X[1], b.V[0], b.V[1],
                                                 b.V[0], b.V[1], b.E[0]);
    else
        k_lf_second64<7><<<g, t, b.lds_push, s>>>(b.insts, b.st, b.tiles, b.X[0], b.X[1], b.V[0], b.V[1],
                                                 b.V[0], b.V[1], b.E[0]);
}

void launch_diag(const RunBuffers& b, hipStream_t s) {
    // the mixed midpoint and Simpson pushes sum the kinetic terms in their
    // final sweep
    const bool fused = b.precision == 1 && (b.scheme == 1 || b.scheme == 2);
    if (b.precision == 0)
        k_diag_chunks64<<<b.nchunks, 256, 0, s>>>(b.insts, b.st, b.chunks, b.V[0], b.V[1], b.dpart, b.halt);
    k_diag_final<<<b.ninst, 256, 0, s>>>(b.insts, b.st, b.dpart, b.E[0], b.E[1], b.diag, b.sweeps,
                                         b.diag_cap, b.scheme, b.halt, fused ? 1 : 0);
}

int simpson_push_lds_bytes(int N) { return 64 * N; }  // two fields + two 3-limb charges

void launch_simp_first(const RunBuffers& b, hipStream_t s) {
    const dim3 g(b.ntiles), t(b.block_push);
    const int la = 24 * b.maxN, lb = 56 * b.maxN;
    u64* r1 = (u64*)b.racc;
    u64* r2 = (u64*)b.racc2;
    u64* r3 = (u64*)b.racc3;
    if (b.precision == 1) {
        launch_simp32_first(b, s);
        k_simp_solve1<<<b.ninst, kSolveThreads, solve_lds_bytes(b.maxN), s>>>(b.insts, b.st, r1, 1, b.Es1, b.E32s1,
                                                                             (const double2*)b.tw, b.halt);
        launch_simp32_start(b, s);
        return;
    }
    if (b.R == 1) k_simp_first<1><<<g, t, la, s>>>(b.insts, b.st, b.tiles, b.X[0], b.X[1], r1, b.halt);
    else if (b.R == 6) k_simp_first<6><<<g, t, la, s>>>(b.insts, b.st, b.tiles, b.X[0], b.X[1], r1, b.halt);
    else k_simp_first<7><<<g, t, la, s>>>(b.insts, b.st, b.tiles, b.X[0], b.X[1], r1, b.halt);
    k_simp_solve1<<<b.ninst, kSolveThreads, solve_lds_bytes(b.maxN), s>>>(b.insts, b.st, r1, 0, b.Es1, nullptr,
                                                                         (const double2*)b.tw, b.halt);
    if (b.R == 1)
        k_simp_start<1><<<g, t, lb, s>>>(b.insts, b.st, b.tiles, b.X[0], b.X[1], b.V[0], b.V[1], b.Vg, b.Es1, r2,
                                         r3, b.halt);
    else if (b.R == 6)
        k_simp_start<6><<<g, t, lb, s>>>(b.insts, b.st, b.tiles, b.X[0], b.X[1], b.V[0], b.V[1], b.Vg, b.Es1, r2,
                                         r3, b.halt);
    else
        k_simp_start<7><<<g, t, lb, s>>>(b.insts, b.st, b.tiles, b.X[0], b.X[1], b.V[0], b.V[1], b.Vg, b.Es1, r2,
                                         r3, b.halt);
}

void launch_simp_solve(const RunBuffers& b, int k, bool stall, hipStream_t s) {
    k_simp_solve<<<b.ninst, kSolveThreads, solve_lds_bytes(b.maxN), s>>>(
        b.insts, b.st, (u64*)b.racc2, (u64*)b.racc3, b.precision, b.Es1, b.Es2[0], b.Es2[1], b.E[0], b.E[1],
        b.E32s2, b.E32, (const double2*)b.tw, k, b.halt, stall ? 1 : 0);
}

void launch_simp_push(const RunBuffers& b, int k, bool last_slot, hipStream_t s) {
    if (b.precision == 1) return launch_simp32_push(b, k, last_slot, s);
    const dim3 g(b.ntiles), t(b.block_push);
    const int lds = simpson_push_lds_bytes(b.maxN), ls = last_slot ? 1 : 0;
    u64* r2 = (u64*)b.racc2;
    u64* r3 = (u64*)b.racc3;
    if (b.R == 1)
        k_simp_push<1><<<g, t, lds, s>>>(b.insts, b.st, b.tiles, b.X[0], b.X[1], b.V[0], b.V[1], b.vit[0],
                                         b.vit[1], b.Vg, b.Es2[0], b.Es2[1], b.E[0], b.E[1], r2, r3, k, b.halt, ls);
    else if (b.R == 6)
        k_simp_push<6><<<g, t, lds, s>>>(b.insts, b.st, b.tiles, b.X[0], b.X[1], b.V[0], b.V[1], b.vit[0],
                                         b.vit[1], b.Vg, b.Es2[0], b.Es2[1], b.E[0], b.E[1], r2, r3, k, b.halt, ls);
    else
        k_simp_push<7><<<g, t, lds, s>>>(b.insts, b.st, b.tiles, b.X[0], b.X[1], b.V[0], b.V[1], b.vit[0],
                                         b.vit[1], b.Vg, b.Es2[0], b.Es2[1], b.E[0], b.E[1], r2, r3, k, b.halt, ls);
}

// ---------------------------------------------------------------------------
// Operators: the scripts' closures one at a time on contiguous arrays.
// ---------------------------------------------------------------------------
constexpr int kOpTileP = 16384;

int op_tiles(int64_t P) {
    int64_t t = (P + kOpTileP - 1) / kOpTileP;
    return (int)(t < 1 ? 1 : t);
}

template <int R>
__global__ void __launch_bounds__(256)
k_op_deposit(const double* __restrict__ x, const double* __restrict__ y, int64_t P, double w,
             int q, int N, u128* __restrict__ part) {
    extern __shared__ u64 lds[];
    zero_lds(lds, R == 0 ? N : 3 * N);
    __syncthreads();
    const double Nd = (double)N;
    const int32_t nmask = N - 1;
    const int64_t start = (int64_t)blockIdx.x * kOpTileP;
    const int64_t end = start + kOpTileP < P ? start + kOpTileP : P;
    for (int64_t j = start + threadIdx.x; j < end; j += blockDim.x) {
        const double xj = x[j];
        if (R == 0) {
            atomicAdd(&lds[ngp_cell(xj, Nd, nmask)], 1ull);
        } else {
            const double c = y ? (xj + y[j]) / 2.0 : (xj + xj) / 2.0;
            deposit_shape<R>(lds, N, c, Nd, nmask, w, q);
        }
    }
    __syncthreads();
    u128* out = part + (int64_t)blockIdx.x * N;
    if (R == 0) {
        const u128 one = fxp_from(w, q);
        for (int c = threadIdx.x; c < N; c += blockDim.x) out[c] = (u128)lds[c] * one;
    } else {
        for (int c = threadIdx.x; c < N; c += blockDim.x)
            out[c] = limbs_to_u128(lds[c], lds[N + c], lds[2 * N + c]);
    }
}

__global__ void __launch_bounds__(256)
k_op_reduce(const u128* __restrict__ part, int ntiles, int N, int q, double* __restrict__ rho) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    u128 s = 0;
    for (int t = 0; t < ntiles; ++t) s += part[(int64_t)t * N + c];
    rho[c] = fxp_to(s, q);
}

__global__ void __launch_bounds__(kSolveThreads)
k_op_fieldsolve(const double* __restrict__ rho, double* __restrict__ E, const double2* __restrict__ twp,
                int N, int log2N) {
    extern __shared__ double sm[];
    const int NP = solve_np(N);
    double* re = sm;
    double* im = sm + NP;
    SolveRegs R;
    solve_prefetch(R, twp, N, log2N);
    for (int j = threadIdx.x; j < N; j += blockDim.x) {
        const int b = padi((int)brev((uint32_t)j, log2N));
        re[b] = rho[j];
        im[b] = 0.0;
    }
    __syncthreads();
    solve_lds(re, im, R, twp, N, log2N);
    for (int j = threadIdx.x; j < N; j += blockDim.x) E[j] = solve_at(re, j);
}

__global__ void k_op_drift_half(double* __restrict__ x, const double* __restrict__ v, int64_t P, double dt) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < P) x[j] = jl_mod1f(x[j] + (v[j] / 2.0) * dt);
}

template <int R>
__global__ void __launch_bounds__(256)
k_op_push_leapfrog(double* __restrict__ x, double* __restrict__ v, const double* __restrict__ E,
                   int64_t P, double dt, int N) {
    extern __shared__ double Eld[];
    for (int i = threadIdx.x; i < N; i += blockDim.x) Eld[i] = E[i];
    __syncthreads();
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= P) return;
    const double Nd = (double)N;
    const double vj = v[j];
    const double xj = jl_mod1f(x[j] + (vj / 2.0) * dt);
    x[j] = xj;
    const double Sg = R == 0 ? Eld[ngp_cell(xj, Nd, N - 1)] : gather_gauss<R>(Eld, xj, Nd, N - 1);
    v[j] = vj + Sg * dt;
}

__global__ void k_op_midpos(const double* __restrict__ X, const double* __restrict__ V,
                            const double* __restrict__ v, double* __restrict__ x, int64_t P, double dt) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < P) x[j] = X[j] + ((v[j] + V[j]) / 2.0) * dt;
}

template <int R>
__global__ void __launch_bounds__(256)
k_op_push_midpoint(const double* __restrict__ X, const double* __restrict__ V,
                   const double* __restrict__ x, double* __restrict__ v,
                   const double* __restrict__ E, int64_t P, double dt, int N) {
    extern __shared__ double Eld[];
    for (int i = threadIdx.x; i < N; i += blockDim.x) Eld[i] = E[i];
    __syncthreads();
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= P) return;
    const double Sg = gather_gauss<R>(Eld, (x[j] + X[j]) / 2.0, (double)N, N - 1);
    v[j] = V[j] + Sg * dt;
}

// v_out[j] = v_in[j] + sum(k->E[k[1]]*k[2], d((a[j]+b[j])/2))*coef/6: one
// stage kick of the Simpson-1/3 sweep (src/GaussianFixedPointQuietSimpson13.jl:12,14,16
// with (a, b, coef) = (X, X, dt), (X, x, 4dt), (x, x, dt)).
template <int SH>
__global__ void __launch_bounds__(256)
k_op_stage_kick(const double* __restrict__ a, const double* __restrict__ b, const double* __restrict__ vin,
                double* __restrict__ vout, const double* __restrict__ E, int64_t P, double coef, int N) {
    extern __shared__ double Eld[];
    for (int i = threadIdx.x; i < N; i += blockDim.x) Eld[i] = E[i];
    __syncthreads();
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= P) return;
    const double g = gather_shape<SH>(Eld, (a[j] + b[j]) / 2.0, (double)N, N - 1);
    vout[j] = vin[j] + (g * coef) / 6.0;
}

__global__ void k_op_wrap(double* __restrict__ x, int64_t P) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < P) x[j] = jl_mod1f(x[j]);
}

__global__ void __launch_bounds__(256)
k_op_isapprox(const double* __restrict__ F, const double* __restrict__ E, int n, double rtol,
              double atol, int32_t* out) {
    __shared__ double red[kRedDoubles];
    double me;
    const bool r = block_isapprox(red, [&](int i) { return F[i]; }, [&](int i) { return E[i]; }, n, rtol, atol, me);
    if (threadIdx.x == 0) *out = r ? 1 : 0;
}

__global__ void __launch_bounds__(256)
k_op_diag_chunks(const double* __restrict__ v, int64_t P, double* __restrict__ dpart) {
    __shared__ double red[4 * kTreeLanes + 8];
    const int64_t start = (int64_t)blockIdx.x * kTreeChunk;
    const int64_t n = P - start < kTreeChunk ? P - start : kTreeChunk;
    const double* vv = v + start;
    double s2, s1;
    chunk_sums_v(red, n, (double)P, [&](int64_t i) { return vv[i]; }, s2, s1);
    if (threadIdx.x == 0) {
        dpart[2 * blockIdx.x] = s2;
        dpart[2 * blockIdx.x + 1] = s1;
    }
}

__global__ void __launch_bounds__(256)
k_op_diag_final(const double* __restrict__ E, int N, double* __restrict__ dpart, int64_t nchunks,
                int64_t P, double W, double* __restrict__ out) {
    __shared__ double red[4 * kTreeLanes + 8];
    double se, sv, sp;
    diag_sums(red, E, N, dpart, nchunks, se, sv, sp);
    if (threadIdx.x == 0) {
        const double d1 = (se / (double)N) / 2.0;
        const double d2 = ((sv * W) / (double)P) / 2.0;
        const double d3 = d1 + d2;
        const double s = 2.0 / W;
        out[0] = d1 * s;
        out[1] = d2 * s;
        out[2] = d3 * s;
        out[3] = sp;
    }
}

static inline int blocks_for(int64_t P, int t) { return (int)((P + t - 1) / t); }

void launch_op_deposit(const OpBuffers& o, const double* x, const double* y, int64_t P, double w,
                       int R, int q, double* rho, hipStream_t s) {
    const int nt = op_tiles(P);
    const int lds = (R == 0 ? 1 : 3) * o.N * 8;
    u128* part = (u128*)o.part;
    if (R == 0) k_op_deposit<0><<<nt, 256, lds, s>>>(x, y, P, w, q, o.N, part);
    else if (R == 1) k_op_deposit<1><<<nt, 256, lds, s>>>(x, y, P, w, q, o.N, part);
    else if (R == 6) k_op_deposit<6><<<nt, 256, lds, s>>>(x, y, P, w, q, o.N, part);
    else k_op_deposit<7><<<nt, 256, lds, s>>>(x, y, P, w, q, o.N, part);
    k_op_reduce<<<(o.N + 255) / 256, 256, 0, s>>>(part, nt, o.N, q, rho);
}

void launch_op_fieldsolve(const OpBuffers& o, const double* rho, double* E, hipStream_t s) {
    k_op_fieldsolve<<<1, kSolveThreads, solve_lds_bytes(o.N), s>>>(rho, E, (const double2*)o.tw, o.N, o.log2N);
}

void launch_op_drift_half(double* x, const double* v, int64_t P, double dt, hipStream_t s) {
    k_op_drift_half<<<blocks_for(P, 256), 256, 0, s>>>(x, v, P, dt);
}

void launch_op_push_leapfrog(const OpBuffers& o, double* x, double* v, const double* E, int64_t P,
                             double dt, int R, hipStream_t s) {
    const int lds = o.N * 8;
    if (R == 0) k_op_push_leapfrog<0><<<blocks_for(P, 256), 256, lds, s>>>(x, v, E, P, dt, o.N);
    else if (R == 6) k_op_push_leapfrog<6><<<blocks_for(P, 256), 256, lds, s>>>(x, v, E, P, dt, o.N);
    else k_op_push_leapfrog<7><<<blocks_for(P, 256), 256, lds, s>>>(x, v, E, P, dt, o.N);
}

void launch_op_midpos(const double* X, const double* V, const double* v, double* x, int64_t P,
                      double dt, hipStream_t s) {
    k_op_midpos<<<blocks_for(P, 256), 256, 0, s>>>(X, V, v, x, P, dt);
}

void launch_op_push_midpoint(const OpBuffers& o, const double* X, const double* V, const double* x,
                             double* v, const double* E, int64_t P, double dt, int R, hipStream_t s) {
    const int lds = o.N * 8;
    if (R == 6) k_op_push_midpoint<6><<<blocks_for(P, 256), 256, lds, s>>>(X, V, x, v, E, P, dt, o.N);
    else k_op_push_midpoint<7><<<blocks_for(P, 256), 256, lds, s>>>(X, V, x, v, E, P, dt, o.N);
}

void launch_op_stage_kick(const OpBuffers& o, const double* a, const double* b, const double* vin, double* vout,
                          const double* E, int64_t P, double coef, int shape, hipStream_t s) {
    const int lds = o.N * 8, g = blocks_for(P, 256);
    if (shape == 1) k_op_stage_kick<1><<<g, 256, lds, s>>>(a, b, vin, vout, E, P, coef, o.N);
    else if (shape == 6) k_op_stage_kick<6><<<g, 256, lds, s>>>(a, b, vin, vout, E, P, coef, o.N);
    else k_op_stage_kick<7><<<g, 256, lds, s>>>(a, b, vin, vout, E, P, coef, o.N);
}

void launch_op_wrap(double* x, int64_t P, hipStream_t s) {
    k_op_wrap<<<blocks_for(P, 256), 256, 0, s>>>(x, P);
}

void launch_op_isapprox(const double* F, const double* E, int32_t n, double rtol, double atol,
                        int32_t* out, hipStream_t s) {
    k_op_isapprox<<<1, 256, 0, s>>>(F, E, n, rtol, atol, out);
}

void launch_op_diagnostics(const OpBuffers& o, const double* E, const double* v, int64_t P, double W,
                           double* out4, hipStream_t s) {
    const int64_t nch = (P + kTreeChunk - 1) / kTreeChunk;
    k_op_diag_chunks<<<(int)nch, 256, 0, s>>>(v, P, o.dpart);
    k_op_diag_final<<<1, 256, 0, s>>>(E, o.N, o.dpart, nch, P, W, out4);
}

bool mixed_available() { return true; }

template <typename K>
static hipError_t allow_lds(K kernel, int bytes) {
    return hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                               hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

// Largest dynamic LDS any launch here asks for: the FP64 push at N = 4096
// (E + three limb arrays, 32 bytes per cell).  A refusal is not fatal: a
// launch that really lacks LDS fails with its own error.
int configure_kernels() {
    const int big = 32 * kMaxN;
    hipError_t e = hipSuccess;
    auto chk = [&](hipError_t r) { if (r != hipSuccess) e = r; };
    chk(allow_lds(k_mid_push64<6, 0>, big));
    chk(allow_lds(k_mid_push64<7, 0>, big));
    chk(allow_lds(k_mid_push64<6, 1>, big));
    chk(allow_lds(k_mid_push64<7, 1>, big));
    chk(allow_lds(k_lf_first64<0>, big));
    chk(allow_lds(k_lf_first64<6>, big));
    chk(allow_lds(k_lf_first64<7>, big));
    chk(allow_lds(k_lf_second64<0>, big));
    chk(allow_lds(k_lf_second64<6>, big));
    chk(allow_lds(k_lf_second64<7>, big));
    chk(allow_lds(k_simp_first<1>, big));
    chk(allow_lds(k_simp_first<6>, big));
    chk(allow_lds(k_simp_first<7>, big));
    chk(allow_lds(k_simp_start<1>, big));
    chk(allow_lds(k_simp_start<6>, big));
    chk(allow_lds(k_simp_start<7>, big));
    chk(allow_lds(k_simp_push<1>, big));
    chk(allow_lds(k_simp_push<6>, big));
    chk(allow_lds(k_simp_push<7>, big));
    chk(allow_lds(k_simp_solve1, big));
    chk(allow_lds(k_simp_solve, big));
    chk(allow_lds(k_solve<true>, big));
    chk(allow_lds(k_solve<false>, big));
    chk(allow_lds(k_op_deposit<0>, big));
    chk(allow_lds(k_op_deposit<1>, big));
    chk(allow_lds(k_op_deposit<6>, big));
    chk(allow_lds(k_op_stage_kick<1>, big));
    chk(allow_lds(k_op_stage_kick<6>, big));
    chk(allow_lds(k_op_stage_kick<7>, big));
    chk(allow_lds(k_op_deposit<7>, big));
    chk(allow_lds(k_op_fieldsolve, big));
    (void)hipGetLastError();
    (void)e;
    if (configure_kernels_small() != 0) return -1;
    return configure_kernels_mixed();
}

}  // namespace pic
