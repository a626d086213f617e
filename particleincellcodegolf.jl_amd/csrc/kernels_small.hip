// kernels_small.hip -- persistent FP64 kernels for small single instances: the
// whole step loop of a run (src/NGPFourier.jl:4-7, src/Gaussian.jl:8-11,
// src/GaussianFixedPointQuiet.jl:7-13) in one launch, with the Picard sweeps
// and their convergence test on the device and no host round trip or graph
// node between steps.
//
// At N <= 256 and a few thousand particles a step is a chain of tiny
// dependent phases (deposit, solve, test, gather) whose cost in the captured
// graph is launch latency, not work.  Two kernels remove it:
//
//  k_small_ngp    one workgroup of 512 threads; the particles stay in
//                 registers for the whole run (thread t owns particles
//                 t + 512 j), counts per cell in LDS, the solve in LDS.
//  k_small_gauss  the Gaussian stencils, whose 16 erf per centre are too much
//                 arithmetic for one compute unit: K particle workgroups with
//                 16 lanes per particle (lane e evaluates edge e of the
//                 stencil, so one wave carries 4 particles and a stencil costs
//                 one erf of latency), plus one diagnostics workgroup.  Each
//                 Picard iteration deposits into LDS limbs, adds them to a
//                 device accumulator, meets at one grid barrier, and then
//                 every workgroup solves the field redundantly (N <= 256 is a
//                 few microseconds in one wave), runs the same convergence test on
//                 the same bits and gathers.  The launch is cooperative, so
//                 all workgroups are resident; every barrier wait is bounded.
//
// Both produce the same bits as the graph path (and the CPU restatement):
// the same stencil expressions, exact integer charge sums, the same solve and
// the same reduction trees.  Compiled with -ffp-contract=off.
#include "pic_device.h"
#include "pic_internal.h"
#include "solve1d.h"
#include "stencil1d.h"

#pragma clang fp contract(off)

namespace pic {

constexpr int kSmallThreads = kSolveThreads;  // threads per workgroup of both kernels
constexpr int kSegLanes = 16;       // lanes per particle in k_small_gauss
constexpr int kSegs = kSmallThreads / kSegLanes;

// LDS layout shared by both kernels (doubles): the forward and the inverse
// transform (re, im each), the previous field, reduction scratch, the field,
// diagnostics partials, the solve's tables (stage twiddles of an N-point
// transform as double2, then the 1/(2 pi k) factors), the u64 charge
// counters (NGP: N counts; Gaussian: 3 N limbs), and (NGP) the exchange of
// the upper half's velocities for the diagnostics tree.
constexpr int kNgpXchg = (kSmallThreads / 2) * kSmallNgpSlots;
__host__ __device__ constexpr int small_lds_doubles(int N) {
    return 4 * N + 2 * N + kRedDoubles + 8 + 2 * N + (N / 2 + 1) + 3 * N;
}
int small_lds_bytes(int N, bool ngp) { return 8 * (small_lds_doubles(N) + (ngp ? kNgpXchg : 0)); }

struct SmallLds {
    double *re, *im, *bre, *bim, *Fl, *red, *Eld, *dp, *xi, *xchg;
    double2* tw;
    u64* acc;
};

__device__ __forceinline__ SmallLds small_lds(double* sm, int N) {
    SmallLds L;
    L.re = sm;
    L.im = sm + N;
    L.bre = sm + 2 * N;
    L.bim = sm + 3 * N;
    L.Fl = sm + 4 * N;
    L.red = L.Fl + N;
    L.Eld = L.red + kRedDoubles;
    L.dp = L.Eld + N;
    L.tw = reinterpret_cast<double2*>(L.dp + 8);
    L.xi = L.dp + 8 + 2 * N;
    L.acc = reinterpret_cast<u64*>(L.xi + N / 2 + 1);
    L.xchg = reinterpret_cast<double*>(L.acc + 3 * N);
    return L;
}

// Stages the solve's tables in LDS: the stage twiddles an N-point transform
// reads (entries below N - 1) and the factors 1/(2 pi k), k <= N/2.
__device__ __forceinline__ void stage_tables(const SmallLds& L, const double2* __restrict__ tw, int N) {
    const double* xi = reinterpret_cast<const double*>(tw) + kXiOffset;
    for (int i = threadIdx.x; i < N; i += blockDim.x) L.tw[i] = tw[i];
    for (int i = threadIdx.x; i <= N / 2; i += blockDim.x) L.xi[i] = xi[i];
    __syncthreads();
}

// ---------------------------------------------------------------------------
// The solve in one wave.  Element i of an N-point (N <= 256) transform lives in
// lane i % 64, slot i / 64 of wave 0.  A radix-2 stage with half-size h < 64
// pairs lanes l and l ^ h (lane shuffles), with h >= 64 slots q and q + h/64
// of one lane; every butterfly is the stage-by-stage transform's (operands
// b = the upper element, a = the lower, t = b w, a + t and a - t; twiddle of
// stage s at table entry 2^(s-1) - 1 + j, conjugated for the inverse), so
// the result is that of the CPU restatement bit for bit.
// ---------------------------------------------------------------------------
constexpr int kWaveSlots = kSmallMaxN / 64;

__device__ __forceinline__ void wave_radix2(double (&re)[kWaveSlots], double (&im)[kWaveSlots],
                                            const double2* __restrict__ tw, int N, int log2N, bool inverse) {
    const int l = threadIdx.x & 63;
    const int E = N >= 64 ? N >> 6 : 1;
    for (int st = 1; st <= log2N; ++st) {
        const int h = 1 << (st - 1);
        if (h < 64) {
            const bool lower = (l & h) == 0;
            const double2 w = tw[h - 1 + (l & (h - 1))];
            const double wr = w.x, wi = inverse ? -w.y : w.y;
#pragma unroll
            for (int q = 0; q < kWaveSlots; ++q) {
                if (q >= E) break;
                const double pr = __shfl_xor(re[q], h, 64), pi = __shfl_xor(im[q], h, 64);
                const double br = lower ? pr : re[q], bi = lower ? pi : im[q];
                const double ar = lower ? re[q] : pr, ai = lower ? im[q] : pi;
                const double tr = br * wr - bi * wi;
                const double tq = br * wi + bi * wr;
                re[q] = lower ? ar + tr : ar - tr;
                im[q] = lower ? ai + tq : ai - tq;
            }
        } else {
            const int hs = h >> 6;
#pragma unroll
            for (int q = 0; q < kWaveSlots; ++q) {
                if ((q & hs) || q + hs >= E) continue;
                const double2 w = tw[h - 1 + ((q * 64 + l) & (h - 1))];
                const double wr = w.x, wi = inverse ? -w.y : w.y;
#pragma unroll
                for (int qq = 0; qq < kWaveSlots; ++qq) {
                    if (qq != q + hs) continue;
                    const double br = re[qq], bi = im[qq];
                    const double tr = br * wr - bi * wi;
                    const double tq = br * wi + bi * wr;
                    const double ar = re[q], ai = im[q];
                    re[q] = ar + tr;
                    im[q] = ai + tq;
                    re[qq] = ar - tr;
                    im[qq] = ai - tq;
                }
            }
        }
    }
}

// The canonical reduction tree (tree256 over lanes t < 256 that each fold the
// elements t, t + 256, ... in order from +0.0) over the N values x of wave 0
// (N <= 256 here; the loop also folds slot u + 4, which N = 512 would fill):
// tree lane t = l + 64u folds slot u of lane l, the two
// halving steps combine (u, u + 2) and then (0, 1) in lane l, and the last
// six steps are tree256's own shuffles.  MAX: the NaN-propagating maximum.
// Every value is >= 0 (squares, magnitudes, flags), so folding +0.0 for a
// missing element leaves a sum or maximum unchanged.  Lane 0's result is
// returned to every lane.
template <bool MAX>
__device__ __forceinline__ double wave_canon(const double (&x)[kWaveSlots], int N) {
    auto op = [](double a, double b) { return MAX ? nanmax(a, b) : a + b; };
    const int l = threadIdx.x & 63;
    double P[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        P[u] = 0.0;
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int q = u + 4 * k;
            if (q < kWaveSlots && q * 64 + l < N) P[u] = op(P[u], x[q]);
        }
    }
    double r = op(op(P[0], P[2]), op(P[1], P[3]));
#pragma unroll
    for (int h = 32; h > 0; h >>= 1) r = op(r, __shfl_down(r, h, 64));
    return __shfl(r, 0, 64);
}

// The solve of the charge staged in re/im (bit-reversed), in wave 0:
// E = real(ifft((xi = fft(rho) ./ ik; xi[1] *= 0; xi))) into Eld, with ik and
// Julia's division by (0, 2 pi k') as in k_solve; with TEST also the
// convergence test isapprox(Fl, E) of the midpoint loop, with block_isapprox's
// norms and decisions (Fl then takes E).  Returns the test's result; emax
// gets max |E|.  Every thread of the block calls it.
template <bool TEST>
__device__ __forceinline__ bool small_solve(const SmallLds& L, int N, int log2N, double rtol, double& emax) {
    if (threadIdx.x < 64) {
        const int l = threadIdx.x;
        double re[kWaveSlots], im[kWaveSlots];
#pragma unroll
        for (int q = 0; q < kWaveSlots; ++q) {
            const int i = q * 64 + l;
            re[q] = i < N ? L.re[i] : 0.0;
            im[q] = i < N ? L.im[i] : 0.0;
        }
        wave_radix2(re, im, L.tw, N, log2N, false);
#pragma unroll
        for (int q = 0; q < kWaveSlots; ++q) {
            const int m = q * 64 + l;
            if (m >= N) continue;
            const double t = m == 0 ? L.xi[1] : (m <= N / 2 ? L.xi[m] : -L.xi[N - m]);
            double a = im[q] * t;
            double b = -(re[q] * t);
            if (m == 0) {
                a = a * 0.0;
                b = b * 0.0;
            }
            const int p = (int)brev((uint32_t)m, log2N);
            L.bre[p] = a;
            L.bim[p] = b;
        }
        // the permutation stays inside the wave: LDS operations of one wave
        // complete in order; the fences keep the compiler from moving them
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
        for (int q = 0; q < kWaveSlots; ++q) {
            const int i = q * 64 + l;
            re[q] = i < N ? L.bre[i] : 0.0;
            im[q] = i < N ? L.bim[i] : 0.0;
        }
        wave_radix2(re, im, L.tw, N, log2N, true);
        const double scale = 1.0 / (double)N;
        double e[kWaveSlots];
#pragma unroll
        for (int q = 0; q < kWaveSlots; ++q) {
            const int i = q * 64 + l;
            e[q] = re[q] * scale;
            if (i < N) L.Eld[i] = e[q];
        }
        bool conv = false;
        double me;
        if (TEST) {
            double f[kWaveSlots], x[kWaveSlots];
#pragma unroll
            for (int q = 0; q < kWaveSlots; ++q) {
                const int i = q * 64 + l;
                f[q] = i < N ? L.Fl[i] : 0.0;
            }
            // the three norms of block_norms3: d = F - E, F, E
            double m3[3], s3[3];
#pragma unroll
            for (int v = 0; v < 3; ++v) {
#pragma unroll
                for (int q = 0; q < kWaveSlots; ++q) {
                    const double xv = v == 0 ? f[q] - e[q] : (v == 1 ? f[q] : e[q]);
                    x[q] = fabs(xv);
                }
                m3[v] = wave_canon<true>(x, N);
#pragma unroll
                for (int q = 0; q < kWaveSlots; ++q) {
                    const double xv = v == 0 ? f[q] - e[q] : (v == 1 ? f[q] : e[q]);
                    x[q] = xv * xv;
                }
                s3[v] = wave_canon<false>(x, N);
            }
            double nrm[3];
#pragma unroll
            for (int v = 0; v < 3; ++v) {
                const bool direct = m3[v] == 0.0 || isinf(m3[v]);
                const bool scaled = !direct && !(isfinite((double)N * m3[v] * m3[v]) && m3[v] * m3[v] != 0.0);
                if (scaled) {
#pragma unroll
                    for (int q = 0; q < kWaveSlots; ++q) {
                        const double xv = v == 0 ? f[q] - e[q] : (v == 1 ? f[q] : e[q]);
                        const double y = fabs(xv) / m3[v];
                        x[q] = y * y;
                    }
                    s3[v] = wave_canon<false>(x, N);
                }
                nrm[v] = direct ? m3[v] : (scaled ? m3[v] * sqrt(s3[v]) : sqrt(s3[v]));
            }
            me = m3[2];
            const double dn = nrm[0], nf = nrm[1], ne = nrm[2];
            if (isfinite(dn)) {
                double mx = nf > ne ? nf : ne;
                if (nf != nf || ne != ne) mx = __longlong_as_double(0x7ff8000000000000LL);
                const double tol = rtol * mx;
                double lim = 0.0 > tol ? 0.0 : tol;
                if (tol != tol) lim = tol;
                conv = dn <= lim;
            } else {
#pragma unroll
                for (int q = 0; q < kWaveSlots; ++q) {
                    const double a = f[q], b = e[q];
                    bool good = a == b;
                    if (!good && isfinite(a) && isfinite(b)) {
                        const double ma = fabs(a) > fabs(b) ? fabs(a) : fabs(b);
                        const double t = rtol * ma;
                        good = fabs(a - b) <= (0.0 > t ? 0.0 : t);
                    }
                    x[q] = good ? 0.0 : 1.0;
                }
                conv = wave_canon<true>(x, N) == 0.0;
            }
#pragma unroll
            for (int q = 0; q < kWaveSlots; ++q) {
                const int i = q * 64 + l;
                if (i < N) L.Fl[i] = e[q];
            }
        } else {
            double x[kWaveSlots];
#pragma unroll
            for (int q = 0; q < kWaveSlots; ++q) x[q] = fabs(e[q]);
            me = wave_canon<true>(x, N);
        }
        if (l == 0) {
            L.red[0] = me;
            L.red[1] = conv ? 1.0 : 0.0;
        }
    }
    __syncthreads();
    emax = L.red[0];
    const bool conv = L.red[1] != 0.0;
    __syncthreads();
    return conv;
}

// D[t,1:4] of one step from the field in Eld and the chunk partials in dp
// (src/GaussianFixedPointQuiet.jl:11,13), as k_diag_final writes it.
__device__ __forceinline__ void small_diag_row(const SmallLds& L, const InstDesc& I, int64_t nchunks,
                                               double* __restrict__ row4, int32_t* __restrict__ sweeps,
                                               int32_t nsweeps) {
    double se, sv, sp;
    diag_sums(L.red, L.Eld, I.N, L.dp, nchunks, se, sv, sp);
    if (threadIdx.x == 0) {
        const double d1 = (se / (double)I.N) / 2.0;
        const double d2 = ((sv * I.W) / (double)I.P) / 2.0;
        const double d3 = d1 + d2;
        const double s = 2.0 / I.W;
        row4[0] = d1 * s;
        row4[1] = d2 * s;
        row4[2] = d3 * s;
        row4[3] = sp;
        *sweeps = nsweeps;
    }
}

// ---------------------------------------------------------------------------
// NGP leapfrog (src/NGPFourier.jl:4-6), one workgroup, J particles per thread.
// ---------------------------------------------------------------------------
template <int JMAX>
__global__ void __launch_bounds__(kSmallThreads)
k_small_ngp(SmallRun a) {
    extern __shared__ double sm[];
    const InstDesc I = a.I;
    InstState S = *a.st;
    const int N = I.N, log2N = I.log2N;
    const SmallLds L = small_lds(sm, N);
    const int t = threadIdx.x;
    const int64_t P = I.P;
    const double Nd = (double)N, dt = I.dt, Pd = (double)P;
    const int32_t nmask = I.nmask;
    const int par0 = S.step & 1;
    const double* X = par0 ? a.X1 : a.X0;
    const double* V = par0 ? a.V1 : a.V0;
    // particle t + 512 j
    double x[JMAX], v[JMAX];
    bool has[JMAX];
#pragma unroll
    for (int j = 0; j < JMAX; ++j) {
        const int64_t p = t + (int64_t)kSmallThreads * j;
        has[j] = p < P;
        x[j] = has[j] ? X[p] : 0.0;
        v[j] = has[j] ? V[p] : 0.0;
    }
    stage_tables(L, a.tw, N);
    const u128 one = fxp_from(I.w, I.q);
    const int64_t nch = (P + kTreeChunk - 1) / kTreeChunk;
    constexpr int kPerChunk = kTreeChunk / kSmallThreads;  // slots j of one diagnostics chunk
    bool bad = false;
    for (int64_t s = 0; s < a.nsteps; ++s) {
        for (int c = t; c < N; c += kSmallThreads) L.acc[c] = 0;
        __syncthreads();
        // u(); n[f(j)] += w (the count; times w exactly below)
#pragma unroll
        for (int j = 0; j < JMAX; ++j)
            if (has[j]) {
                x[j] = jl_mod1f(x[j] + (v[j] / 2.0) * dt);
                atomicAdd(&L.acc[ngp_cell(x[j], Nd, nmask)], 1ull);
            }
        __syncthreads();
        for (int c = t; c < N; c += kSmallThreads) {
            const int b = (int)brev((uint32_t)c, log2N);
            L.re[b] = fxp_to((u128)L.acc[c] * one, I.q);
            L.im[b] = 0.0;
        }
        __syncthreads();
        double emax;
        small_solve<false>(L, N, log2N, 0.0, emax);
        bad = bad || !isfinite(emax);
        // u(); v += E[f.(x)]*dt
#pragma unroll
        for (int j = 0; j < JMAX; ++j)
            if (has[j]) {
                x[j] = jl_mod1f(x[j] + (v[j] / 2.0) * dt);
                v[j] = v[j] + L.Eld[ngp_cell(x[j], Nd, nmask)] * dt;
            }
        // chunk partials of v^2 and v/P.  Lane l < 256 of chunk c sums particles
        // 4096 c + l + 256 i, i = 0..15, in order: those of thread l (i even,
        // slot 8c + i/2) and of thread l + 256 (i odd), whose values come
        // through LDS.
        if (t >= kTreeLanes)
#pragma unroll
            for (int j = 0; j < JMAX; ++j) L.xchg[(t - kTreeLanes) * JMAX + j] = v[j];
        __syncthreads();
#pragma unroll
        for (int c = 0; c < JMAX / kPerChunk; ++c) {
            if (c < nch) {
                double s2 = 0.0, s1 = 0.0, z = 0.0;
                if (t < kTreeLanes) {
#pragma unroll
                    for (int i = 0; i < kTreeChunk / kTreeLanes; ++i) {
                        const int j = kPerChunk * c + i / 2;
                        const int64_t p = (int64_t)kTreeChunk * c + t + (int64_t)kTreeLanes * i;
                        if (p < P) {
                            const double vv = (i & 1) ? L.xchg[t * JMAX + j] : v[j];
                            s2 = s2 + vv * vv;
                            s1 = s1 + vv / Pd;
                        }
                    }
                }
                tree256x3<false>(L.red, s2, s1, z);
                if (t == 0) {
                    L.dp[2 * c] = s2;
                    L.dp[2 * c + 1] = s1;
                }
            }
        }
        __syncthreads();
        small_diag_row(L, I, nch, a.diag + 4 * s, a.sweeps + s, 1);
        __syncthreads();
    }
    const int parN = (int)((S.step + a.nsteps) & 1);
    double* Xo = parN ? a.X1 : a.X0;
    double* Vo = parN ? a.V1 : a.V0;
#pragma unroll
    for (int j = 0; j < JMAX; ++j) {
        const int64_t p = t + (int64_t)kSmallThreads * j;
        if (has[j]) {
            Xo[p] = x[j];
            Vo[p] = v[j];
        }
    }
    for (int c = t; c < N; c += kSmallThreads) a.E0[c] = L.Eld[c];
    if (t == 0) {
        S.step += (int32_t)a.nsteps;
        S.e_cur = 0;
        S.last_sweep = 0;
        if (bad) S.err |= 1;
        *a.st = S;
    }
}

// ---------------------------------------------------------------------------
// Grid barrier of k_small_gauss: a counter and a generation word, agent-scope
// release/acquire (the pattern of a cooperative grid sync).  Every wait is
// bounded: past spin_limit polls, or once another workgroup has given up, the
// barrier reports failure and the caller leaves the kernel.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool grid_barrier(uint32_t* bar, uint32_t nblocks, uint32_t& gen, uint32_t spin_limit,
                                             int32_t* okflag) {
    __syncthreads();
    if (threadIdx.x == 0) {
        int32_t ok = 1;
        __threadfence();
        const uint32_t arrived = __hip_atomic_fetch_add(&bar[0], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (arrived == nblocks - 1) {
            __hip_atomic_store(&bar[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&bar[1], gen + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            uint32_t n = 0;
            while (__hip_atomic_load(&bar[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == gen) {
                if (++n > spin_limit ||
                    __hip_atomic_load(&bar[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
                    ok = 0;
                    __hip_atomic_store(&bar[2], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        __threadfence();
        *okflag = ok;
    }
    ++gen;
    __syncthreads();
    return *okflag != 0;
}

__device__ __forceinline__ u64 ld_agent(const u64* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(u64* p, u64 v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Edge e of the stencil of centre c: f((i - 1/2)/N, c) with i = i0 - R + e
// (src/GaussianFixedPointQuiet.jl:5); lane e of a particle's 16 holds it, and
// weight e = f_{e+1} - f_e comes from the next lane.
template <int R>
__device__ __forceinline__ double seg_weight(int64_t i0, int e, double c, double Nd) {
    const double f = edge_half_erf(i0 - R + e, c, Nd);
    const double fn = __shfl_down(f, 1, kSegLanes);
    return fn - f;
}

// sum(k -> E[k[1]]*k[2], d(c)): the products on lanes 0..2R, folded left in
// cell order by every lane of the particle.
template <int R>
__device__ __forceinline__ double seg_gather(const double* __restrict__ Eld, double c, double Nd, int32_t nmask,
                                             int e) {
    const int64_t i0 = (int64_t)rint(c * Nd);
    const double wt = seg_weight<R>(i0, e, c, Nd);
    const double term = Eld[cell0(i0 - R + e, nmask)] * wt;
    double S = __shfl(term, 0, kSegLanes);
#pragma unroll
    for (int q = 1; q <= 2 * R; ++q) S = S + __shfl(term, q, kSegLanes);
    return S;
}

// r[k[1]] += k[2]*w: lane e <= 2R adds cell i0 - R + e in exact limbs.
template <int R>
__device__ __forceinline__ void seg_deposit(u64* __restrict__ acc, int32_t N, double c, double Nd, int32_t nmask,
                                            double w, int q, int e) {
    const int64_t i0 = (int64_t)rint(c * Nd);
    const double wt = seg_weight<R>(i0, e, c, Nd);
    const double a = wt * w;
    if (e <= 2 * R && a != 0.0) lds_add_fxp(acc, N, cell0(i0 - R + e, nmask), fxp_from(a, q));
}

// Adds this workgroup's LDS limbs into the device accumulator (re-split so
// that the device-scope adds never carry, as flush_limbs does).
__device__ __forceinline__ void small_flush(const u64* __restrict__ acc, int N, u64* __restrict__ g) {
    for (int c = threadIdx.x; c < N; c += blockDim.x) {
        const u128 v = limbs_to_u128(acc[c], acc[N + c], acc[2 * N + c]);
        const u64 l0 = (u64)v & kLimbMask, l1 = (u64)(v >> 43) & kLimbMask, l2 = (u64)(v >> 86);
        if (l0) __hip_atomic_fetch_add(g + 3 * c, l0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (l1) __hip_atomic_fetch_add(g + 3 * c + 1, l1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (l2) __hip_atomic_fetch_add(g + 3 * c + 2, l2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// ---------------------------------------------------------------------------
// Gaussian stencils (R = 6, 7): SCHEME 0 leapfrog (src/Gaussian.jl:8-10),
// SCHEME 1 implicit midpoint with the Picard loop
// (src/GaussianFixedPointQuiet.jl:8-10).  Workgroups 0..K-1 hold particles
// (particle p on segment p mod 32K, slot p / 32K); workgroup K computes the
// diagnostics of each step from the velocities the others store.
// ---------------------------------------------------------------------------

// Per-workgroup state of k_small_gauss that is not per particle.
struct GaussCtl {
    int32_t* okflag;  // LDS: the barrier's verdict for the workgroup
    uint32_t nblocks, gen;
    int64_t nb;       // barriers passed: accumulator nb % 3 is filled before barrier nb
    bool diagwg;
};

// Deposit of every held particle's centre (the caller's positions), the
// flush into device accumulator nb % 3, the barrier, and the summed charge
// staged bit-reversed for the solve.  Returns false if the barrier failed.
template <int R, int JMAX>
__device__ __forceinline__ bool small_deposit_and_meet(const SmallRun& a, const SmallLds& L, GaussCtl& C,
                                                       const double (&cen)[JMAX], const bool (&has)[JMAX], int e) {
    const InstDesc& I = a.I;
    const int N = I.N;
    if (!C.diagwg) {
        for (int c = threadIdx.x; c < 3 * N; c += blockDim.x) L.acc[c] = 0;
        __syncthreads();
#pragma unroll
        for (int j = 0; j < JMAX; ++j)
            if (has[j]) seg_deposit<R>(L.acc, N, cen[j], (double)N, I.nmask, I.w, I.q, e);
        __syncthreads();
        small_flush(L.acc, N, a.gacc + 3 * (int64_t)N * (C.nb % 3));
    }
    if (!grid_barrier(a.bar, C.nblocks, C.gen, a.spin_limit, C.okflag)) return false;
    // the accumulator read two barriers ago is free again
    if (blockIdx.x == 0)
        for (int c = threadIdx.x; c < 3 * N; c += blockDim.x) st_agent(a.gacc + 3 * (int64_t)N * ((C.nb + 2) % 3) + c, 0);
    const u64* g = a.gacc + 3 * (int64_t)N * (C.nb % 3);
    for (int c = threadIdx.x; c < N; c += blockDim.x) {
        const int b = (int)brev((uint32_t)c, I.log2N);
        L.re[b] = fxp_to(limbs_to_u128(ld_agent(g + 3 * c), ld_agent(g + 3 * c + 1), ld_agent(g + 3 * c + 2)), I.q);
        L.im[b] = 0.0;
    }
    __syncthreads();
    ++C.nb;
    return true;
}

// Diagnostics workgroup: the row of step `row`, which ended before the last
// barrier, from the velocities stored then (its field still in Eld).
__device__ __forceinline__ void small_gauss_diag(const SmallRun& a, const SmallLds& L, int64_t row, int32_t nsweeps,
                                                 int32_t step0) {
    const InstDesc& I = a.I;
    const int64_t P = I.P;
    const int64_t nch = (P + kTreeChunk - 1) / kTreeChunk;
    const int parw = (int)((step0 + row + 1) & 1);
    const u64* vg = reinterpret_cast<const u64*>(parw ? a.V1 : a.V0);
    for (int64_t c = 0; c < nch; ++c) {
        const int64_t n = (P - c * kTreeChunk) < kTreeChunk ? (P - c * kTreeChunk) : kTreeChunk;
        const u64* vc = vg + c * kTreeChunk;
        double s2, s1;
        chunk_sums_v(L.red, n, (double)P, [&](int64_t i) { return __longlong_as_double((long long)ld_agent(vc + i)); },
                     s2, s1);
        if (threadIdx.x == 0) {
            L.dp[2 * c] = s2;
            L.dp[2 * c + 1] = s1;
        }
    }
    __syncthreads();
    small_diag_row(L, I, nch, a.diag + 4 * row, a.sweeps + row, nsweeps);
    __syncthreads();
}

template <int R, int SCHEME>
__global__ void __launch_bounds__(kSmallThreads)
k_small_gauss(SmallRun a) {
    extern __shared__ double sm[];
    __shared__ int32_t okflag;
    const InstDesc I = a.I;
    const InstState S0 = *a.st;
    const int N = I.N, log2N = I.log2N;
    const SmallLds L = small_lds(sm, N);
    const int K = a.K;
    GaussCtl C;
    C.okflag = &okflag;
    C.nblocks = (uint32_t)K + 1;
    C.gen = 0;
    C.nb = 0;
    C.diagwg = (int)blockIdx.x == K;
    const int e = threadIdx.x & (kSegLanes - 1);
    const int64_t seg = (int64_t)blockIdx.x * kSegs + (threadIdx.x >> 4);
    const int64_t G = (int64_t)K * kSegs;
    const int64_t P = I.P;
    const double Nd = (double)N, dt = I.dt;
    const int32_t nmask = I.nmask;
    constexpr int JMAX = kSmallMaxSlots;
    // particle state: X, V of the step and the Picard iterate v_k
    double Xp[JMAX], Vp[JMAX], vk[JMAX], cen[JMAX];
    bool has[JMAX];
    {
        const int par = S0.step & 1;
        const double* X = par ? a.X1 : a.X0;
        const double* V = par ? a.V1 : a.V0;
#pragma unroll
        for (int j = 0; j < JMAX; ++j) {
            const int64_t p = seg + G * j;
            has[j] = !C.diagwg && p < P;
            Xp[j] = has[j] ? X[p] : 0.0;
            Vp[j] = has[j] ? V[p] : 0.0;
            vk[j] = Vp[j];
            cen[j] = 0.0;
        }
    }
    {
        const double* Ein = S0.e_cur ? a.E1 : a.E0;
        for (int j = threadIdx.x; j < N; j += blockDim.x) {
            L.Fl[j] = Ein[j];
            L.Eld[j] = Ein[j];
        }
    }
    stage_tables(L, a.tw, N);
    bool bad = false, ok = true;
    int64_t pend_row = -1;  // diagnostics workgroup: step whose row is still to be written
    int32_t pend_sweeps = 0;

    for (int64_t s = 0; s < a.nsteps && ok; ++s) {
        const int paro = (int)((S0.step + s + 1) & 1);
        u64* Xo = reinterpret_cast<u64*>(paro ? a.X1 : a.X0);
        u64* Vo = reinterpret_cast<u64*>(paro ? a.V1 : a.V0);
        if (SCHEME == 0) {
            // u(); deposit at x (centre (x + x)/2)
#pragma unroll
            for (int j = 0; j < JMAX; ++j) {
                Xp[j] = has[j] ? jl_mod1f(Xp[j] + (Vp[j] / 2.0) * dt) : 0.0;
                cen[j] = (Xp[j] + Xp[j]) / 2.0;
            }
            if (!(ok = small_deposit_and_meet<R, JMAX>(a, L, C, cen, has, e))) break;
            if (C.diagwg && pend_row >= 0) small_gauss_diag(a, L, pend_row, pend_sweeps, S0.step);
            double emax;
            small_solve<false>(L, N, log2N, 0.0, emax);
            bad = bad || !isfinite(emax);
            // u(); v += gather(E, x)*dt
#pragma unroll
            for (int j = 0; j < JMAX; ++j)
                if (has[j]) {
                    Xp[j] = jl_mod1f(Xp[j] + (Vp[j] / 2.0) * dt);
                    Vp[j] = Vp[j] + seg_gather<R>(L.Eld, Xp[j], Nd, nmask, e) * dt;
                    if (e == 0) {
                        st_agent(Xo + seg + G * j, (u64)__double_as_longlong(Xp[j]));
                        st_agent(Vo + seg + G * j, (u64)__double_as_longlong(Vp[j]));
                    }
                }
            pend_row = s;
            pend_sweeps = 1;
        } else {
            // X .= x; V .= v; the loop's first iterate is v_0 = V:
            // x_0 = X + ((V + V)/2)*dt, centre (x_0 + X)/2
#pragma unroll
            for (int j = 0; j < JMAX; ++j) {
                vk[j] = Vp[j];
                const double x = Xp[j] + ((vk[j] + Vp[j]) / 2.0) * dt;
                cen[j] = (x + Xp[j]) / 2.0;
            }
            if (!(ok = small_deposit_and_meet<R, JMAX>(a, L, C, cen, has, e))) break;
            if (C.diagwg && pend_row >= 0) small_gauss_diag(a, L, pend_row, pend_sweeps, S0.step);
            for (int k = 0;; ++k) {
                double emax;
                const bool conv = small_solve<true>(L, N, log2N, I.rtol, emax);
                bad = bad || !isfinite(emax);
                const bool done = conv || k == I.max_sweeps - 1;
                // v_{k+1} = V + gather(E_k, c_k)*dt; unless the loop ends here
                // the next centre c_{k+1} = (x_{k+1} + X)/2
#pragma unroll
                for (int j = 0; j < JMAX; ++j)
                    if (has[j]) {
                        const double x = Xp[j] + ((vk[j] + Vp[j]) / 2.0) * dt;
                        const double vn = Vp[j] + seg_gather<R>(L.Eld, (x + Xp[j]) / 2.0, Nd, nmask, e) * dt;
                        if (done) {
                            Xp[j] = jl_mod1f(x);
                            Vp[j] = vn;
                            if (e == 0) {
                                st_agent(Xo + seg + G * j, (u64)__double_as_longlong(Xp[j]));
                                st_agent(Vo + seg + G * j, (u64)__double_as_longlong(Vp[j]));
                            }
                        } else {
                            vk[j] = vn;
                            const double x2 = Xp[j] + ((vn + Vp[j]) / 2.0) * dt;
                            cen[j] = (x2 + Xp[j]) / 2.0;
                        }
                    }
                if (done) {
                    pend_row = s;
                    pend_sweeps = k + 1;
                    break;
                }
                if (!(ok = small_deposit_and_meet<R, JMAX>(a, L, C, cen, has, e))) break;
            }
        }
    }
    if (ok) ok = grid_barrier(a.bar, C.nblocks, C.gen, a.spin_limit, C.okflag);
    if (ok && C.diagwg && pend_row >= 0) small_gauss_diag(a, L, pend_row, pend_sweeps, S0.step);
    if (blockIdx.x == 0 && ok) {
        double* Eo = S0.e_cur ? a.E1 : a.E0;
        for (int j = threadIdx.x; j < N; j += blockDim.x) Eo[j] = L.Eld[j];
        if (threadIdx.x == 0) {
            InstState S = S0;
            S.step += (int32_t)a.nsteps;
            S.last_sweep = -1;
            if (bad) S.err |= 1;
            *a.st = S;
        }
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
bool small_eligible(const InstDesc& I, int scheme, int precision, int64_t* K_out) {
    if (precision != 0 || I.N > kSmallMaxN) return false;
    if (I.R == 0) {
        if (scheme != 0 || I.P > (int64_t)kSmallThreads * kSmallNgpSlots) return false;
        *K_out = 1;
        return true;
    }
    if ((I.R != 6 && I.R != 7) || (scheme != 0 && scheme != 1)) return false;
    // one slot per segment while the particles fill at most kSmallMaxGroups
    // workgroups, then more slots per segment
    int64_t K = (I.P + kSegs - 1) / kSegs;
    if (K > kSmallMaxGroups) K = kSmallMaxGroups;
    if ((I.P + K * kSegs - 1) / (K * kSegs) > kSmallMaxSlots) return false;
    *K_out = K;
    return true;
}

int small_workgroups_ok(const InstDesc& I, int64_t K) {
    if (I.R == 0) return K == 1;
    return K >= 1 && K <= kSmallMaxGroups && (I.P + K * kSegs - 1) / (K * kSegs) <= kSmallMaxSlots;
}

hipError_t launch_small_run(const SmallRun& a, int scheme, hipStream_t s) {
    const int lds = small_lds_bytes(a.I.N, a.I.R == 0);
    SmallRun arg = a;
    void* args[] = {&arg};
    if (a.I.R == 0) {
        k_small_ngp<kSmallNgpSlots><<<1, kSmallThreads, lds, s>>>(a);
        return hipGetLastError();
    }
    const void* f;
    if (a.I.R == 6) f = scheme ? reinterpret_cast<const void*>(k_small_gauss<6, 1>)
                               : reinterpret_cast<const void*>(k_small_gauss<6, 0>);
    else f = scheme ? reinterpret_cast<const void*>(k_small_gauss<7, 1>)
                    : reinterpret_cast<const void*>(k_small_gauss<7, 0>);
    return hipLaunchCooperativeKernel(f, dim3((unsigned)a.K + 1), dim3(kSmallThreads), args, (unsigned)lds, s);
}

int configure_kernels_small() {
    const int bytes = small_lds_bytes(kSmallMaxN, true);
    hipError_t e = hipSuccess;
    auto chk = [&](hipError_t r) { if (r != hipSuccess) e = r; };
    chk(hipFuncSetAttribute(reinterpret_cast<const void*>(k_small_ngp<kSmallNgpSlots>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
    chk(hipFuncSetAttribute(reinterpret_cast<const void*>(k_small_gauss<6, 0>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
    chk(hipFuncSetAttribute(reinterpret_cast<const void*>(k_small_gauss<6, 1>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
    chk(hipFuncSetAttribute(reinterpret_cast<const void*>(k_small_gauss<7, 0>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
    chk(hipFuncSetAttribute(reinterpret_cast<const void*>(k_small_gauss<7, 1>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
    (void)hipGetLastError();
    return e == hipSuccess ? 0 : -1;
}

}  // namespace pic
