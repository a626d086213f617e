// solve1d.h -- the LDS-resident spectral solve of the 1D kernels and the
// block reductions of the convergence test (one block of kSolveThreads
// threads per instance).  Compiled with -ffp-contract=off by every includer.
#pragma once
#include "pic_device.h"
#include "pic_internal.h"

#pragma clang fp contract(off)

namespace pic {
// ---------------------------------------------------------------------------
// Spectral solve in LDS: E = real(ifft((xi = fft(rho) ./ ik; xi[1] *= 0; xi)))
// with the radix-2 decimation-in-time transform of the CPU restatement.
// Stages are applied up to three at a time: a group of 2^M elements spaced
// 2^s0 apart is closed under stages s0+1..s0+M, so one thread loads it into
// registers and applies exactly the radix-2 butterflies (same operands, same
// twiddles per element) before writing back.  The forward transform runs its
// log2N % 3 remainder stages first and the inverse runs them last, so the
// forward's final pass and the inverse's first pass are both three stages
// wide: the eight spectrum values m = g + j 2^(L-3) a thread ends the forward
// transform with are, bit-reversed, exactly the contiguous group
// 8 brev(g) + brev3(j) of the inverse's first pass, and the division by ik
// happens between the two in registers.  A 4096-point solve is seven LDS
// passes; the last one writes only the real parts, times 1/N.
// ---------------------------------------------------------------------------

// Stage-s twiddles (s = 1..12) W_{2^s}^k, k < 2^(s-1), sit at table entry
// 2^(s-1) - 1 + k; local stage t of a pass starting at stage s0+1 uses
// k = jlo 2^s0 + lo for jlo < 2^t.
__device__ __forceinline__ int stage_tw(int s0, int t, int jlo, int lo) {
    return (1 << (s0 + t)) - 1 + (jlo << s0) + lo;
}

// Twiddles of a pass of M (runtime, <= 3) stages for group offset lo.
__device__ __forceinline__ void load_pass_tw(double2 (&w)[7], const double2* __restrict__ tw, int s0, int M,
                                             int lo) {
#pragma unroll
    for (int t = 0; t < kPassStages; ++t)
#pragma unroll
        for (int jlo = 0; jlo < (1 << t); ++jlo)
            if (t < M) w[(1 << t) - 1 + jlo] = tw[stage_tw(s0, t, jlo, lo)];
}

// Radix-2 butterflies of local stages 0..M-1 of a group held in registers.
template <int M>
__device__ __forceinline__ void group_butterflies(double* ar, double* ai, const double2 (&w)[7], bool inverse) {
    constexpr int G = 1 << M;
#pragma unroll
    for (int t = 0; t < M; ++t) {
#pragma unroll
        for (int j1 = 0; j1 < G; ++j1) {
            if (j1 & (1 << t)) continue;
            const int j2 = j1 | (1 << t);
            const double2 wv = w[(1 << t) - 1 + (j1 & ((1 << t) - 1))];
            const double wr = wv.x;
            const double wi = inverse ? -wv.y : wv.y;
            const double br = ar[j2], bi = ai[j2];
            const double tr = br * wr - bi * wi;
            const double tq = br * wi + bi * wr;
            const double ur = ar[j1], ui = ai[j1];
            ar[j1] = ur + tr;
            ai[j1] = ui + tq;
            ar[j2] = ur - tr;
            ai[j2] = ui - tq;
        }
    }
}

// LDS layouts of the transform: the forward phase pads one double every 16
// (padi); from the fused pass on, the inverse phase also pads one every 64
// (padj), so that the fused pass's bit-reversed group stores (stride 64
// doubles across a wave) spread over the banks like the strided passes do.
__device__ __forceinline__ int padj(int i) { return i + (i >> 4) + (i >> 6); }
// Doubles per padded re/im array (either layout).
__host__ __device__ constexpr int solve_np(int N) { return N + N / 16 + N / 64; }
// Element j of the solve's result (in the inverse layout).
__device__ __forceinline__ double solve_at(const double* re, int j) { return re[padj(j)]; }

// One LDS pass of M stages from stage s0+1; thread g owns group g (every pass
// has N >> M <= 512 = kSolveThreads groups).  INV: the inverse-phase layout.
// With scale_out the pass writes only the real parts, times `scale`.
template <int M, bool INV>
__device__ __forceinline__ void lds_pass(double* re, double* im, const double2 (&w)[7], int N, int s0,
                                         bool inverse, bool scale_out, double scale) {
    constexpr int G = 1 << M;
    const int g = threadIdx.x;
    if (g < (N >> M)) {
        const int lo = g & ((1 << s0) - 1);
        const int base = ((g >> s0) << (s0 + M)) + lo;
        double ar[G], ai[G];
#pragma unroll
        for (int j = 0; j < G; ++j) {
            const int id = INV ? padj(base + (j << s0)) : padi(base + (j << s0));
            ar[j] = re[id];
            ai[j] = im[id];
        }
        group_butterflies<M>(ar, ai, w, inverse);
#pragma unroll
        for (int j = 0; j < G; ++j) {
            const int id = INV ? padj(base + (j << s0)) : padi(base + (j << s0));
            if (scale_out) {
                re[id] = ar[j] * scale;
            } else {
                re[id] = ar[j];
                im[id] = ai[j];
            }
        }
    }
    __syncthreads();
}

template <bool INV>
__device__ __forceinline__ void lds_pass_m(int M, double* re, double* im, const double2 (&w)[7], int N, int s0,
                                           bool inverse, bool scale_out, double scale) {
    if (M == 3) lds_pass<3, INV>(re, im, w, N, s0, inverse, scale_out, scale);
    else if (M == 2) lds_pass<2, INV>(re, im, w, N, s0, inverse, scale_out, scale);
    else lds_pass<1, INV>(re, im, w, N, s0, inverse, scale_out, scale);
}

// Pass plans for L = log2 N >= 3, r = L % 3, np = ceil(L / 3) passes each:
// forward pass p starts at stage fwd_s0(p) + 1 and is r stages wide if it is
// the first and r > 0, else 3; inverse pass q starts at 3q + 1 and is r wide
// if it is the last and r > 0, else 3.
__device__ __forceinline__ int fwd_s0(int p, int r) { return p == 0 ? 0 : (r ? r + 3 * (p - 1) : 3 * p); }

// Register operands of one thread's share of the solve.
struct SolveRegs {
    double2 w[4][7];  // forward pass twiddles; after the forward transform, the inverse's
    double2 w0[7];    // the inverse's first pass (group offset 0: the same for every thread)
    double xt[8];     // 1/(2 pi |m'|) of the thread's eight modes
};

// Issues the solve's table loads (call before the charge loads so that their
// latency overlaps).  Requires blockDim.x == kSolveThreads.  xi: the
// 1/(2 pi k) factors, k = 0..N/2 (NULL: the ones after the twiddle table).
__device__ __forceinline__ void solve_prefetch(SolveRegs& R, const double2* __restrict__ tw, int N, int L,
                                               const double* __restrict__ xi_in = nullptr) {
    const int r = L % 3, np = (L + 2) / 3;
#pragma unroll
    for (int p = 0; p < 4; ++p)
        if (p < np) {
            const int s0 = fwd_s0(p, r);
            load_pass_tw(R.w[p], tw, s0, (p == 0 && r) ? r : 3, threadIdx.x & ((1 << s0) - 1));
        }
    load_pass_tw(R.w0, tw, 0, 3, 0);
    const double* xi = xi_in ? xi_in : reinterpret_cast<const double*>(tw) + kXiOffset;
    const int g = threadIdx.x, s = L - 3;
    if (g < (N >> 3)) {
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int m = g + (j << s);
            R.xt[j] = xi[m == 0 ? 1 : (m <= N / 2 ? m : N - m)];
        }
    }
}

// LDS doubles used by the solve: padded re/im, previous field, reduction scratch
__host__ __device__ constexpr int solve_lds_doubles(int N) { return 2 * solve_np(N) + N + kRedDoubles; }

// On entry re/im hold rho in bit-reversed order (padi layout, arrays of
// solve_np(N) doubles); on exit solve_at(re, j) is E[j].
// Requires blockDim.x == kSolveThreads and R from solve_prefetch.
__device__ __forceinline__ void solve_lds(double* re, double* im, SolveRegs& R, const double2* __restrict__ tw,
                                          int N, int L) {
    const int r = L % 3, np = (L + 2) / 3;
    const double scale = 1.0 / (double)N;
#pragma unroll
    for (int p = 0; p < 3; ++p)
        if (p < np - 1)
            lds_pass_m<false>((p == 0 && r) ? r : 3, re, im, R.w[p], N, fwd_s0(p, r), false, false, 0.0);
    // the forward transform's last pass, xi = rho^ ./ ik, the inverse's first pass
    const int g = threadIdx.x, s = L - 3;
    const bool own = g < (N >> 3);
    double br[8], bi[8];
    if (own) {
        double ar[8], ai[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int id = padi(g + (j << s));
            ar[j] = re[id];
            ai[j] = im[id];
        }
        double2 wl[7];
#pragma unroll
        for (int p = 0; p < 4; ++p)
            if (p == np - 1)
#pragma unroll
                for (int e = 0; e < 7; ++e) wl[e] = R.w[p][e];
        group_butterflies<3>(ar, ai, wl, false);
        // ik = 2 pi i [1, 1:N/2, -N/2+1:-1]; Julia's complex division by (0, k)
        // is (Im * (1/k), -(Re * (1/k))); xi[1] *= 0.
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int m = g + (j << s);
            const double t = m > N / 2 ? -R.xt[j] : R.xt[j];
            double a = ai[j] * t;
            double b = -(ar[j] * t);
            if (m == 0) {
                a = a * 0.0;
                b = b * 0.0;
            }
            const int q = ((j & 1) << 2) | (j & 2) | ((j >> 2) & 1);  // brev3(j)
            br[q] = a;
            bi[q] = b;
        }
        group_butterflies<3>(br, bi, R.w0, true);
    }
    // the inverse's pass twiddles (the same as the forward's when r == 0)
    if (r)
#pragma unroll
        for (int q = 1; q < 4; ++q)
            if (q < np) load_pass_tw(R.w[q], tw, 3 * q, q == np - 1 ? r : 3, threadIdx.x & ((1 << (3 * q)) - 1));
    __syncthreads();
    if (own) {
        const int base = (s ? (int)brev((uint32_t)g, s) : 0) << 3;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int id = padj(base + q);
            if (np == 1) {
                re[id] = br[q] * scale;
            } else {
                re[id] = br[q];
                im[id] = bi[q];
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 1; q < 4; ++q)
        if (q < np) {
            const bool last = q == np - 1;
            lds_pass_m<true>(last && r ? r : 3, re, im, R.w[q], N, 3 * q, true, last, scale);
        }
}

// Six canonical trees at once: NaN-propagating maxima of m[0..2] and sums of
// s[0..2], with the operands of tree256; red holds >= 6*256 + 6 doubles.
__device__ __forceinline__ void tree256_max3_sum3(double* red, double (&m)[3], double (&s)[3]) {
    const int t = threadIdx.x;
    if (t < kTreeLanes) {
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            red[q * kTreeLanes + t] = m[q];
            red[(3 + q) * kTreeLanes + t] = s[q];
        }
    }
    __syncthreads();
    if (t < 128) {
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            red[q * kTreeLanes + t] = nanmax(red[q * kTreeLanes + t], red[q * kTreeLanes + t + 128]);
            red[(3 + q) * kTreeLanes + t] = red[(3 + q) * kTreeLanes + t] + red[(3 + q) * kTreeLanes + t + 128];
        }
    }
    __syncthreads();
    if (t < 64) {
        double a[6];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            a[q] = nanmax(red[q * kTreeLanes + t], red[q * kTreeLanes + t + 64]);
            a[3 + q] = red[(3 + q) * kTreeLanes + t] + red[(3 + q) * kTreeLanes + t + 64];
        }
#pragma unroll
        for (int h = 32; h > 0; h >>= 1)
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                a[q] = nanmax(a[q], __shfl_down(a[q], h, 64));
                a[3 + q] = a[3 + q] + __shfl_down(a[3 + q], h, 64);
            }
        if (t == 0)
#pragma unroll
            for (int q = 0; q < 6; ++q) red[6 * kTreeLanes + q] = a[q];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        m[q] = red[6 * kTreeLanes + q];
        s[q] = red[6 * kTreeLanes + 3 + q];
    }
    __syncthreads();
}

// Julia's norm of three vectors (d = F - E, F, E) from their max-abs values:
// m == 0 or Inf -> m; n*m^2 finite and m^2 != 0 -> sqrt(sum x^2); otherwise
// m*sqrt(sum (|x|/m)^2).  Sums use the canonical tree.  One pass forms the
// maxima and the plain sums together (for n <= 4096 with all of a lane's
// loads issued before its sums); a second pass runs only for a vector that
// needs the scaled sum.
template <typename FF, typename FE>
__device__ void block_norms3(double* red, int n, FF&& F, FE&& E, double& nd, double& nf, double& ne, double& me) {
    const int t = threadIdx.x;
    double m[3] = {0.0, 0.0, 0.0}, s[3] = {0.0, 0.0, 0.0};
    auto add = [&](double f, double e) {
        const double x[3] = {f - e, f, e};
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            m[q] = nanmax(m[q], fabs(x[q]));
            s[q] = s[q] + x[q] * x[q];
        }
    };
    if (t < kTreeLanes) {
        if (n <= kTreeChunk) {
            constexpr int K = kTreeChunk / kTreeLanes;
            double xf[K], xe[K];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const int i = t + k * kTreeLanes;
                xf[k] = i < n ? F(i) : 0.0;
                xe[k] = i < n ? E(i) : 0.0;
            }
            // a value past n is +0.0, which leaves every maximum (>= +0 or
            // NaN) and every sum of squares (>= +0, Inf or NaN) as it is
#pragma unroll
            for (int k = 0; k < K; ++k) add(xf[k], xe[k]);
        } else {
            for (int i = t; i < n; i += kTreeLanes) add(F(i), E(i));
        }
    }
    tree256_max3_sum3(red, m, s);
    bool direct[3], scaled[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        direct[q] = m[q] == 0.0 || isinf(m[q]);
        scaled[q] = !direct[q] && !(isfinite((double)n * m[q] * m[q]) && m[q] * m[q] != 0.0);
    }
    if (scaled[0] || scaled[1] || scaled[2]) {  // uniform across the block
        double z[3] = {0.0, 0.0, 0.0};
        if (t < kTreeLanes)
            for (int i = t; i < n; i += kTreeLanes) {
                const double f = F(i), e = E(i);
                const double x[3] = {f - e, f, e};
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    const double y = fabs(x[q]) / m[q];
                    z[q] = z[q] + y * y;
                }
            }
        tree256x3<false>(red, z[0], z[1], z[2]);
#pragma unroll
        for (int q = 0; q < 3; ++q)
            if (scaled[q]) s[q] = z[q];
    }
    nd = direct[0] ? m[0] : (scaled[0] ? m[0] * sqrt(s[0]) : sqrt(s[0]));
    nf = direct[1] ? m[1] : (scaled[1] ? m[1] * sqrt(s[1]) : sqrt(s[1]));
    ne = direct[2] ? m[2] : (scaled[2] ? m[2] * sqrt(s[2]) : sqrt(s[2]));
    me = m[2];
}

// isapprox(F, E; rtol, atol) on arrays (src/GaussianFixedPointQuiet.jl:8), as
// restated on the CPU.  All threads of the block call it; me returns max|E|.
template <typename FF, typename FE>
__device__ bool block_isapprox(double* red, FF&& F, FE&& E, int n, double rtol, double atol, double& me) {
    double dn, nf, ne;
    block_norms3(red, n, F, E, dn, nf, ne, me);
    if (isfinite(dn)) {
        double mx = nf > ne ? nf : ne;
        if (nf != nf || ne != ne) mx = __longlong_as_double(0x7ff8000000000000LL);
        const double tol = rtol * mx;
        double lim = atol > tol ? atol : tol;
        if (tol != tol) lim = tol;
        return dn <= lim;
    }
    double bad = 0.0;
    if (threadIdx.x < kTreeLanes)
        for (int i = threadIdx.x; i < n; i += kTreeLanes) {
            const double a = F(i), b = E(i);
            if (a == b) continue;
            bool good = isfinite(a) && isfinite(b);
            if (good) {
                const double ma = fabs(a) > fabs(b) ? fabs(a) : fabs(b);
                const double t = rtol * ma;
                good = fabs(a - b) <= (atol > t ? atol : t);
            }
            if (!good) bad = 1.0;
        }
    return tree256<true>(red, bad) == 0.0;
}

// One block of kSolveThreads threads per instance.  Reads the charge, solves,
// and (TEST) runs the midpoint's convergence test between the previous field
// (e_cur) and the new one -- the test the scripts run at the top of sweep
// k+1 -- ending the loop on success or at the last sweep.

// ---------------------------------------------------------------------------
// Diagnostics sums (src/GaussianFixedPointQuiet.jl:11,13) with the canonical
// tree: sum(E.^2) over the grid and the trees over the per-chunk partials of
// v^2 and v/P (chunks of 4096 particles).
// ---------------------------------------------------------------------------
// Canonical tree over n values spread with stride 2 (in place in dpart).
__device__ inline double tree_over_partials(double* red, double* vals, int64_t n) {
    while (n > kTreeChunk) {
        const int64_t C = (n + kTreeChunk - 1) / kTreeChunk;
        for (int64_t c = 0; c < C; ++c) {
            const int64_t m = n - c * kTreeChunk < kTreeChunk ? n - c * kTreeChunk : kTreeChunk;
            const double s = block_chunk_sum(red, m, [&](int64_t i) { return vals[2 * (c * kTreeChunk + i)]; });
            if (threadIdx.x == 0) vals[2 * c] = s;
            __syncthreads();
        }
        n = C;
    }
    return block_chunk_sum(red, n, [&](int64_t i) { return vals[2 * i]; });
}

// sum(E.^2) over N <= 4096 cells and the canonical tree over the chunk
// partials of v^2 and v/P; the common case (<= 4096 chunks) in one pass.
__device__ inline void diag_sums(double* red, const double* E, int N, double* dp, int64_t nchunks, double& se,
                          double& sv, double& sp) {
    if (nchunks > kTreeChunk) {
        se = block_chunk_sum(red, N, [&](int64_t i) { return E[i] * E[i]; });
        sv = tree_over_partials(red, dp, nchunks);
        sp = tree_over_partials(red, dp + 1, nchunks);
        return;
    }
    double a = 0.0, b = 0.0, c = 0.0;
    if (threadIdx.x < kTreeLanes) {
        constexpr int K = kTreeChunk / kTreeLanes;
        double e[K], x[K], y[K];
#pragma unroll
        for (int i = 0; i < K; ++i) {
            const int64_t j = threadIdx.x + (int64_t)i * kTreeLanes;
            e[i] = j < N ? E[j] : 0.0;
            x[i] = j < nchunks ? dp[2 * j] : 0.0;
            y[i] = j < nchunks ? dp[2 * j + 1] : 0.0;
        }
#pragma unroll
        for (int i = 0; i < K; ++i) {
            const int64_t j = threadIdx.x + (int64_t)i * kTreeLanes;
            if (j < N) a = a + e[i] * e[i];
            if (j < nchunks) {
                b = b + x[i];
                c = c + y[i];
            }
        }
    }
    tree256x3<false>(red, a, b, c);
    se = a;
    sv = b;
    sp = c;
}


}  // namespace pic
