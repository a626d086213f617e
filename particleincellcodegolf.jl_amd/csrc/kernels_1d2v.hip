// kernels_1d2v.hip -- FP64 gfx950 kernels of the magnetised 1D2V step
// (src/NGP1D2V.jl, src/NGP1D2V2S.jl): explicit Boris push with the Gaussian
// stencil, one or two species, and the time-binned field.  Compiled with
// -ffp-contract=off: results are bitwise equal to the CPU restatement.
//
// Per step: the solve reads the exact signed charge deposited by the previous
// push (or by the prime deposit after an upload), writes E, adds it into its
// time bin and sums E^2; the push gathers E at x, applies the Boris rotation,
// moves and wraps the particle and deposits its charge for the next step; the
// diagnostics (predicated on t % bin == 0) sum the kinetic terms and write a
// row of D, and advance the step counter.
#include "pic1d2v_internal.h"
#include "pic_device.h"
#include "solve1d.h"
#include "stencil1d.h"

#pragma clang fp contract(off)

namespace pic {

// r[k[1]] += q*k[2]*w over d(c) (src/NGP1D2V2S.jl:28): each contribution's
// magnitude on the 2^-qb grid into the LDS limbs of the species' sign.
template <int R>
__device__ __forceinline__ void deposit_gauss_q(u64* __restrict__ acc, int32_t N, double c, double Nd,
                                                int32_t nmask, double q, double w, int qb) {
    const int64_t i0 = (int64_t)rint(c * Nd);
    double prev = edge_half_erf(i0 - R, c, Nd);
#pragma unroll 1
    for (int j = -R; j <= R; ++j) {
        const double next = edge_half_erf(i0 + j + 1, c, Nd);
        const double wt = next - prev;
        prev = next;
        const double a = (q * wt) * w;
        if (a != 0.0) lds_add_fxp(acc, N, cell0(i0 + j, nmask), fxp_from(fabs(a), qb));
    }
}

// boris(vx, vy, E, B, dt, q_m) of src/NGP1D2V2S.jl:5-11 with every product of
// the SVector cross products made, zeros included; q_m = 1 is
// src/NGP1D2V.jl:5-10 (E*dt/2 == E*(dt/2)).
__device__ __forceinline__ void boris12(double& vx, double& vy, double E, double B, double dt, double q_m) {
    const double h = (dt / 2.0) * q_m;
    const double m1 = vx + E * h, m2 = vy, m3 = 0.0;
    const double t1 = 0.0, t2 = 0.0, t3 = B * h;
    const double c1 = m2 * t3 - m3 * t2, c2 = m3 * t1 - m1 * t3, c3 = m1 * t2 - m2 * t1;
    const double a1 = m1 + c1, a2 = m2 + c2, a3 = m3 + c3;
    const double d1 = a2 * t3 - a3 * t2, d2 = a3 * t1 - a1 * t3;
    const double s = 1.0 + ((t1 * t1 + t2 * t2) + t3 * t3);
    vx = (m1 + (2.0 * d1) / s) + E * h;
    vy = m2 + (2.0 * d2) / s;
}

// The charge of every tile's particles at their current positions.
template <int R>
__global__ void __launch_bounds__(1024)
k12_deposit(const Tile12* __restrict__ tiles, const double* __restrict__ x, u64* __restrict__ racc, int N,
            double q0, double q1, double w, int qb) {
    extern __shared__ u64 lds[];
    const Tile12 T = tiles[blockIdx.x];
    const double q = T.species ? q1 : q0;
    zero_lds(lds, 3 * N);
    __syncthreads();
    const double Nd = (double)N;
    for (int64_t p = threadIdx.x; p < T.count; p += blockDim.x)
        deposit_gauss_q<R>(lds, N, x[T.start + p], Nd, N - 1, q, w, qb);
    __syncthreads();
    flush_limbs(lds, N, racc + (q < 0.0 ? 3 * N : 0));
}

// E = real(ifft(fft(rho) ./ ik)) from the exact signed charge (which it
// clears), Es[:, ceil(t/bin)] .+= E and sum(abs2, E).  One block.
// Mixed (racc32 set): the charge is per species u64 quanta, cq[s] each
// (q_s * w * 2^-30), summed in FP64; E32 gets E * escale2 in FP32.
__global__ void __launch_bounds__(kSolveThreads)
k12_solve(u64* __restrict__ racc, double* __restrict__ E, double* __restrict__ Es, int64_t cap_bins,
          State12* __restrict__ st, const double2* __restrict__ twp, int N, int log2N, int qb, int bin,
          u64* __restrict__ racc32, double cq0, double cq1, float* __restrict__ E32, double escale2) {
    extern __shared__ double sm[];
    const int NP = solve_np(N);
    double* re = sm;
    double* im = sm + NP;
    double* red = sm + 2 * NP + N;
    SolveRegs Rg;
    solve_prefetch(Rg, twp, N, log2N);
    for (int j = threadIdx.x; j < N; j += blockDim.x) {
        double r;
        if (racc32) {
            const u64 a0 = racc32[j], a1 = racc32[N + j];
            racc32[j] = 0;
            racc32[N + j] = 0;
            r = cq0 * (double)a0 + cq1 * (double)a1;
        } else {
            u64* a = racc + 3 * j;
            u64* b = racc + 3 * N + 3 * j;
            const u128 pos = limbs_to_u128(a[0], a[1], a[2]), neg = limbs_to_u128(b[0], b[1], b[2]);
            a[0] = 0;
            a[1] = 0;
            a[2] = 0;
            b[0] = 0;
            b[1] = 0;
            b[2] = 0;
            r = pos >= neg ? fxp_to(pos - neg, qb) : -fxp_to(neg - pos, qb);
        }
        const int k = padi((int)brev((uint32_t)j, log2N));
        re[k] = r;
        im[k] = 0.0;
    }
    __syncthreads();
    solve_lds(re, im, Rg, twp, N, log2N);
    const int64_t t = st->step + 1;
    const int64_t bi = (t + bin - 1) / bin - 1;
    for (int j = threadIdx.x; j < N; j += blockDim.x) {
        const double e = solve_at(re, j);
        E[j] = e;
        if (E32) E32[j] = (float)(e * escale2);
        if (bi < cap_bins) Es[bi * N + j] = Es[bi * N + j] + e;
    }
    const double se = block_chunk_sum(red, N, [&](int64_t i) {
        const double e = solve_at(re, (int)i);
        return e * e;
    });
    const double emax = block_maxabs(red, N, [&](int64_t i) { return solve_at(re, (int)i); });
    if (threadIdx.x == 0) {
        st->se = se;
        if (!isfinite(emax)) st->err |= 1;
    }
}

// Ej = sum(k->E[k[1]]*k[2], d(x[j])); (vx, vy) = boris(...); x = mod(x + vx*dt, 1);
// then the particle's charge at the new x for the next step.
template <int R>
__global__ void __launch_bounds__(1024)
k12_push(const Tile12* __restrict__ tiles, double* __restrict__ x, double* __restrict__ vx,
         double* __restrict__ vy, const double* __restrict__ E, u64* __restrict__ racc, int N, double q0,
         double q1, double qm0, double qm1, double w, int qb, double B0, double dt) {
    extern __shared__ u64 lds[];
    const Tile12 T = tiles[blockIdx.x];
    const double q = T.species ? q1 : q0, qm = T.species ? qm1 : qm0;
    double* Eld = reinterpret_cast<double*>(lds);
    u64* acc = lds + N;
    for (int i = threadIdx.x; i < N; i += blockDim.x) Eld[i] = E[i];
    zero_lds(acc, 3 * N);
    __syncthreads();
    const double Nd = (double)N;
    const int32_t nmask = N - 1;
    for (int64_t p = threadIdx.x; p < T.count; p += blockDim.x) {
        const int64_t g = T.start + p;
        const double xj = x[g];
        double u = vx[g], v = vy[g];
        const double Ej = gather_gauss<R>(Eld, xj, Nd, nmask);
        boris12(u, v, Ej, B0, dt, qm);
        const double xn = jl_mod1f(xj + u * dt);
        x[g] = xn;
        vx[g] = u;
        vy[g] = v;
        deposit_gauss_q<R>(acc, N, xn, Nd, nmask, q, w, qb);
    }
    __syncthreads();
    flush_limbs(acc, N, racc + (q < 0.0 ? 3 * N : 0));
}

// Diagnostics partials over chunks of 4096 particle indices, when t % bin == 0:
// sum(vy^2 + vx^2 [+ M*(vy2^2 + vx2^2)]), sum(vx/P [+ M*vx2/P]), same for vy.
__global__ void __launch_bounds__(256)
k12_diag_chunks(const double* __restrict__ vx, const double* __restrict__ vy, int64_t P, int ns, double m0,
                double m1, int bin, const State12* __restrict__ st, double* __restrict__ dpart) {
    __shared__ double red[4 * kTreeLanes + 8];
    if ((st->step + 1) % bin != 0) return;
    const int64_t s = (int64_t)blockIdx.x * kTreeChunk;
    const int64_t n = P - s < kTreeChunk ? P - s : kTreeChunk;
    const double Pd = (double)P;
    double a = 0.0, b = 0.0, c = 0.0;
    if (threadIdx.x < kTreeLanes)
        for (int64_t i = threadIdx.x; i < n; i += kTreeLanes) {
            const int64_t j = s + i;
            const double x0 = vx[j], y0 = vy[j];
            double k = m0 * ((y0 * y0) + (x0 * x0));
            double mx = (m0 * x0) / Pd, my = (m0 * y0) / Pd;
            if (ns == 2) {
                const double x1 = vx[P + j], y1 = vy[P + j];
                k = k + m1 * ((y1 * y1) + (x1 * x1));
                mx = mx + (m1 * x1) / Pd;
                my = my + (m1 * y1) / Pd;
            }
            a = a + k;
            b = b + mx;
            c = c + my;
        }
    tree256x3<false>(red, a, b, c);
    if (threadIdx.x == 0) {
        dpart[3 * blockIdx.x] = a;
        dpart[3 * blockIdx.x + 1] = b;
        dpart[3 * blockIdx.x + 2] = c;
    }
}

// Canonical tree over n values spread with stride 3 (in place).
__device__ double tree_stride3(double* red, double* vals, int64_t n) {
    while (n > kTreeChunk) {
        const int64_t C = (n + kTreeChunk - 1) / kTreeChunk;
        for (int64_t c = 0; c < C; ++c) {
            const int64_t m = n - c * kTreeChunk < kTreeChunk ? n - c * kTreeChunk : kTreeChunk;
            const double s = block_chunk_sum(red, m, [&](int64_t i) { return vals[3 * (c * kTreeChunk + i)]; });
            if (threadIdx.x == 0) vals[3 * c] = s;
            __syncthreads();
        }
        n = C;
    }
    return block_chunk_sum(red, n, [&](int64_t i) { return vals[3 * i]; });
}

// D[t/bin, 1:5] (src/NGP1D2V2S.jl:52-56) when t % bin == 0; then step += 1.
__global__ void __launch_bounds__(256)
k12_diag_final(double* __restrict__ dpart, int64_t nchunks, int64_t P, int N, double n0, int bin,
               State12* __restrict__ st, double* __restrict__ D, int64_t cap_bins) {
    __shared__ double red[kTreeLanes + 8];
    const int64_t t = st->step + 1;
    if (t % bin == 0) {
        const double sk = tree_stride3(red, dpart, nchunks);
        const double sx = tree_stride3(red, dpart + 1, nchunks);
        const double sy = tree_stride3(red, dpart + 2, nchunks);
        const int64_t bi = t / bin - 1;
        if (threadIdx.x == 0 && bi < cap_bins) {
            double* row = D + 5 * bi;
            row[0] = row[0] + (st->se / (double)N) / 2.0;
            row[1] = row[1] + ((sk * n0) / (double)P) / 2.0;
            row[3] = row[3] + sx;
            row[4] = row[4] + sy;
            row[2] = row[2] + (row[0] + row[1]);
            for (int c = 0; c < 3; ++c) row[c] = row[c] * (2.0 / n0);
        }
    }
    if (threadIdx.x == 0) st->step = t;
}

// rho from the exact signed charge (clearing it), for the deposit operator.
__global__ void k12_rho(u64* __restrict__ racc, double* __restrict__ rho, int N, int qb) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
    u64* a = racc + 3 * j;
    u64* b = racc + 3 * N + 3 * j;
    const u128 pos = limbs_to_u128(a[0], a[1], a[2]), neg = limbs_to_u128(b[0], b[1], b[2]);
    a[0] = 0;
    a[1] = 0;
    a[2] = 0;
    b[0] = 0;
    b[1] = 0;
    b[2] = 0;
    rho[j] = pos >= neg ? fxp_to(pos - neg, qb) : -fxp_to(neg - pos, qb);
}

template <int R>
__global__ void __launch_bounds__(256)
k12_op_gather(const double* __restrict__ E, const double* __restrict__ x, int64_t n, double* __restrict__ Ej,
              int N) {
    extern __shared__ double Eld[];
    for (int i = threadIdx.x; i < N; i += blockDim.x) Eld[i] = E[i];
    __syncthreads();
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) Ej[j] = gather_gauss<R>(Eld, x[j], (double)N, N - 1);
}

__global__ void k12_op_boris(double* __restrict__ vx, double* __restrict__ vy, const double* __restrict__ Ej,
                             int64_t n, double B, double dt, double q_m) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    double u = vx[j], v = vy[j];
    boris12(u, v, Ej[j], B, dt, q_m);
    vx[j] = u;
    vy[j] = v;
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
int push12_lds_bytes(int N) { return 32 * N; }  // E + three limb arrays

void launch12_deposit(const Buffers12& b, hipStream_t s) {
    if (b.precision == 1) return launch12_deposit32(b, s);
    const int lds = 24 * b.N;
    u64* r = (u64*)b.racc;
    if (b.R == 6)
        k12_deposit<6><<<b.ntiles, b.block, lds, s>>>(b.tiles, b.x, r, b.N, b.q[0], b.q[1], b.w, b.qbits);
    else
        k12_deposit<7><<<b.ntiles, b.block, lds, s>>>(b.tiles, b.x, r, b.N, b.q[0], b.q[1], b.w, b.qbits);
}

void launch12_solve(const Buffers12& b, hipStream_t s) {
    const int lds = solve_lds_doubles(b.N) * 8;
    const bool m = b.precision == 1;
    const double cq0 = m ? b.q[0] * b.w * 0x1p-30 : 0.0, cq1 = m ? b.q[1] * b.w * 0x1p-30 : 0.0;
    k12_solve<<<1, kSolveThreads, lds, s>>>((u64*)b.racc, b.E, b.Es, b.cap_bins, b.st, b.tw, b.N, b.log2N,
                                            b.qbits, b.bin, m ? (u64*)b.racc32 : nullptr, cq0, cq1,
                                            m ? b.E32 : nullptr, b.escale2);
}

void launch12_push(const Buffers12& b, hipStream_t s) {
    if (b.precision == 1) return launch12_push32(b, s);
    const int lds = push12_lds_bytes(b.N);
    u64* r = (u64*)b.racc;
    if (b.R == 6)
        k12_push<6><<<b.ntiles, b.block, lds, s>>>(b.tiles, b.x, b.vx, b.vy, b.E, r, b.N, b.q[0], b.q[1],
                                                   b.q_m[0], b.q_m[1], b.w, b.qbits, b.B0, b.dt);
    else
        k12_push<7><<<b.ntiles, b.block, lds, s>>>(b.tiles, b.x, b.vx, b.vy, b.E, r, b.N, b.q[0], b.q[1],
                                                   b.q_m[0], b.q_m[1], b.w, b.qbits, b.B0, b.dt);
}

void launch12_diag(const Buffers12& b, hipStream_t s) {
    if (b.precision == 1) return launch12_diag32(b, s);
    k12_diag_chunks<<<(unsigned)b.nchunks, 256, 0, s>>>(b.vx, b.vy, b.P, b.ns, b.mass[0], b.mass[1], b.bin,
                                                       b.st, b.dpart);
    k12_diag_final<<<1, 256, 0, s>>>(b.dpart, b.nchunks, b.P, b.N, b.n0, b.bin, b.st, b.D, b.cap_bins);
}

void launch12_rho(const Buffers12& b, double* rho, hipStream_t s) {
    k12_rho<<<(b.N + 255) / 256, 256, 0, s>>>((u64*)b.racc, rho, b.N, b.qbits);
}

void launch12_op_gather(const Buffers12& b, const double* E, const double* x, int64_t n, double* Ej,
                        hipStream_t s) {
    const unsigned g = (unsigned)((n + 255) / 256);
    if (b.R == 6) k12_op_gather<6><<<g, 256, 8 * b.N, s>>>(E, x, n, Ej, b.N);
    else k12_op_gather<7><<<g, 256, 8 * b.N, s>>>(E, x, n, Ej, b.N);
}

void launch12_op_boris(double* vx, double* vy, const double* Ej, int64_t n, double B, double dt, double q_m,
                       hipStream_t s) {
    k12_op_boris<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(vx, vy, Ej, n, B, dt, q_m);
}

template <typename K>
static hipError_t allow12(K kernel, int bytes) {
    return hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                               bytes);
}

int configure_kernels_1d2v() {
    const int big = 32 * kMaxN;
    hipError_t e = hipSuccess;
    auto chk = [&](hipError_t r) { if (r != hipSuccess) e = r; };
    chk(allow12(k12_deposit<6>, big));
    chk(allow12(k12_deposit<7>, big));
    chk(allow12(k12_push<6>, big));
    chk(allow12(k12_push<7>, big));
    chk(allow12(k12_solve, big));
    (void)hipGetLastError();
    return e == hipSuccess ? 0 : -1;
}

}  // namespace pic
