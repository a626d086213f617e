// kernels_2d3v_lg.hip -- FP64 gfx950 kernels of the Lorenz-gauge
// electromagnetic path of the structured 2D3V module (src/PIC2D3V.jl,
// LorenzGaugeStaggeredField as src/2D3V.jl drives it).  Compiled with
// contraction off; the gather's @muladd is an explicit fma.  Every result is
// bitwise equal to the CPU restatement.
//
// One step (loop!, :875-979):
//   k_lg_spectral   per mode: the wave update of phi, Ax, Ay, Az for any ImEx
//                   theta (:629-635) and E = -im k phi+ - (A+ - A0)/dt,
//                   B = curl (A+ + A0)/2 in the spectrum (:917-925)
//   k_lg_cols_inv,  the inverse transforms of the six (:927-934), the row
//   k_lg_rows_inv   pass scaling by 1/(NX*NY) and writing update!'s EBxyz
//                   (:468-479) beside the complex fields
//   k_lg_push       gather (six components, the field staged in an LDS window
//                   of the chunk's cells), ElectromagneticBoris, the two half
//                   moves; the half-step positions go to xh, yh
//   k_lg_deposit    the charge at (xh, yh) and the currents (v*qw) at the new
//                   positions: exact signed limbs in an LDS window, flushed
//                   with integer atomics
//   k_lg_finish     rho0, J0 from the limbs (imaginary parts 0) and the
//                   potentials' buffer copies (:969-978)
// diagnose! (:1332-1398): the row (field energy over the ghosted EBxyz, field
// momentum, the species sums), the in-place inverse transforms of Ax0, Ay0,
// Az0, phi0, rho0, Jx0, Jy0, Jz0, and the binned averages.
//
// LorenzGaugeField's centred loop! (:665-745) reuses them: the push (with
// the centred move) comes first, then the deposits at the new positions,
// k_lg_finish<2>, the forward transforms of rho+, J+, k_lg_spectral<true> and
// the inverse transforms with update!.
// LorenzGaugeSemiImplicitField's iterated loop! (:993-1100) has kernels of its
// own at the end of the file: the push into work arrays, exact deposits into
// the buffered grid, the grid and its isapprox test, the periodic reduction,
// the wave update with three source weights and the rotation of the
// potentials.
#include <climits>

#include "fft_lines.h"
#include "pic_device.h"
#include "pic2d3v_lg_internal.h"
#include "shapes23.h"

#pragma clang fp contract(off)

namespace pic {

// Window budgets, sized so that four blocks share a CU's 160 KB of LDS: the
// push stages six doubles per cell (39 KB), the current deposit three grids
// of three limbs (39 KB).  A chunk of 1024 particles sorted per cell spans
// about 6 x 69 cells with the quadratic B-spline, a regrouped one 13 x 13.
constexpr int kLGWinCells = 832;
constexpr int kLGDepCells = 552;

__host__ __device__ inline int lg_cols(int NX, int NY) {
    const int c = kLGLinePoints / NY < 1 ? 1 : kLGLinePoints / NY;
    return c < NX ? c : NX;
}
__host__ __device__ inline int lg_rows(int NX, int NY) {
    const int r = kLGLinePoints / NX < 1 ? 1 : kLGLinePoints / NX;
    return r < NY ? r : NY;
}
static size_t lg_cols_lds(int NX, int NY) { return sizeof(double) * 2 * (size_t)lg_cols(NX, NY) * line_stride(NY); }
static size_t lg_rows_lds(int NX, int NY) { return sizeof(double) * 2 * (size_t)lg_rows(NX, NY) * line_stride(NX); }

struct LGList {
    double2* a[8];
};

__device__ __forceinline__ double bits_to_d(u64 b) { return __longlong_as_double((long long)b); }
__device__ __forceinline__ u64 d_to_bits(double d) { return (u64)__double_as_longlong(d); }
__device__ __forceinline__ double fmax_nn(double a, double b) { return b > a ? b : a; }

// the exact sum of a cell (three 38-bit two's-complement limb counters,
// modulo 2^114) rounded once
__device__ __forceinline__ double lg_limbs_v(u64 a0, u64 a1, u64 a2, int q) {
    u128 v = (u128)a0 + ((u128)a1 << kLimbS) + ((u128)a2 << (2 * kLimbS));
    v &= ((u128)1 << (3 * kLimbS)) - 1;
    const i128 s = (v >> (3 * kLimbS - 1)) ? (i128)v - ((i128)1 << (3 * kLimbS)) : (i128)v;
    return i128_to_double(s, q);
}
__device__ __forceinline__ double lg_limbs(const u64* r, int64_t G, int64_t k, int q) {
    return lg_limbs_v(r[k], r[G + k], r[2 * G + k], q);
}

// q of a sum bounded by sum_s P_s*(|qw_s|*M_s) (M = 1 for the charge)
__device__ __forceinline__ int lg_frac_bits(const LG23& L, int comp) {
    double total = 0.0;
    for (int s = 0; s < L.ns; ++s)
        total += (double)L.P[s] * (fabs(L.qw[s]) * (comp < 0 ? 1.0 : bits_to_d(L.umax[3 * s + comp])));
    if (!(total > 0.0) || isinf(total) || total != total) return 64;
    const int q = 110 - ilogb(total);
    return q > 1100 ? 1100 : q;
}

// ---------------------------------------------------------------------------
// the field solve
// ---------------------------------------------------------------------------
__device__ __forceinline__ void cmul_lg(double ar, double ai, double br, double bi, double& cr, double& ci) {
    cr = ar * br - ai * bi;
    ci = ar * bi + ai * br;
}

// CENTRED (LorenzGaugeField, :711-725): the sources are already in the
// spectrum and their [1,1] is cleared here; new = (num*x+ + dt^2*s)/den - x0,
// x0 = x+, x+ = new (the state slots PHI0.. hold x+, PHIM.. x0), and E, B use
// (phi+ + phi0)/2 and A+ - A0, A+ + A0.
template <bool CENTRED>
__global__ void __launch_bounds__(256) k_lg_spectral(LG23 L) {
    const int NX = L.NX;
    const int64_t G = (int64_t)NX * L.NY;
    if (blockIdx.x == 0 && threadIdx.x < 3 * kMaxSpecies23) L.umax[threadIdx.x] = 0;
    const double dt = L.dt, dt2 = dt * dt, th = L.theta;
    double2* F = L.F;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < G; k += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(k & (NX - 1)), j = (int)(k >> L.lx);
        const double kxi = L.kx[i], kyj = L.ky[j];
        const double k2 = kxi * kxi + kyj * kyj;
        const double num = 2.0 - dt2 * k2 * (1.0 - th);
        const double den = 1.0 + dt2 * k2 * th / 2.0;
        const int x0s[4] = {kLGPhi0, kLGAx0, kLGAy0, kLGAz0};
        const int xms[4] = {kLGPhiM, kLGAxM, kLGAyM, kLGAzM};
        const int srcs[4] = {kLGRho0, kLGJx0, kLGJy0, kLGJz0};
        double2 p[4], q0[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const double2 x0 = F[x0s[a] * G + k], xm = F[xms[a] * G + k];
            double2 sr = F[srcs[a] * G + k];
            if (CENTRED && k == 0) {
                sr = make_double2(0.0, 0.0);
                F[srcs[a] * G] = sr;
            }
            p[a].x = (num * x0.x + dt2 * sr.x) / den - xm.x;
            p[a].y = (num * x0.y + dt2 * sr.y) / den - xm.y;
            if (CENTRED) {
                q0[a] = x0;
                F[xms[a] * G + k] = x0;
                F[x0s[a] * G + k] = p[a];
            } else {
                q0[a] = x0;
                L.plus[a * G + k] = p[a];
            }
        }
        // A0 of the E and B formulas: the state's A0 (staggered) or the old A+ (centred)
        const double2 ax0 = q0[1], ay0 = q0[2], az0 = q0[3];
        double er, ei, fr, fi;
        double phr = p[0].x, phi_ = p[0].y;
        if (CENTRED) {
            phr = p[0].x + q0[0].x;
            phi_ = p[0].y + q0[0].y;
        }
        cmul_lg(0.0 * kxi, -kxi, phr, phi_, er, ei);
        if (CENTRED) {
            er = er / 2.0;
            ei = ei / 2.0;
        }
        F[kLGEx * G + k] = make_double2(er - (p[1].x - ax0.x) / dt, ei - (p[1].y - ax0.y) / dt);
        cmul_lg(0.0 * kyj, -kyj, phr, phi_, er, ei);
        if (CENTRED) {
            er = er / 2.0;
            ei = ei / 2.0;
        }
        F[kLGEy * G + k] = make_double2(er - (p[2].x - ay0.x) / dt, ei - (p[2].y - ay0.y) / dt);
        F[kLGEz * G + k] = make_double2(-(p[3].x - az0.x) / dt, -(p[3].y - az0.y) / dt);
        const double szr = p[3].x + az0.x, szi = p[3].y + az0.y;
        cmul_lg(0.0 * kyj, kyj, szr, szi, er, ei);
        F[kLGBx * G + k] = make_double2(er / 2.0, ei / 2.0);
        cmul_lg(0.0 * kxi, -kxi, szr, szi, er, ei);
        F[kLGBy * G + k] = make_double2(er / 2.0, ei / 2.0);
        cmul_lg(0.0 * kxi, kxi, p[2].x + ay0.x, p[2].y + ay0.y, er, ei);
        cmul_lg(0.0 * kyj, kyj, p[1].x + ax0.x, p[1].y + ax0.y, fr, fi);
        F[kLGBz * G + k] = make_double2(er / 2.0 - fr / 2.0, ei / 2.0 - fi / 2.0);
    }
}

// In-place transform along y of C columns of array blockIdx.y (INV: the
// unnormalised inverse).
template <bool INV>
__global__ void __launch_bounds__(256) k_lg_cols(LGList lst, int NX, int NY, int ly, const double2* tw) {
    extern __shared__ double lds_lg[];
    const int C = lg_cols(NX, NY), st = line_stride(NY), lc = __ffs(C) - 1;
    double* re = lds_lg;
    double* im = lds_lg + C * st;
    double2* Z = lst.a[blockIdx.y];
    // a contiguous range of columns per XCD (blocks go round-robin to the 8
    // XCDs): neighbouring columns share cache lines, read through one L2
    const int nb = gridDim.x;
    const int bi = (nb & 7) == 0 ? (blockIdx.x & 7) * (nb >> 3) + (blockIdx.x >> 3) : blockIdx.x;
    const int i0 = bi * C;
    for (int e = threadIdx.x; e < C * NY; e += blockDim.x) {
        const int c = e & (C - 1), j = e >> lc;
        const double2 v = Z[(int64_t)j * NX + i0 + c];
        const int pos = c * st + padi((int)brev((uint32_t)j, ly));
        re[pos] = v.x;
        im[pos] = v.y;
    }
    __syncthreads();
    fft_lines(re, im, tw, ly, C, st, INV);
    for (int e = threadIdx.x; e < C * NY; e += blockDim.x) {
        const int c = e & (C - 1), j = e >> lc;
        const int pos = c * st + padi(j);
        Z[(int64_t)j * NX + i0 + c] = make_double2(re[pos], im[pos]);
    }
}

// In-place transform along x of R rows of array blockIdx.y.  INV: the inverse,
// scaled by 1/(NX*NY) (FFTW's ifft); else the unnormalised forward transform.
// EBOUT: the list is Ex, Ey, Ez, Bx, By, Bz and the pass also writes update!'s
// EBxyz = 0.0 + real(F) (+ B0 for B).
template <bool INV, bool EBOUT>
__global__ void __launch_bounds__(256) k_lg_rows(LGList lst, LG23 L) {
    extern __shared__ double lds_lg[];
    const int NX = L.NX, lx = L.lx;
    const int R = lg_rows(NX, L.NY), st = line_stride(NX);
    double* re = lds_lg;
    double* im = lds_lg + R * st;
    const int a = blockIdx.y;
    double2* Z = lst.a[a];
    const int64_t G = (int64_t)NX * L.NY;
    const int64_t k0 = (int64_t)blockIdx.x * R * NX;
    for (int e = threadIdx.x; e < R * NX; e += blockDim.x) {
        const int r = e >> lx, i = e & (NX - 1);
        const int pos = r * st + padi((int)brev((uint32_t)i, lx));
        const double2 v = Z[k0 + e];
        re[pos] = v.x;
        im[pos] = v.y;
    }
    __syncthreads();
    fft_lines(re, im, L.tw, lx, R, st, INV);
    const double scale = 1.0 / (double)G;
    for (int e = threadIdx.x; e < R * NX; e += blockDim.x) {
        const int r = e >> lx, i = e & (NX - 1);
        const int pos = r * st + padi(i);
        const double zr = INV ? re[pos] * scale : re[pos], zi = INV ? im[pos] * scale : im[pos];
        Z[k0 + e] = make_double2(zr, zi);
        if (EBOUT) {
            const double ev = 0.0 + zr;
            L.EB[a * G + k0 + e] = a < 3 ? ev : ev + L.B0[a - 3];
        }
    }
}

// ifft! in place (along y, then along x and the scale)
static void lg_inverse(const LGList& lst, int n, const LG23& L, bool ebout, hipStream_t st) {
    const int C = lg_cols(L.NX, L.NY), R = lg_rows(L.NX, L.NY);
    const dim3 gc((unsigned)(L.NX / C), (unsigned)n), gr((unsigned)(L.NY / R), (unsigned)n);
    k_lg_cols<true><<<gc, 256, lg_cols_lds(L.NX, L.NY), st>>>(lst, L.NX, L.NY, L.ly, L.tw);
    if (ebout) k_lg_rows<true, true><<<gr, 256, lg_rows_lds(L.NX, L.NY), st>>>(lst, L);
    else k_lg_rows<true, false><<<gr, 256, lg_rows_lds(L.NX, L.NY), st>>>(lst, L);
}

// fft! in place (along x, then along y)
static void lg_forward(const LGList& lst, int n, const LG23& L, hipStream_t st) {
    const int C = lg_cols(L.NX, L.NY), R = lg_rows(L.NX, L.NY);
    const dim3 gc((unsigned)(L.NX / C), (unsigned)n), gr((unsigned)(L.NY / R), (unsigned)n);
    k_lg_rows<false, false><<<gr, 256, lg_rows_lds(L.NX, L.NY), st>>>(lst, L);
    k_lg_cols<false><<<gc, 256, lg_cols_lds(L.NX, L.NY), st>>>(lst, L.NX, L.NY, L.ly, L.tw);
}

static unsigned grid_blocks(int64_t n) {
    const int64_t b = (n + 255) / 256;
    return (unsigned)(b < 8192 ? (b > 0 ? b : 1) : 8192);
}

void lg_launch_solve(const LG23& L, bool centred, hipStream_t st) {
    const int64_t G = (int64_t)L.NX * L.NY;
    if (centred) {
        LGList src{};
        for (int c = 0; c < 4; ++c) src.a[c] = L.F + (size_t)(kLGRho0 + c) * G;
        lg_forward(src, 4, L, st);
        k_lg_spectral<true><<<grid_blocks(G), 256, 0, st>>>(L);
    } else {
        k_lg_spectral<false><<<grid_blocks(G), 256, 0, st>>>(L);
    }
    LGList lst{};
    for (int c = 0; c < 6; ++c) lst.a[c] = L.F + (size_t)(kLGEx + c) * G;
    lg_inverse(lst, 6, L, true, st);
}

// update! alone: EBxyz from the six fields of the state
__global__ void __launch_bounds__(256) k_lg_update_eb(LG23 L) {
    const int64_t G = (int64_t)L.NX * L.NY;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < 6 * G; e += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(e / G);
        const double ev = 0.0 + L.F[(kLGEx + c) * G + (e - c * G)].x;
        L.EB[e] = c < 3 ? ev : ev + L.B0[c - 3];
    }
}

void lg_launch_update_eb(const LG23& L, hipStream_t st) {
    k_lg_update_eb<<<grid_blocks(6 * (int64_t)L.NX * L.NY), 256, 0, st>>>(L);
}

// ---------------------------------------------------------------------------
// the particle loop
// ---------------------------------------------------------------------------
// ElectromagneticBoris (:56-69)
__device__ __forceinline__ void lg_boris(double dt, double q_m, const double* f, double& vx, double& vy, double& vz) {
    const double th = dt / 2.0 * q_m;
    const double t1 = f[3] * th, t2 = f[4] * th, t3 = f[5] * th;
    const double sc = 2.0 / (1.0 + (t1 * t1 + t2 * t2 + t3 * t3));
    const double e1 = f[0] * th, e2 = f[1] * th, e3 = f[2] * th;
    const double m1 = vx + e1, m2 = vy + e2, m3 = vz + e3;
    const double c1 = m2 * t3 - m3 * t2, c2 = m3 * t1 - m1 * t3, c3 = m1 * t2 - m2 * t1;
    const double u1 = m1 + c1, u2 = m2 + c2, u3 = m3 + c3;
    const double d1 = u2 * t3 - u3 * t2, d2 = u3 * t1 - u1 * t3, d3 = u1 * t2 - u2 * t1;
    vx = (m1 + d1 * sc) + e1;
    vy = (m2 + d2 * sc) + e2;
    vz = (m3 + d3 * sc) + e3;
}

// The window of a chunk: the block-reduced extent of the particles' base
// cells (relative to (ax, ay), periodic), widened by the stencil and a halo;
// win = (ox, oy, wx, wy, fits in CELLS).
template <int SH, int CELLS>
__device__ __forceinline__ void lg_window(const Grid23& g, const double* xs, const double* ys, int nv, int ax, int ay,
                                          int (*red)[4], int* win) {
    constexpr int LO = ShapeExtent<SH>::lo, HI = ShapeExtent<SH>::hi;
    const int NX = g.NX, NY = g.NY, t = threadIdx.x;
    int mnx = INT_MAX, mxx = INT_MIN, mny = INT_MAX, mxy = INT_MIN;
#pragma unroll
    for (int k = 0; k < kTile23PPT; ++k) {
        if (k >= nv) break;
        const int dx = ((base_cell(xs[k], g.NX_Lx, NX) - ax + NX / 2) & (NX - 1)) - NX / 2;
        const int dy = ((base_cell(ys[k], g.NY_Ly, NY) - ay + NY / 2) & (NY - 1)) - NY / 2;
        mnx = min(mnx, dx);
        mxx = max(mxx, dx);
        mny = min(mny, dy);
        mxy = max(mxy, dy);
    }
    mnx = wave_min_i(mnx);
    mxx = wave_max_i(mxx);
    mny = wave_min_i(mny);
    mxy = wave_max_i(mxy);
    const int wv = t >> 6;
    if ((t & 63) == 0) {
        red[wv][0] = mnx;
        red[wv][1] = mxx;
        red[wv][2] = mny;
        red[wv][3] = mxy;
    }
    __syncthreads();
    if (t == 0) {
        for (int w = 1; w < kTile23Threads / 64; ++w) {
            mnx = min(mnx, red[w][0]);
            mxx = max(mxx, red[w][1]);
            mny = min(mny, red[w][2]);
            mxy = max(mxy, red[w][3]);
        }
        int wx = mxx - mnx + LO + HI + 2 * kTile23Halo + 1, sx = ax + mnx - LO - kTile23Halo;
        int wy = mxy - mny + LO + HI + 2 * kTile23Halo + 1, sy = ay + mny - LO - kTile23Halo;
        if (wx >= NX) {
            wx = NX;
            sx = 0;
        }
        if (wy >= NY) {
            wy = NY;
            sy = 0;
        }
        win[0] = sx & (NX - 1);
        win[1] = sy & (NY - 1);
        win[2] = wx;
        win[3] = wy;
        win[4] = wx * wy <= CELLS;
    }
    __syncthreads();
}

// MODE 0: the staggered loop's two half moves, the first written to (xh, yh);
// 1 (LorenzGaugeField, :683-690): x = unimod(x + (vx0 + vx)/2*dt); 2
// (LorenzGaugeSemiImplicitField, :1036-1040): xw = unimod(x + vx0*dt) and vw
// into the work arrays o, s unchanged (modes 0 and 1 write s, o unused).
// (the quadratic B-spline's loop is capped at 128 registers: four waves per
// SIMD, which the LDS window allows)
template <int SH, int MODE>
__global__ void __launch_bounds__(kTile23Threads) __attribute__((amdgpu_waves_per_eu(SH == 4 ? 4 : 1)))
k_lg_push(LG23 L, Grid23 g, Species23 s, Species23 o, int sidx, double* __restrict__ xh, double* __restrict__ yh) {
    constexpr bool CENTRED = MODE == 1;
    constexpr int n = ShapeCount<SH>::value;
    __shared__ double Ew[6 * kLGWinCells];
    __shared__ int red[4][4];
    __shared__ int win[5];
    __shared__ u64 mred[kTile23Threads / 64];
    const int t = threadIdx.x;
    const int NX = g.NX, NY = g.NY;
    const int64_t G = (int64_t)NX * NY;
    const int64_t c0 = (int64_t)blockIdx.x * kTile23Chunk;
    double xs[kTile23PPT], ys[kTile23PPT];
    int nv = 0;
#pragma unroll
    for (int k = 0; k < kTile23PPT; ++k) {
        const int64_t p = c0 + t + (int64_t)k * kTile23Threads;
        xs[k] = 0.0;
        ys[k] = 0.0;
        if (p < s.P) {
            xs[k] = s.x[p];
            ys[k] = s.y[p];
            nv = k + 1;
        }
    }
    const int ax = base_cell(s.x[c0], g.NX_Lx, NX), ay = base_cell(s.y[c0], g.NY_Ly, NY);
    lg_window<SH, kLGWinCells>(g, xs, ys, nv, ax, ay, red, win);
    const int ox = win[0], oy = win[1], wx = win[2], wy = win[3];
    const bool tile = win[4] != 0;
    if (tile)
        for (int e = t; e < wx * wy; e += kTile23Threads) {
            const int ry = e / wx, rx = e - ry * wx;
            const int64_t k = ((ox + rx) & (NX - 1)) + (int64_t)NX * ((oy + ry) & (NY - 1));
#pragma unroll
            for (int c = 0; c < 6; ++c) Ew[c * kLGWinCells + e] = L.EB[c * G + k];
        }
    __syncthreads();
    double m1 = 0.0, m2 = 0.0, m3 = 0.0;
    const double dt = g.dt;
#pragma unroll 1
    for (int k = 0; k < kTile23PPT; ++k) {
        const int64_t p = c0 + t + (int64_t)k * kTile23Threads;
        if (p >= s.P) break;
        double x = xs[k], y = ys[k], vx = s.vx[p], vy = s.vy[p], vz = s.vz[p];
        int32_t cx[n], cy[n];
        double wxs[n], wys[n];
        int lx[n], ly[n];
        shape_w<SH>(x, g.NX_Lx, NX, cx, wxs);
        shape_w<SH>(y, g.NY_Ly, NY, cy, wys);
        bool in = tile;
#pragma unroll
        for (int q = 0; q < n; ++q) {
            lx[q] = (cx[q] - ox) & (NX - 1);
            ly[q] = (cy[q] - oy) & (NY - 1);
            in = in && lx[q] < wx && ly[q] < wy;
        }
        double f[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        // the window and the fallback in blocks of their own (one branch per
        // particle, not one per stencil cell)
        if (in) {
#pragma unroll
            for (int jj = 0; jj < n; ++jj)
#pragma unroll
                for (int ii = 0; ii < n; ++ii) {
                    const double wxy = wxs[ii] * wys[jj];
                    const int e = lx[ii] + wx * ly[jj];
#pragma unroll
                    for (int c = 0; c < 6; ++c) f[c] = fma(Ew[c * kLGWinCells + e], wxy, f[c]);
                }
        } else {
#pragma unroll
            for (int jj = 0; jj < n; ++jj)
#pragma unroll
                for (int ii = 0; ii < n; ++ii) {
                    const double wxy = wxs[ii] * wys[jj];
                    const int64_t kk = cx[ii] + (int64_t)NX * cy[jj];
#pragma unroll
                    for (int c = 0; c < 6; ++c) f[c] = fma(L.EB[c * G + kk], wxy, f[c]);
                }
        }
        const double vx0 = vx, vy0 = vy;
        lg_boris(dt, s.q_m, f, vx, vy, vz);
        if (MODE == 2) {
            o.x[p] = unimod23(x + vx0 * dt, g.Lx);
            o.y[p] = unimod23(y + vy0 * dt, g.Ly);
            o.vx[p] = vx;
            o.vy[p] = vy;
            o.vz[p] = vz;
        } else if (CENTRED) {
            x = unimod23(x + (vx0 + vx) / 2.0 * dt, g.Lx);
            y = unimod23(y + (vy0 + vy) / 2.0 * dt, g.Ly);
        } else {
            x = unimod23(x + vx * dt / 2.0, g.Lx);
            y = unimod23(y + vy * dt / 2.0, g.Ly);
            xh[p] = x;
            yh[p] = y;
            x = unimod23(x + vx * dt / 2.0, g.Lx);
            y = unimod23(y + vy * dt / 2.0, g.Ly);
        }
        if (MODE != 2) {
            s.x[p] = x;
            s.y[p] = y;
            s.vx[p] = vx;
            s.vy[p] = vy;
            s.vz[p] = vz;
        }
        m1 = fmax_nn(m1, fabs(vx));
        m2 = fmax_nn(m2, fabs(vy));
        m3 = fmax_nn(m3, fabs(vz));
    }
    const double ms[3] = {m1, m2, m3};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        u64 b = d_to_bits(ms[c]);
#pragma unroll
        for (int h = 32; h > 0; h >>= 1) {
            const u64 o = __shfl_xor(b, h, 64);
            b = o > b ? o : b;
        }
        if ((t & 63) == 0) mred[t >> 6] = b;
        __syncthreads();
        if (t == 0) {
            u64 m = mred[0];
            for (int w = 1; w < kTile23Threads / 64; ++w) m = mred[w] > m ? mred[w] : m;
            atomicMax(L.umax + 3 * sidx + c, m);
        }
        __syncthreads();
    }
}

// NG = 1: the charge (weight qw) at (X, Y); NG = 3: the currents (weights
// vx*qw, vy*qw, vz*qw) at (X, Y) = the species' positions.  GHOST (the
// semi-implicit field's buffered grid): a stencil cell outside the NX x NY
// interior also adds its limbs to that ghost cell of eacc, and a stencil
// cell outside the buffered grid drops the particle and sets eflag.
template <int SH, int NG, bool GHOST>
__global__ void __launch_bounds__(kTile23Threads)
k_lg_deposit(LG23 L, Grid23 g, Species23 s, const double* __restrict__ X, const double* __restrict__ Y) {
    constexpr int n = ShapeCount<SH>::value;
    constexpr int g0 = NG == 1 ? 0 : 1;  // first grid: 0 rho, 1..3 J
    __shared__ u64 Rw[NG * 3 * kLGDepCells];
    __shared__ int red[4][4];
    __shared__ int win[5];
    const int t = threadIdx.x;
    const int NX = g.NX, NY = g.NY;
    const int64_t G = (int64_t)NX * NY;
    const int64_t c0 = (int64_t)blockIdx.x * kTile23Chunk;
    // Lane l of wave w takes particles l*16 + w*4 + k: the particles are
    // sorted per cell, so in each iteration a wave's lanes sit in different
    // cells and few of them add to the same LDS word at once.
    const int64_t pb = c0 + (int64_t)(t & 63) * (kTile23Threads / 64 * kTile23PPT) + (t >> 6) * kTile23PPT;
    double xs[kTile23PPT], ys[kTile23PPT];
    int nv = 0;
#pragma unroll
    for (int k = 0; k < kTile23PPT; ++k) {
        const int64_t p = pb + k;
        xs[k] = 0.0;
        ys[k] = 0.0;
        if (p < s.P) {
            xs[k] = X[p];
            ys[k] = Y[p];
            nv = k + 1;
        }
    }
    const int ax = base_cell(X[c0], g.NX_Lx, NX), ay = base_cell(Y[c0], g.NY_Ly, NY);
    lg_window<SH, kLGDepCells>(g, xs, ys, nv, ax, ay, red, win);
    const int ox = win[0], oy = win[1], wx = win[2], wy = win[3];
    const bool tile = win[4] != 0;
    if (tile)
        for (int e = t; e < NG * 3 * kLGDepCells; e += kTile23Threads) Rw[e] = 0;
    __syncthreads();
    int q[NG];
#pragma unroll
    for (int c = 0; c < NG; ++c) q[c] = NG == 1 ? L.qr : lg_frac_bits(L, c);
    constexpr u64 M = (1ull << kLimbS) - 1;
#pragma unroll 1
    for (int k = 0; k < kTile23PPT; ++k) {
        const int64_t p = pb + k;
        if (p >= s.P) break;
        double w[NG];
        if (NG == 1) {
            w[0] = s.qw;
        } else {
            const double v[3] = {s.vx[p], s.vy[p], s.vz[p]};
#pragma unroll
            for (int c = 0; c < NG; ++c) w[c] = v[c] * s.qw;
        }
        int32_t cx[n], cy[n];
        double wxs[n], wys[n];
        int lx[n], ly[n];
        int ux[n], uy[n];  // GHOST: the buffered rows (unwrapped cell + b)
        if (GHOST) {
            shape_w<SH>(xs[k], g.NX_Lx, 0, cx, wxs);
            shape_w<SH>(ys[k], g.NY_Ly, 0, cy, wys);
            bool ok = true;
#pragma unroll
            for (int qq = 0; qq < n; ++qq) {
                ux[qq] = cx[qq] + L.b;
                uy[qq] = cy[qq] + L.b;
                ok = ok && (unsigned)ux[qq] < (unsigned)L.NXe && (unsigned)uy[qq] < (unsigned)L.NYe;
                cx[qq] &= NX - 1;
                cy[qq] &= NY - 1;
            }
            if (!ok) {
                atomicOr(L.eflag, 1);
                continue;
            }
        } else {
            shape_w<SH>(xs[k], g.NX_Lx, NX, cx, wxs);
            shape_w<SH>(ys[k], g.NY_Ly, NY, cy, wys);
        }
        bool in = tile;
#pragma unroll
        for (int qq = 0; qq < n; ++qq) {
            lx[qq] = (cx[qq] - ox) & (NX - 1);
            ly[qq] = (cy[qq] - oy) & (NY - 1);
            in = in && lx[qq] < wx && ly[qq] < wy;
        }
#pragma unroll
        for (int jj = 0; jj < n; ++jj)
#pragma unroll
            for (int ii = 0; ii < n; ++ii) {
                const double wxy = wxs[ii] * wys[jj];
                // GHOST: the cell lies outside the interior of the buffered grid
                const bool ghost = GHOST && ((unsigned)(ux[ii] - L.b) >= (unsigned)NX ||
                                             (unsigned)(uy[jj] - L.b) >= (unsigned)NY);
#pragma unroll
                for (int c = 0; c < NG; ++c) {
                    const double term = wxy * w[c];
                    if (term == 0.0) continue;
                    const u128 u = (u128)fix_round128(term, q[c]);
                    const u64 l[3] = {(u64)u & M, (u64)(u >> kLimbS) & M, (u64)(u >> (2 * kLimbS)) & M};
                    if (ghost) {
                        const int64_t Ge = (int64_t)L.NXe * L.NYe;
                        u64* r = L.eacc + (size_t)(g0 + c) * 3 * Ge;
                        const int64_t kk = ux[ii] + (int64_t)L.NXe * uy[jj];
#pragma unroll
                        for (int b = 0; b < 3; ++b)
                            if (l[b]) atomicAdd(r + b * Ge + kk, l[b]);
                    }
                    if (in) {
                        const int e = lx[ii] + wx * ly[jj];
#pragma unroll
                        for (int b = 0; b < 3; ++b)
                            if (l[b]) atomicAdd(Rw + (c * 3 + b) * kLGDepCells + e, l[b]);
                    } else {
                        u64* r = L.racc + (size_t)(g0 + c) * 3 * G;
                        const int64_t kk = cx[ii] + (int64_t)NX * cy[jj];
#pragma unroll
                        for (int b = 0; b < 3; ++b)
                            if (l[b]) atomicAdd(r + b * G + kk, l[b]);
                    }
                }
            }
    }
    __syncthreads();
    if (tile)
        for (int e = t; e < wx * wy; e += kTile23Threads) {
            const int ry = e / wx, rx = e - ry * wx;
            const int64_t k = ((ox + rx) & (NX - 1)) + (int64_t)NX * ((oy + ry) & (NY - 1));
#pragma unroll
            for (int c = 0; c < NG; ++c)
#pragma unroll
                for (int b = 0; b < 3; ++b) {
                    const u64 v = Rw[(c * 3 + b) * kLGDepCells + e];
                    if (v) atomicAdd(L.racc + (size_t)(g0 + c) * 3 * G + b * G + k, v);
                }
        }
}

// rho0, J0 from the limbs (cleared).  MODE 0: then the staggered buffer
// copies of the potentials; 1 (warmup!): rho0 = 0, as warmup!'s second
// deposit! discards the charge; 2 (centred): rho+, J+ only.
template <int MODE>
__global__ void __launch_bounds__(256) k_lg_finish(LG23 L) {
    constexpr bool WARM = MODE == 1;
    const int64_t G = (int64_t)L.NX * L.NY;
    int qj[3];
    for (int c = 0; c < 3; ++c) qj[c] = lg_frac_bits(L, c);
    double2* F = L.F;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < G; k += (int64_t)gridDim.x * blockDim.x) {
        if (WARM) {
            F[kLGRho0 * G + k] = make_double2(0.0, 0.0);
        } else {
            F[kLGRho0 * G + k] = make_double2(lg_limbs(L.racc, G, k, L.qr), 0.0);
            for (int b = 0; b < 3; ++b) L.racc[b * G + k] = 0;
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            u64* r = L.racc + (size_t)(1 + c) * 3 * G;
            F[(kLGJx0 + c) * G + k] = make_double2(lg_limbs(r, G, k, qj[c]), 0.0);
            for (int b = 0; b < 3; ++b) r[b * G + k] = 0;
        }
        if (MODE == 0) {
            const int x0s[4] = {kLGPhi0, kLGAx0, kLGAy0, kLGAz0};
            const int xms[4] = {kLGPhiM, kLGAxM, kLGAyM, kLGAzM};
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                F[xms[a] * G + k] = F[x0s[a] * G + k];
                F[x0s[a] * G + k] = L.plus[a * G + k];
            }
        }
    }
}

__global__ void __launch_bounds__(256) k_lg_advect(Grid23 g, Species23 s, double h) {
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < s.P; p += (int64_t)gridDim.x * blockDim.x) {
        s.x[p] = unimod23(s.x[p] + s.vx[p] * h, g.Lx);
        s.y[p] = unimod23(s.y[p] + s.vy[p] * h, g.Ly);
    }
}

__global__ void k_lg_zero_umax(LG23 L) {
    if (threadIdx.x < 3 * kMaxSpecies23) L.umax[threadIdx.x] = 0;
}

__global__ void __launch_bounds__(256) k_lg_vmax(LG23 L, Species23 s, int sidx) {
    __shared__ u64 mred[256 / 64];
    const int t = threadIdx.x;
    double m[3] = {0.0, 0.0, 0.0};
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + t; p < s.P; p += (int64_t)gridDim.x * blockDim.x) {
        m[0] = fmax_nn(m[0], fabs(s.vx[p]));
        m[1] = fmax_nn(m[1], fabs(s.vy[p]));
        m[2] = fmax_nn(m[2], fabs(s.vz[p]));
    }
    for (int c = 0; c < 3; ++c) {
        u64 b = d_to_bits(m[c]);
#pragma unroll
        for (int h = 32; h > 0; h >>= 1) {
            const u64 o = __shfl_xor(b, h, 64);
            b = o > b ? o : b;
        }
        if ((t & 63) == 0) mred[t >> 6] = b;
        __syncthreads();
        if (t == 0) {
            u64 mm = mred[0];
            for (int w = 1; w < 4; ++w) mm = mred[w] > mm ? mred[w] : mm;
            atomicMax(L.umax + 3 * sidx + c, mm);
        }
        __syncthreads();
    }
}

static unsigned chunks23(int64_t P) { return (unsigned)((P + kTile23Chunk - 1) / kTile23Chunk); }

template <template <int> class K, typename... A>
static void by_shape(int shape, A... a) {
    switch (shape) {
        case 0: K<0>::go(a...); break;
        case 1: K<1>::go(a...); break;
        case 2: K<2>::go(a...); break;
        case 3: K<3>::go(a...); break;
        case 4: K<4>::go(a...); break;
        case 5: K<5>::go(a...); break;
        case 6: K<6>::go(a...); break;
        default: K<7>::go(a...); break;
    }
}

template <int SH>
struct PushL {
    static void go(const LG23& L, const Grid23& g, const Species23& s, int sidx, double* xh, double* yh,
                   bool centred, hipStream_t st) {
        if (centred) k_lg_push<SH, 1><<<chunks23(s.P), kTile23Threads, 0, st>>>(L, g, s, s, sidx, xh, yh);
        else k_lg_push<SH, 0><<<chunks23(s.P), kTile23Threads, 0, st>>>(L, g, s, s, sidx, xh, yh);
    }
};
template <int SH>
struct DepRhoL {
    static void go(const LG23& L, const Grid23& g, const Species23& s, const double* xh, const double* yh,
                   hipStream_t st) {
        k_lg_deposit<SH, 1, false><<<chunks23(s.P), kTile23Threads, 0, st>>>(L, g, s, xh, yh);
    }
};
template <int SH>
struct DepJL {
    static void go(const LG23& L, const Grid23& g, const Species23& s, hipStream_t st) {
        k_lg_deposit<SH, 3, false><<<chunks23(s.P), kTile23Threads, 0, st>>>(L, g, s, s.x, s.y);
    }
};

void lg_launch_push(const LG23& L, const Grid23& g, const Species23& s, int sidx, double* xh, double* yh,
                    bool centred, hipStream_t st) {
    if (s.P > 0) by_shape<PushL>(s.shape, L, g, s, sidx, xh, yh, centred, st);
}

void lg_launch_deposit_rho(const LG23& L, const Grid23& g, const Species23& s, const double* xh, const double* yh,
                           hipStream_t st) {
    if (s.P > 0) by_shape<DepRhoL>(s.shape, L, g, s, xh, yh, st);
}

void lg_launch_deposit_j(const LG23& L, const Grid23& g, const Species23& s, hipStream_t st) {
    if (s.P > 0) by_shape<DepJL>(s.shape, L, g, s, st);
}

void lg_launch_finish(const LG23& L, bool centred, hipStream_t st) {
    if (centred) k_lg_finish<2><<<grid_blocks((int64_t)L.NX * L.NY), 256, 0, st>>>(L);
    else k_lg_finish<0><<<grid_blocks((int64_t)L.NX * L.NY), 256, 0, st>>>(L);
}

void lg_launch_advect(const LG23& L, const Grid23& g, const Species23& s, double h, hipStream_t st) {
    (void)L;
    if (s.P > 0) k_lg_advect<<<grid_blocks(s.P), 256, 0, st>>>(g, s, h);
}

void lg_launch_vmax(const LG23& L, const Species23& s, int sidx, bool clear, hipStream_t st) {
    if (clear) k_lg_zero_umax<<<1, 64, 0, st>>>(L);
    if (s.P > 0) k_lg_vmax<<<grid_blocks(s.P), 256, 0, st>>>(L, s, sidx);
}

void lg_launch_warm_finish(const LG23& L, hipStream_t st) {
    k_lg_finish<1><<<grid_blocks((int64_t)L.NX * L.NY), 256, 0, st>>>(L);
}

// ---------------------------------------------------------------------------
// diagnose!
// ---------------------------------------------------------------------------
// mean(abs2, EBxyz) over the ghosted 6 x NXp x NYp array: element e is
// component e % 6 of padded cell e / 6, whose periodic image is interior.
__global__ void __launch_bounds__(256) k_lg_diag_field(LG23 L, int32_t buffer, int64_t nf) {
    __shared__ double red[kTreeLanes + 8];
    const int32_t NXp = L.NX + 2 * buffer;
    const int64_t G = (int64_t)L.NX * L.NY;
    const int64_t s0 = (int64_t)blockIdx.x * kTreeChunk;
    const int64_t n = nf - s0 < kTreeChunk ? nf - s0 : kTreeChunk;
    const double r = block_chunk_sum(red, n, [&](int64_t i) {
        const int64_t e = s0 + i;
        const int64_t q = e / 6;
        const int c = (int)(e - 6 * q);
        const int32_t jp = (int32_t)(q / NXp), ip = (int32_t)(q - (int64_t)jp * NXp);
        const int32_t ci = (ip - buffer) & (L.NX - 1), cj = (jp - buffer) & (L.NY - 1);
        const double v = L.EB[c * G + ci + (int64_t)L.NX * cj];
        return v * v;
    });
    if (threadIdx.x == 0) L.fpart[blockIdx.x] = r;
}

// the field momentum's three sums over the cells (real parts of E and B)
__global__ void __launch_bounds__(256) k_lg_diag_fmom(LG23 L, int64_t nch) {
    __shared__ double red[kTreeLanes + 8];
    const int64_t G = (int64_t)L.NX * L.NY;
    const int64_t s0 = (int64_t)blockIdx.x * kTreeChunk;
    const int64_t n = G - s0 < kTreeChunk ? G - s0 : kTreeChunk;
    const double2* F = L.F;
    for (int c = 0; c < 3; ++c) {
        const double r = block_chunk_sum(red, n, [&](int64_t i) {
            const int64_t k = s0 + i;
            const double Ex = F[kLGEx * G + k].x, Ey = F[kLGEy * G + k].x, Ez = F[kLGEz * G + k].x;
            const double Bx = F[kLGBx * G + k].x, By = F[kLGBy * G + k].x, Bz = F[kLGBz * G + k].x;
            return c == 0 ? Ey * Bz - Ez * By : (c == 1 ? Ez * Bx - Ex * Bz : Ex * By - Ey * Bx);
        });
        if (threadIdx.x == 0) L.mfpart[c * nch + blockIdx.x] = r;
    }
}

struct LGDiagFinal {
    int ns;
    int64_t nf, nfc, G, nmc;
    int64_t nk[kMaxSpecies23], nm[kMaxSpecies23];
    double mass[kMaxSpecies23], weight[kMaxSpecies23];
};

__global__ void __launch_bounds__(256) k_lg_diag_final(LG23 L, Diag23 d, LGDiagFinal f) {
    __shared__ double red[kTreeLanes + 8];
    const double fe = tree_strided(red, L.fpart, 1, f.nfc) / (double)f.nf / 2.0;
    double fm[3];
    for (int c = 0; c < 3; ++c) fm[c] = tree_strided(red, L.mfpart + c * f.nmc, 1, f.nmc) / (double)f.G;
    double ke = 0.0, pm[3] = {0.0, 0.0, 0.0}, pc[3] = {0.0, 0.0, 0.0};
    for (int s = 0; s < f.ns; ++s) {
        const double kes = tree_strided(red, d.kpart[s], 1, f.nk[s]) * f.mass[s] / 2.0 * f.weight[s];
        double ms[3], cs[3];
        for (int c = 0; c < 3; ++c) {
            ms[c] = tree_strided(red, d.mpart[s] + c, 6, f.nm[s]) * (f.mass[s] * f.weight[s]);
            cs[c] = tree_strided(red, d.mpart[s] + 3 + c, 6, f.nm[s]) * (f.mass[s] * f.weight[s]);
        }
        if (s == 0) {
            ke = kes;
            for (int c = 0; c < 3; ++c) {
                pm[c] = ms[c];
                pc[c] = cs[c];
            }
        } else {
            ke = ke + kes;
            for (int c = 0; c < 3; ++c) {
                pm[c] = pm[c] + ms[c];
                pc[c] = pc[c] + cs[c];
            }
        }
    }
    if (threadIdx.x == 0) {
        const int64_t b = *d.t / d.ntskip;
        if (b < L.cap) {
            double* o = L.rows + kLGRow * b;
            o[0] = ke;
            o[1] = fe;
            for (int c = 0; c < 3; ++c) {
                o[2 + c] = pm[c];
                o[5 + c] = fm[c];
                o[8 + c] = pc[c];
            }
        }
    }
}

void lg_launch_diag_row(const LG23& L, const Grid23& g, const Diag23& d, const Species23* sp, int ns,
                        hipStream_t st) {
    (void)g;
    const int64_t nf = 6 * (int64_t)(L.NX + 2 * d.buffer) * (L.NY + 2 * d.buffer);
    const int64_t G = (int64_t)L.NX * L.NY;
    LGDiagFinal f{};
    f.ns = ns;
    f.nf = nf;
    f.nfc = (nf + kTreeChunk - 1) / kTreeChunk;
    f.G = G;
    f.nmc = (G + kTreeChunk - 1) / kTreeChunk;
    k_lg_diag_field<<<(unsigned)f.nfc, 256, 0, st>>>(L, d.buffer, nf);
    k_lg_diag_fmom<<<(unsigned)f.nmc, 256, 0, st>>>(L, f.nmc);
    for (int s = 0; s < ns; ++s) {
        f.nk[s] = (3 * sp[s].P + kTreeChunk - 1) / kTreeChunk;
        f.nm[s] = (sp[s].P + kTreeChunk - 1) / kTreeChunk;
        f.mass[s] = sp[s].mass;
        f.weight[s] = sp[s].weight;
        launch23_diag_species(sp[s], d.kpart[s], d.mpart[s], st);
    }
    k_lg_diag_final<<<1, 256, 0, st>>>(L, d, f);
}

// bins[a][t/ntskip][il, jl] += real(Z_a[il*g + ii, jl*g + jj]) * factor, jj
// outer, ii inner (:1370-1386)
__global__ void __launch_bounds__(256) k_lg_bins(LG23 L, Diag23 d) {
    const int64_t b = *d.t / d.ntskip;
    if (b >= L.cap) return;
    const int NXd = d.NXd, NYd = d.NYd, gs = d.ngskip, NX = L.NX;
    const int64_t per = (int64_t)NXd * NYd;
    const int64_t G = (int64_t)NX * L.NY;
    const double factor = 1.0 / (double)((int64_t)d.ntskip * gs * gs);
    const int slots[kLGBins] = {kLGEx, kLGEy, kLGEz, kLGBx, kLGBy, kLGBz, kLGAx0,
                                kLGAy0, kLGAz0, kLGPhi0, kLGRho0, kLGJx0, kLGJy0, kLGJz0};
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < kLGBins * per;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int a = (int)(e / per);
        const int64_t r = e - a * per;
        const int jl = (int)(r / NXd), il = (int)(r - (int64_t)jl * NXd);
        const double2* z = L.F + (size_t)slots[a] * G;
        double* o = L.bins + ((size_t)a * L.cap + b) * per + r;
        double acc = *o;
        for (int jj = 0; jj < gs; ++jj)
            for (int ii = 0; ii < gs; ++ii) {
                const int64_t k = (int64_t)il * gs + ii + (int64_t)NX * ((int64_t)jl * gs + jj);
                acc = acc + z[k].x * factor;
            }
        *o = acc;
    }
}

void lg_launch_diag_fields(const LG23& L, const Diag23& d, hipStream_t st) {
    const int64_t G = (int64_t)L.NX * L.NY;
    LGList lst{};
    const int tr[8] = {kLGAx0, kLGAy0, kLGAz0, kLGPhi0, kLGRho0, kLGJx0, kLGJy0, kLGJz0};
    for (int a = 0; a < 8; ++a) lst.a[a] = L.F + (size_t)tr[a] * G;
    lg_inverse(lst, 8, L, false, st);
    k_lg_bins<<<grid_blocks(kLGBins * (int64_t)d.NXd * d.NYd), 256, 0, st>>>(L, d);
}

// ---------------------------------------------------------------------------
// LorenzGaugeSemiImplicitField (:403-461, warmup! :864-871, loop! :993-1100)
// ---------------------------------------------------------------------------
// The deposit grid is the reference's buffered one: Julia cell c (1-based,
// unwrapped, axes -(b-1):NX+b) is row c - 1 + b of the (NX+2b)-wide grid.
// The deposits (k_lg_deposit<.., GHOST>) add every contribution to the
// periodic limbs racc, as the other fields' deposits do, and a contribution
// to a ghost cell also to that cell of eacc.  The periodic reduction is then
// racc itself; an interior buffered cell's counters are its periodic cell's
// minus those of the cell's ghost images (exact: the counters are sums
// modulo 2^64), a ghost cell's are eacc's.

// The value of buffered cell e of component comp (see above)
__device__ __forceinline__ double si_cell(const LG23& L, int64_t Ge, int comp, int64_t e, const int* q) {
    const int NX = L.NX, NY = L.NY, NXe = L.NXe, NYe = L.NYe, b = L.b;
    const int v = (int)(e / NXe), u = (int)(e - (int64_t)v * NXe);
    const u64* ge = L.eacc + (size_t)comp * 3 * Ge;
    if ((unsigned)(u - b) >= (unsigned)NX || (unsigned)(v - b) >= (unsigned)NY)
        return lg_limbs(ge, Ge, e, q[comp]);
    const int64_t G = (int64_t)NX * NY;
    const int64_t k = (u - b) + (int64_t)NX * (v - b);
    const u64* r = L.racc + (size_t)comp * 3 * G;
    u64 a0 = r[k], a1 = r[G + k], a2 = r[2 * G + k];
    // the ghost images (u + m NX, v + n NY) of the cell; the cell itself holds 0
    for (int vv = v & (NY - 1); vv < NYe; vv += NY)
        for (int uu = u & (NX - 1); uu < NXe; uu += NX) {
            const int64_t f = uu + (int64_t)NXe * vv;
            a0 -= ge[f];
            a1 -= ge[Ge + f];
            a2 -= ge[2 * Ge + f];
        }
    return lg_limbs_v(a0, a1, a2, q[comp]);
}

// The grid rhoJs+ from the limbs into gridP (element i = component i % 4 of
// buffered cell i / 4, the reference's order) and, TEST, the chunk partials
// of norm(gridT - gridP), norm(gridT), norm(gridP): canonical sums of squares
// and the maxima of |.| (generic_norm2).
template <bool TEST>
__global__ void __launch_bounds__(256) k_lg_si_norm(LG23 L, int64_t nch) {
    __shared__ double red[3 * kTreeLanes + 8];
    const int64_t Ge = (int64_t)L.NXe * L.NYe, n = 4 * Ge;
    const int64_t s0 = (int64_t)blockIdx.x * kTreeChunk;
    const int64_t m = n - s0 < kTreeChunk ? n - s0 : kTreeChunk;
    int q[4];
    q[0] = L.qr;
    for (int c = 0; c < 3; ++c) q[c + 1] = lg_frac_bits(L, c);
    double sd = 0.0, st = 0.0, sp = 0.0, md = 0.0, mt = 0.0, mp = 0.0;
    if (threadIdx.x < kTreeLanes)
        for (int64_t i = threadIdx.x; i < m; i += kTreeLanes) {
            const int64_t idx = s0 + i;
            const double p = si_cell(L, Ge, (int)(idx & 3), idx >> 2, q);
            L.gridP[idx] = p;
            if (TEST) {
                const double tv = L.gridT[idx], d = tv - p;
                sd = sd + d * d;
                st = st + tv * tv;
                sp = sp + p * p;
                md = nanmax(md, fabs(d));
                mt = nanmax(mt, fabs(tv));
                mp = nanmax(mp, fabs(p));
            }
        }
    if (!TEST) return;
    tree256x3<false>(red, sd, st, sp);
    tree256x3<true>(red, md, mt, mp);
    if (threadIdx.x == 0) {
        const double v[6] = {sd, st, sp, md, mt, mp};
        for (int k = 0; k < 6; ++k) L.npart[k * nch + blockIdx.x] = v[k];
    }
}

// isapprox(gridT, gridP; rtol, atol = 0) (:1000) as LinearAlgebra's array
// isapprox: d = norm(T - P) finite -> d <= rtol*max(norm(T), norm(P)), else
// elementwise.  One block; res = (d, norm T, norm P, converged, eflag).
__global__ void __launch_bounds__(256) k_lg_si_test(LG23 L, int64_t nch) {
    __shared__ double red[kTreeLanes + 8];
    __shared__ int bad;
    const int t = threadIdx.x;
    const int64_t Ge = (int64_t)L.NXe * L.NYe, n = 4 * Ge;
    double sums[3], maxs[3], nrm[3];
    for (int k = 0; k < 3; ++k) sums[k] = tree_strided(red, L.npart + (size_t)k * nch, 1, nch);
    for (int k = 0; k < 3; ++k) {
        double mm = 0.0;
        if (t < kTreeLanes)
            for (int64_t i = t; i < nch; i += kTreeLanes) mm = nanmax(mm, L.npart[(size_t)(3 + k) * nch + i]);
        maxs[k] = tree256<true>(red, mm);
    }
    for (int k = 0; k < 3; ++k) {
        const double m = maxs[k];
        if (m == 0.0 || isinf(m)) {
            nrm[k] = m;
        } else if (isfinite((double)n * m * m) && m * m != 0.0) {
            nrm[k] = sqrt(sums[k]);
        } else {
            // scaled: m*sqrt(sum (|x|/m)^2), the partials' storage reused
            for (int64_t c = 0; c < nch; ++c) {
                const int64_t s0 = c * kTreeChunk, mc = n - s0 < kTreeChunk ? n - s0 : kTreeChunk;
                const double r = block_chunk_sum(red, mc, [&](int64_t i) {
                    const double tv = L.gridT[s0 + i], pv = L.gridP[s0 + i];
                    const double x = k == 0 ? tv - pv : (k == 1 ? tv : pv);
                    const double qv = fabs(x) / m;
                    return qv * qv;
                });
                if (t == 0) L.npart[(size_t)k * nch + c] = r;
                __syncthreads();
            }
            nrm[k] = m * sqrt(tree_strided(red, L.npart + (size_t)k * nch, 1, nch));
        }
    }
    const double dn = nrm[0], rtol = L.rtol;
    int conv;
    if (isfinite(dn)) {
        const double nf = nrm[1], ne = nrm[2];
        double mx = nf > ne ? nf : ne;
        if (nf != nf || ne != ne) mx = NAN;
        const double tol = rtol * mx;
        double lim = 0.0 > tol ? 0.0 : tol;
        if (tol != tol) lim = tol;
        conv = dn <= lim;
    } else {
        if (t == 0) bad = 0;
        __syncthreads();
        for (int64_t i = t; i < n; i += blockDim.x) {
            const double a = L.gridT[i], c = L.gridP[i];
            if (a == c) continue;
            if (!(isfinite(a) && isfinite(c))) {
                bad = 1;
                continue;
            }
            const double ma = fabs(a) > fabs(c) ? fabs(a) : fabs(c);
            const double tt = rtol * ma;
            if (!(fabs(a - c) <= (0.0 > tt ? 0.0 : tt))) bad = 1;
        }
        __syncthreads();
        conv = !bad;
    }
    if (t == 0) {
        L.res[0] = dn;
        L.res[1] = nrm[1];
        L.res[2] = nrm[2];
        L.res[3] = conv ? 1.0 : 0.0;
        L.res[4] = (double)*L.eflag;
    }
}

// reduction! (:504-522): each periodic cell's exact sum over its images in
// the buffered grid (the periodic limbs), rounded once, into out (imaginary
// parts 0); the limbs, periodic and ghost, are cleared.
__global__ void __launch_bounds__(256) k_lg_si_wrap(LG23 L, LGOut out) {
    const int NX = L.NX, NY = L.NY, NXe = L.NXe, NYe = L.NYe, b = L.b;
    const int64_t G = (int64_t)NX * NY, Ge = (int64_t)NXe * NYe;
    int q[4];
    q[0] = L.qr;
    for (int c = 0; c < 3; ++c) q[c + 1] = lg_frac_bits(L, c);
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < G; k += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(k & (NX - 1)), j = (int)(k >> L.lx);
        const int u0 = (i + b) & (NX - 1), v0 = (j + b) & (NY - 1);
#pragma unroll
        for (int comp = 0; comp < 4; ++comp) {
            u64* r = L.racc + (size_t)comp * 3 * G;
            out.a[comp][k] = make_double2(lg_limbs(r, G, k, q[comp]), 0.0);
            r[k] = 0;
            r[G + k] = 0;
            r[2 * G + k] = 0;
            u64* ge = L.eacc + (size_t)comp * 3 * Ge;
            for (int v = v0; v < NYe; v += NY)
                for (int u = u0; u < NXe; u += NX) {
                    const int64_t e = u + (int64_t)NXe * v;
                    ge[e] = 0;
                    ge[Ge + e] = 0;
                    ge[2 * Ge + e] = 0;
                }
        }
    }
}

// lorenzgauge! with the three source weights (:638-648) for phi (rho) and A
// (J), then E and B with the field theta (:1066-1077).  src is in the
// spectrum; its [1,1] is cleared here (:1055-1059).
__global__ void __launch_bounds__(256) k_lg_spectral_si(LG23 L) {
    const int NX = L.NX;
    const int64_t G = (int64_t)NX * L.NY;
    const double dt = L.dt, dt2 = dt * dt, dt_2 = 2.0 * dt, thf = L.theta, ths = L.theta_s;
    const double hs = ths / 2.0, ws = 1.0 - ths, hf = thf / 2.0, wf = 1.0 - thf;
    double2* F = L.F;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < G; k += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(k & (NX - 1)), j = (int)(k >> L.lx);
        const double kxi = L.kx[i], kyj = L.ky[j];
        const double k2 = kxi * kxi + kyj * kyj;
        const double num = 2.0 - dt2 * k2 * (1.0 - thf);
        const double den = 1.0 + dt2 * k2 * thf / 2.0;
        const int x0s[4] = {kLGPhi0, kLGAx0, kLGAy0, kLGAz0};
        const int xms[4] = {kLGPhiM, kLGAxM, kLGAyM, kLGAzM};
        const int s0s[4] = {kLGRho0, kLGJx0, kLGJy0, kLGJz0};
        const int sms[4] = {kLGRhoM, kLGJxM, kLGJyM, kLGJzM};
        double2 pl[4], cc[4], xm[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const double2 sm = F[sms[a] * G + k], s0 = F[s0s[a] * G + k];
            double2 sp = L.src[a * G + k];
            if (k == 0) {
                sp = make_double2(0.0, 0.0);
                L.src[a * G] = sp;
            }
            const double2 x0 = F[x0s[a] * G + k];
            xm[a] = F[xms[a] * G + k];
            const double svr = hs * sm.x + ws * s0.x + hs * sp.x;
            const double svi = hs * sm.y + ws * s0.y + hs * sp.y;
            pl[a].x = (num * x0.x + dt2 * svr) / den - xm[a].x;
            pl[a].y = (num * x0.y + dt2 * svi) / den - xm[a].y;
            cc[a].x = hf * (pl[a].x + xm[a].x) + wf * x0.x;
            cc[a].y = hf * (pl[a].y + xm[a].y) + wf * x0.y;
            L.plus[a * G + k] = pl[a];
        }
        double er, ei, fr, fi;
        cmul_lg(0.0 * kxi, -kxi, cc[0].x, cc[0].y, er, ei);
        F[kLGEx * G + k] = make_double2(er - (pl[1].x - xm[1].x) / dt_2, ei - (pl[1].y - xm[1].y) / dt_2);
        cmul_lg(0.0 * kyj, -kyj, cc[0].x, cc[0].y, er, ei);
        F[kLGEy * G + k] = make_double2(er - (pl[2].x - xm[2].x) / dt_2, ei - (pl[2].y - xm[2].y) / dt_2);
        F[kLGEz * G + k] = make_double2(-(pl[3].x - xm[3].x) / dt_2, -(pl[3].y - xm[3].y) / dt_2);
        cmul_lg(0.0 * kyj, kyj, cc[3].x, cc[3].y, er, ei);
        F[kLGBx * G + k] = make_double2(er, ei);
        cmul_lg(0.0 * kxi, -kxi, cc[3].x, cc[3].y, er, ei);
        F[kLGBy * G + k] = make_double2(er, ei);
        cmul_lg(0.0 * kxi, kxi, cc[2].x, cc[2].y, er, ei);
        cmul_lg(0.0 * kyj, kyj, cc[1].x, cc[1].y, fr, fi);
        F[kLGBz * G + k] = make_double2(er - fr, ei - fi);
    }
}

// phi- = phi0, phi0 = phi+, A- = A0, A0 = A+ (:1089-1100)
__global__ void __launch_bounds__(256) k_lg_rotate(LG23 L) {
    const int64_t G = (int64_t)L.NX * L.NY;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < G; k += (int64_t)gridDim.x * blockDim.x) {
        const int x0s[4] = {kLGPhi0, kLGAx0, kLGAy0, kLGAz0};
        const int xms[4] = {kLGPhiM, kLGAxM, kLGAyM, kLGAzM};
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            L.F[xms[a] * G + k] = L.F[x0s[a] * G + k];
            L.F[x0s[a] * G + k] = L.plus[a * G + k];
        }
    }
}

template <int SH>
struct PushSIL {
    static void go(const LG23& L, const Grid23& g, const Species23& s, const Species23& o, int sidx, hipStream_t st) {
        k_lg_push<SH, 2><<<chunks23(s.P), kTile23Threads, 0, st>>>(L, g, s, o, sidx, nullptr, nullptr);
    }
};
template <int SH>
struct DepSIL {
    static void go(const LG23& L, const Grid23& g, const Species23& s, bool charge, hipStream_t st) {
        if (charge) k_lg_deposit<SH, 1, true><<<chunks23(s.P), kTile23Threads, 0, st>>>(L, g, s, s.x, s.y);
        k_lg_deposit<SH, 3, true><<<chunks23(s.P), kTile23Threads, 0, st>>>(L, g, s, s.x, s.y);
    }
};

void lg_launch_si_push(const LG23& L, const Grid23& g, const Species23& s, const Species23& o, int sidx,
                       hipStream_t st) {
    if (s.P > 0) by_shape<PushSIL>(s.shape, L, g, s, o, sidx, st);
}

void lg_launch_si_deposit(const LG23& L, const Grid23& g, const Species23& s, bool charge, hipStream_t st) {
    if (s.P > 0) by_shape<DepSIL>(s.shape, L, g, s, charge, st);
}

void lg_launch_si_reduce(const LG23& L, const LGOut& out, bool test, hipStream_t st) {
    const int64_t n = 4 * (int64_t)L.NXe * L.NYe, nch = (n + kTreeChunk - 1) / kTreeChunk;
    if (test) {
        k_lg_si_norm<true><<<(unsigned)nch, 256, 0, st>>>(L, nch);
        k_lg_si_test<<<1, 256, 0, st>>>(L, nch);
    } else {
        k_lg_si_norm<false><<<(unsigned)nch, 256, 0, st>>>(L, nch);
    }
    k_lg_si_wrap<<<grid_blocks((int64_t)L.NX * L.NY), 256, 0, st>>>(L, out);
}

void lg_launch_si_solve(const LG23& L, hipStream_t st) {
    const int64_t G = (int64_t)L.NX * L.NY;
    LGList src{};
    for (int c = 0; c < 4; ++c) src.a[c] = L.src + (size_t)c * G;
    lg_forward(src, 4, L, st);
    k_lg_spectral_si<<<grid_blocks(G), 256, 0, st>>>(L);
    LGList lst{};
    for (int c = 0; c < 6; ++c) lst.a[c] = L.F + (size_t)(kLGEx + c) * G;
    lg_inverse(lst, 6, L, true, st);
}

void lg_launch_rotate(const LG23& L, hipStream_t st) {
    k_lg_rotate<<<grid_blocks((int64_t)L.NX * L.NY), 256, 0, st>>>(L);
}

void lg_launch_zero_umax(const LG23& L, hipStream_t st) { k_lg_zero_umax<<<1, 64, 0, st>>>(L); }

int configure_kernels_2d3v_lg() {
    const void* ks[] = {reinterpret_cast<const void*>(k_lg_cols<true>), reinterpret_cast<const void*>(k_lg_cols<false>),
                        reinterpret_cast<const void*>(k_lg_rows<true, true>),
                        reinterpret_cast<const void*>(k_lg_rows<true, false>),
                        reinterpret_cast<const void*>(k_lg_rows<false, false>)};
    for (const void* k : ks) (void)hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    (void)hipGetLastError();
    return 0;
}

}  // namespace pic
