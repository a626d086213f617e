// kernels_2d3v.hip -- electrostatic path of the structured 2D3V module
// (src/PIC2D3V.jl): species of any charge and mass in an Lx x Ly box, the
// NGP, area and B-spline (orders 0..5) shapes, Boris about a uniform B of any
// direction, exact signed per-cell charge sums, update! of the gathered
// field, the electrostatic diagnostics and the per-cell sort.
//
// FP64, evaluated in the reference's expression order (contraction off; the
// gather's @muladd is an explicit fma).  Charge contributions are rounded to
// the 2^-q grid and summed exactly in three 38-bit two's-complement limbs per
// cell (device-scope integer atomics), so the charge does not depend on the
// particle order or the schedule.  The field solve is the 2D transform of
// kernels_2d.hip with wavenumbers 2pi/Lx, 2pi/Ly.
#include <climits>

#include "pic_device.h"
#include "pic2d3v_internal.h"
#include "shapes23.h"

#pragma clang fp contract(off)

namespace pic {

// The shapes, Boris forms, signed limbs and push-window helpers are in
// shapes23.h.

// ---------------------------------------------------------------------------
// the particle loop of loop! (:545-555) for one species
//
// A block takes a chunk of kTile23Chunk consecutive particles.  Particles
// sorted per cell (sort!, as the reference's driver does before its loop)
// keep a chunk within a thin strip of cells, so the block stages the field of
// the strip (its window: the chunk's cells plus the stencil and one cell of
// drift) in LDS and sums the charge of the strip in LDS limbs, adding them to
// the grid's limbs once at the end.  A particle whose stencil leaves the
// window (or every particle, when the window would not fit) reads the field
// and adds its charge through global memory instead.  The charge sums are
// exact integers and each particle's update depends only on its own state and
// the field, so the result does not depend on the window.
// ---------------------------------------------------------------------------

// Occupancy: the quadratic B-spline's loop is capped at 128 registers, four
// waves per SIMD (the LDS window allows four blocks per CU).
template <int SH, int BV>
__global__ void __launch_bounds__(kTile23Threads) __attribute__((amdgpu_waves_per_eu(SH == 4 ? 4 : 1)))
k23_push(Grid23 g, Species23 s, const double2* __restrict__ E, u64* __restrict__ racc) {
    constexpr int n = ShapeCount<SH>::value;
    constexpr int LO = ShapeExtent<SH>::lo, HI = ShapeExtent<SH>::hi;
    __shared__ double2 Ew[kTile23Cells];
    __shared__ u64 Rw[3 * kTile23Cells];
    __shared__ int red[4][4];
    __shared__ int win[5];
    const int t = threadIdx.x;
    const int NX = g.NX, NY = g.NY;
    const int64_t G = (int64_t)NX * NY;
    const int64_t c0 = (int64_t)blockIdx.x * kTile23Chunk;
    // the window: the chunk's base cells relative to its first particle's
    const int ax = base_cell(s.x[c0], g.NX_Lx, NX), ay = base_cell(s.y[c0], g.NY_Ly, NY);
    double xs[kTile23PPT], ys[kTile23PPT];
    int mnx = INT_MAX, mxx = INT_MIN, mny = INT_MAX, mxy = INT_MIN;
#pragma unroll
    for (int k = 0; k < kTile23PPT; ++k) {
        const int64_t p = c0 + t + (int64_t)k * kTile23Threads;
        xs[k] = 0.0;
        ys[k] = 0.0;
        if (p < s.P) {
            xs[k] = s.x[p];
            ys[k] = s.y[p];
            const int dx = ((base_cell(xs[k], g.NX_Lx, NX) - ax + NX / 2) & (NX - 1)) - NX / 2;
            const int dy = ((base_cell(ys[k], g.NY_Ly, NY) - ay + NY / 2) & (NY - 1)) - NY / 2;
            mnx = min(mnx, dx);
            mxx = max(mxx, dx);
            mny = min(mny, dy);
            mxy = max(mxy, dy);
        }
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
        win[4] = wx * wy <= kTile23Cells;
    }
    __syncthreads();
    const int ox = win[0], oy = win[1], wx = win[2], wy = win[3];
    const bool tile = win[4] != 0;
    if (tile) {
        for (int e = t; e < wx * wy; e += kTile23Threads) {
            const int ry = e / wx, rx = e - ry * wx;
            Ew[e] = E[((ox + rx) & (NX - 1)) + (int64_t)NX * ((oy + ry) & (NY - 1))];
            Rw[e] = 0;
            Rw[kTile23Cells + e] = 0;
            Rw[2 * kTile23Cells + e] = 0;
        }
    }
    __syncthreads();
    int fb = 0;  // particles that read or added through global memory
#pragma unroll 1
    for (int k = 0; k < kTile23PPT; ++k) {
        const int64_t p = c0 + t + (int64_t)k * kTile23Threads;
        if (p >= s.P) break;
        double x = xs[k], y = ys[k], vx = s.vx[p], vy = s.vy[p], vz = s.vz[p];
        int32_t cx[n], cy[n];
        double wxs[n], wys[n];
        int lx[n], ly[n];
        // gather at the old position (the window and the fallback read
        // through separate loads, LDS and global, not one generic pointer)
        shape_w<SH>(x, g.NX_Lx, NX, cx, wxs);
        shape_w<SH>(y, g.NY_Ly, NY, cy, wys);
        bool in = tile;
#pragma unroll
        for (int q = 0; q < n; ++q) {
            lx[q] = (cx[q] - ox) & (NX - 1);
            ly[q] = (cy[q] - oy) & (NY - 1);
            in = in && lx[q] < wx && ly[q] < wy;
        }
        double a = 0.0, b = 0.0;
        if (in) {
#pragma unroll
            for (int jj = 0; jj < n; ++jj)
#pragma unroll
                for (int ii = 0; ii < n; ++ii) {
                    const double wxy = wxs[ii] * wys[jj];
                    const double2 e = Ew[lx[ii] + wx * ly[jj]];
                    a = fma(e.x, wxy, a);
                    b = fma(e.y, wxy, b);
                }
        } else {
#pragma unroll
            for (int jj = 0; jj < n; ++jj)
#pragma unroll
                for (int ii = 0; ii < n; ++ii) {
                    const double wxy = wxs[ii] * wys[jj];
                    const double2 e = E[cx[ii] + (int64_t)NX * cy[jj]];
                    a = fma(e.x, wxy, a);
                    b = fma(e.y, wxy, b);
                }
        }
        fb += in ? 0 : 1;
        const double vx0 = vx, vy0 = vy;
        boris23<BV>(g, s.q_m, a, b, vx, vy, vz);
        x = unimod23(x + (vx0 + vx) / 2.0 * g.dt, g.Lx);
        y = unimod23(y + (vy0 + vy) / 2.0 * g.dt, g.Ly);
        s.x[p] = x;
        s.y[p] = y;
        s.vx[p] = vx;
        s.vy[p] = vy;
        s.vz[p] = vz;
        // deposit at the new position
        shape_w<SH>(x, g.NX_Lx, NX, cx, wxs);
        shape_w<SH>(y, g.NY_Ly, NY, cy, wys);
        in = tile;
#pragma unroll
        for (int q = 0; q < n; ++q) {
            lx[q] = (cx[q] - ox) & (NX - 1);
            ly[q] = (cy[q] - oy) & (NY - 1);
            in = in && lx[q] < wx && ly[q] < wy;
        }
        fb += in ? 0 : 1;
        if (in) {
#pragma unroll
            for (int jj = 0; jj < n; ++jj)
#pragma unroll
                for (int ii = 0; ii < n; ++ii)
                    add_limbs_lds(Rw, lx[ii] + wx * ly[jj], fix_round128(wxs[ii] * wys[jj] * s.qw, g.q));
        } else {
#pragma unroll
            for (int jj = 0; jj < n; ++jj)
#pragma unroll
                for (int ii = 0; ii < n; ++ii)
                    add_limbs_s(racc, G, cx[ii] + (int64_t)NX * cy[jj], fix_round128(wxs[ii] * wys[jj] * s.qw, g.q));
        }
    }
    count_fallback(g.fallback, fb);
    __syncthreads();
    if (tile) {
        for (int e = t; e < wx * wy; e += kTile23Threads) {
            const int ry = e / wx, rx = e - ry * wx;
            const int64_t k = ((ox + rx) & (NX - 1)) + (int64_t)NX * ((oy + ry) & (NY - 1));
            const u64 l0 = Rw[e], l1 = Rw[kTile23Cells + e], l2 = Rw[2 * kTile23Cells + e];
            if (l0) atomicAdd(racc + k, l0);
            if (l1) atomicAdd(racc + G + k, l1);
            if (l2) atomicAdd(racc + 2 * G + k, l2);
        }
    }
}

template <int SH>
__global__ void __launch_bounds__(kPush23Threads)
k23_deposit_op(Grid23 g, const double* __restrict__ X, const double* __restrict__ Y, int64_t n, double w,
               u64* __restrict__ racc) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += stride)
        deposit23<SH>(g, racc, X[p], Y[p], w);
}

template <int SH>
__global__ void __launch_bounds__(kPush23Threads)
k23_gather_op(Grid23 g, const double2* __restrict__ E, const double* __restrict__ X, const double* __restrict__ Y,
              int64_t n, double* __restrict__ ex, double* __restrict__ ey) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += stride) {
        double a, b;
        gather23<SH>(g, E, X[p], Y[p], a, b);
        ex[p] = a;
        ey[p] = b;
    }
}

template <int BV>
__global__ void __launch_bounds__(kPush23Threads)
k23_boris_op(Grid23 g, double q_m, double* __restrict__ VX, double* __restrict__ VY, double* __restrict__ VZ,
             const double* __restrict__ ex, const double* __restrict__ ey, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += stride) {
        double vx = VX[p], vy = VY[p], vz = VZ[p];
        boris23<BV>(g, q_m, ex[p], ey[p], vx, vy, vz);
        VX[p] = vx;
        VY[p] = vy;
        VZ[p] = vz;
    }
}

__global__ void __launch_bounds__(256) k23_limbs_to_rho(int64_t G, int q, u64* __restrict__ racc,
                                                        double* __restrict__ rho) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < G; k += (int64_t)gridDim.x * blockDim.x) {
        const u64 a = racc[k], b = racc[G + k], c = racc[2 * G + k];
        racc[k] = 0;
        racc[G + k] = 0;
        racc[2 * G + k] = 0;
        u128 v = (u128)a + ((u128)b << kLimbS) + ((u128)c << (2 * kLimbS));
        v &= ((u128)1 << (3 * kLimbS)) - 1;
        const i128 sv = (v >> (3 * kLimbS - 1)) ? (i128)v - ((i128)1 << (3 * kLimbS)) : (i128)v;
        rho[k] = i128_to_double(sv, q);
    }
}

// ---------------------------------------------------------------------------
// update! (:294-297) and the binned fields of diagnose! (:1323-1327)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k23_update(Grid23 g, Diag23 d, const double* __restrict__ Ex, const double* __restrict__ Ey,
           const double* __restrict__ phi, double2* __restrict__ E, float2* __restrict__ E32, int accumulate) {
    const int64_t G = (int64_t)g.NX * g.NY;
    const int64_t b = *d.t / d.ntskip;
    const bool bin = b < d.cap;
    const double nt = (double)d.ntskip;
    const int32_t gm = d.ngskip - 1;
    const int lg = __ffs(d.ngskip) - 1;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < G; k += (int64_t)gridDim.x * blockDim.x) {
        const double ex = Ex[k], ey = Ey[k];
        const double2 e = accumulate ? E[k] : make_double2(0.0, 0.0);
        const double2 en = make_double2(e.x + ex, e.y + ey);
        E[k] = en;
        if (E32) E32[k] = make_float2((float)en.x, (float)en.y);  // the mixed gathers' copy
        const int32_t i = (int32_t)(k & (g.NX - 1)), j = (int32_t)(k >> g.lx);
        if (bin && ((i | j) & gm) == 0) {
            const int64_t o = b * (int64_t)d.NXd * d.NYd + (i >> lg) + (int64_t)d.NXd * (j >> lg);
            d.Exs[o] = d.Exs[o] + ex / nt;
            d.Eys[o] = d.Eys[o] + ey / nt;
            d.phis[o] = d.phis[o] + phi[k] / nt;
        }
    }
}

__global__ void k23_tick(int64_t* t) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *t += 1;
}

// ---------------------------------------------------------------------------
// diagnostics (:1301-1321) with the canonical tree
// ---------------------------------------------------------------------------
// mean(abs2, Exy) over the ghosted 2 x NXp x NYp array: element e is
// component e & 1 of padded cell e >> 1, whose periodic image is interior.
__global__ void __launch_bounds__(256)
k23_diag_field(Grid23 g, Diag23 d, const double2* __restrict__ E, int64_t nf) {
    __shared__ double red[kTreeLanes + 8];
    const int32_t NXp = g.NX + 2 * d.buffer;
    const int64_t s = (int64_t)blockIdx.x * kTreeChunk;
    const int64_t n = nf - s < kTreeChunk ? nf - s : kTreeChunk;
    const double r = block_chunk_sum(red, n, [&](int64_t i) {
        const int64_t e = s + i;
        const int32_t q = (int32_t)(e >> 1);
        const int32_t jp = q / NXp, ip = q - jp * NXp;
        const int32_t ci = (ip - d.buffer) & (g.NX - 1), cj = (jp - d.buffer) & (g.NY - 1);
        const double2 v = E[ci + (int64_t)g.NX * cj];
        const double a = (e & 1) ? v.y : v.x;
        return a * a;
    });
    if (threadIdx.x == 0) d.fpart[blockIdx.x] = r;
}

// sum(abs2, velocities(s)) over the 3P entries in xyv's column order.  A
// block's chunk of entries covers about a third as many particles: they are
// read once each (caller index -> slot -> vx, vy, vz) into LDS as the chunk's
// squares, then summed by the canonical tree in entry order.
__global__ void __launch_bounds__(256)
k23_diag_kinetic(Species23 s, double* __restrict__ kpart) {
    __shared__ double red[kTreeLanes + 8];
    __shared__ double sq[kTreeChunk];
    const int64_t n3 = 3 * s.P;
    const int64_t b0 = (int64_t)blockIdx.x * kTreeChunk;
    const int64_t n = n3 - b0 < kTreeChunk ? n3 - b0 : kTreeChunk;
    const int64_t p0 = b0 / 3, p1 = (b0 + n - 1) / 3;  // the particles of entries b0 .. b0 + n - 1
    for (int64_t p = p0 + threadIdx.x; p <= p1; p += blockDim.x) {
        const int64_t h = s.slot ? (int64_t)s.slot[p] : p;  // caller index p's slot
        const double v[3] = {s.vx[h], s.vy[h], s.vz[h]};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int64_t e = 3 * p + c - b0;
            if (e >= 0 && e < n) sq[e] = v[c] * v[c];
        }
    }
    __syncthreads();
    const double r = block_chunk_sum(red, n, [&](int64_t i) { return sq[i]; });
    if (threadIdx.x == 0) kpart[blockIdx.x] = r;
}

// momentum and characteristic momentum: six sums over P
__global__ void __launch_bounds__(256)
k23_diag_momentum(Species23 s, double* __restrict__ mpart) {
    __shared__ double red[3 * kTreeLanes + 8];
    const int64_t b0 = (int64_t)blockIdx.x * kTreeChunk;
    const int64_t n = s.P - b0 < kTreeChunk ? s.P - b0 : kTreeChunk;
    double a = 0.0, b = 0.0, c = 0.0, e = 0.0, f = 0.0, h = 0.0;
    if (threadIdx.x < kTreeLanes)
        for (int64_t i = threadIdx.x; i < n; i += kTreeLanes) {
            const int64_t hs = s.slot ? (int64_t)s.slot[b0 + i] : b0 + i;  // caller index b0 + i's slot
            const double vx = s.vx[hs], vy = s.vy[hs], vz = s.vz[hs];
            a = a + vx;
            b = b + vy;
            c = c + vz;
            e = e + fabs(vx);
            f = f + fabs(vy);
            h = h + fabs(vz);
        }
    tree256x3<false>(red, a, b, c);
    tree256x3<false>(red, e, f, h);
    if (threadIdx.x == 0) {
        double* o = mpart + 6 * blockIdx.x;
        o[0] = a;
        o[1] = b;
        o[2] = c;
        o[3] = e;
        o[4] = f;
        o[5] = h;
    }
}

struct DiagSpecies23 {
    int64_t nk, nm;  // chunks of the kinetic and momentum sums
    double mass, weight;
};
struct DiagFinal23 {
    int ns;
    int64_t nf, nfc;  // field entries, chunks
    DiagSpecies23 s[kMaxSpecies23];
};

__global__ void __launch_bounds__(256) k23_diag_final(Diag23 d, DiagFinal23 f) {
    __shared__ double red[kTreeLanes + 8];
    const double fe = tree_strided(red, d.fpart, 1, f.nfc) / (double)f.nf / 2.0;
    double ke = 0.0, pm[3] = {0.0, 0.0, 0.0}, pc[3] = {0.0, 0.0, 0.0};
    for (int s = 0; s < f.ns; ++s) {
        const DiagSpecies23& q = f.s[s];
        const double kes = tree_strided(red, d.kpart[s], 1, q.nk) * q.mass / 2.0 * q.weight;
        double ms[3], cs[3];
        for (int c = 0; c < 3; ++c) {
            ms[c] = tree_strided(red, d.mpart[s] + c, 6, q.nm) * (q.mass * q.weight);
            cs[c] = tree_strided(red, d.mpart[s] + 3 + c, 6, q.nm) * (q.mass * q.weight);
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
        if (b < d.cap) {
            double* o = d.rows + 8 * b;
            o[0] = ke;
            o[1] = fe;
            for (int c = 0; c < 3; ++c) {
                o[2 + c] = pm[c];
                o[5 + c] = pc[c];
            }
        }
    }
}

// ---------------------------------------------------------------------------
// xyv (5 x P, column-major) <-> SoA
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k23_unpack(const double* __restrict__ xyv, Species23 s) {
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < s.P; p += (int64_t)gridDim.x * blockDim.x) {
        const double* q = xyv + 5 * p;
        s.x[p] = q[0];
        s.y[p] = q[1];
        s.vx[p] = q[2];
        s.vy[p] = q[3];
        s.vz[p] = q[4];
    }
}

// caller index p from its held slot
__global__ void __launch_bounds__(256) k23_pack(Species23 s, double* __restrict__ xyv) {
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < s.P; p += (int64_t)gridDim.x * blockDim.x) {
        double* q = xyv + 5 * p;
        const int64_t h = s.slot ? (int64_t)s.slot[p] : p;
        q[0] = s.x[h];
        q[1] = s.y[h];
        q[2] = s.vx[h];
        q[3] = s.vy[h];
        q[4] = s.vz[h];
    }
}

// ---------------------------------------------------------------------------
// Regrouping.  The deposits' LDS windows cover a chunk of consecutive held
// particles; particles drift out of the order a sort left them in, and a
// chunk then outgrows its window and falls back to global atomics.  Every few
// steps the held particles are regrouped by 8 x 8-cell tile (a counting sort
// by tile; the order inside a tile is arbitrary).  Every result is
// independent of the held order (exact deposit sums, per-particle updates),
// and the diagnostics' particle sums and the downloads read through slot, in
// the caller's order.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int32_t tile23(const Grid23& g, const Tiles23& T, double x, double y) {
    return base_cell(x, g.NX_Lx, g.NX) / T.TX + T.ntx * (base_cell(y, g.NY_Ly, g.NY) / T.TY);
}

__global__ void __launch_bounds__(256) k23_regroup_count(Grid23 g, Tiles23 T, Species23 s, int64_t per_block) {
    extern __shared__ int32_t hist[];
    for (int i = threadIdx.x; i < T.ntiles; i += blockDim.x) hist[i] = 0;
    __syncthreads();
    const int64_t b0 = (int64_t)blockIdx.x * per_block;
    const int64_t b1 = b0 + per_block < s.P ? b0 + per_block : s.P;
    for (int64_t p = b0 + threadIdx.x; p < b1; p += blockDim.x) atomicAdd(&hist[tile23(g, T, s.x[p], s.y[p])], 1);
    __syncthreads();
    for (int i = threadIdx.x; i < T.ntiles; i += blockDim.x)
        if (hist[i]) atomicAdd(&T.count[i], hist[i]);
}

// one block: exclusive scan of the counts into start and the cursors; the
// counts cleared for the next regrouping
__global__ void __launch_bounds__(1024) k23_regroup_scan(Tiles23 T) {
    __shared__ int32_t part[1024];
    const int per = (T.ntiles + blockDim.x - 1) / blockDim.x;
    const int i0 = threadIdx.x * per;
    int32_t sum = 0;
    for (int i = i0; i < i0 + per && i < T.ntiles; ++i) sum += T.count[i];
    part[threadIdx.x] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        int32_t run = 0;
        for (int k = 0; k < (int)blockDim.x; ++k) {
            const int32_t v = part[k];
            part[k] = run;
            run += v;
        }
        T.start[T.ntiles] = run;
    }
    __syncthreads();
    int32_t run = part[threadIdx.x];
    for (int i = i0; i < i0 + per && i < T.ntiles; ++i) {
        T.start[i] = run;
        T.cursor[i] = run;
        run += T.count[i];
        T.count[i] = 0;
    }
}

// every held particle to a free slot of its tile's range in o; pido and the
// slot of each caller index follow
__global__ void __launch_bounds__(256)
k23_regroup_scatter(Grid23 g, Tiles23 T, Species23 s, Species23 o, const uint32_t* __restrict__ pid,
                    uint32_t* __restrict__ pido, int64_t per_block) {
    extern __shared__ int32_t hist[];
    for (int i = threadIdx.x; i < T.ntiles; i += blockDim.x) hist[i] = 0;
    __syncthreads();
    const int64_t b0 = (int64_t)blockIdx.x * per_block;
    const int64_t b1 = b0 + per_block < s.P ? b0 + per_block : s.P;
    for (int64_t p = b0 + threadIdx.x; p < b1; p += blockDim.x) atomicAdd(&hist[tile23(g, T, s.x[p], s.y[p])], 1);
    __syncthreads();
    for (int i = threadIdx.x; i < T.ntiles; i += blockDim.x)
        if (hist[i]) hist[i] = atomicAdd(&T.cursor[i], hist[i]);
    __syncthreads();
    for (int64_t p = b0 + threadIdx.x; p < b1; p += blockDim.x) {
        const double x = s.x[p], y = s.y[p];
        const int32_t d = atomicAdd(&hist[tile23(g, T, x, y)], 1);
        const uint32_t id = pid[p];
        o.x[d] = x;
        o.y[d] = y;
        o.vx[d] = s.vx[p];
        o.vy[d] = s.vy[p];
        o.vz[d] = s.vz[p];
        pido[d] = id;
        s.slot[id] = (uint32_t)d;
    }
}

// the held particles of s in the caller's order into o
__global__ void __launch_bounds__(256) k23_to_caller(Species23 s, Species23 o) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < s.P; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t h = s.slot[i];
        o.x[i] = s.x[h];
        o.y[i] = s.y[h];
        o.vx[i] = s.vx[h];
        o.vy[i] = s.vy[h];
        o.vz[i] = s.vz[h];
    }
}

__global__ void __launch_bounds__(256) k23_identity(uint32_t* __restrict__ pid, uint32_t* __restrict__ slot, int64_t P) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < P; i += (int64_t)gridDim.x * blockDim.x) {
        pid[i] = (uint32_t)i;
        slot[i] = (uint32_t)i;
    }
}

__global__ void __launch_bounds__(256) k23_pair(const double* __restrict__ Ex, const double* __restrict__ Ey,
                                                double2* __restrict__ E, int64_t G) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < G; k += (int64_t)gridDim.x * blockDim.x)
        E[k] = make_double2(Ex[k], Ey[k]);
}

__global__ void __launch_bounds__(256) k23_split(const double2* __restrict__ E, double* __restrict__ Ex,
                                                 double* __restrict__ Ey, int64_t G) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < G; k += (int64_t)gridDim.x * blockDim.x) {
        const double2 e = E[k];
        if (Ex) Ex[k] = e.x;
        if (Ey) Ey[k] = e.y;
    }
}

// ---------------------------------------------------------------------------
// sort!(s, dx, dy) (:210-215): bucket by cell (ceil(x/dx), ceil(y/dy)) with x
// major, then order each bucket by (vx under isless, index).  Every
// particle's final slot is its bucket start plus the number of bucket members
// before it in that order, so the result is the stable sortperm whatever the
// order of the scatter.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int64_t ceil_index(double a) {
    const double c = ceil(a);
    return (c > -4.0e18 && c < 4.0e18) ? (int64_t)c : 0;
}

// isless on Float64 as an unsigned key: -0.0 < 0.0, every NaN last and equal
__device__ __forceinline__ u64 order_key(double v) {
    if (v != v) return ~0ull;
    const u64 b = (u64)__double_as_longlong(v);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

__global__ void __launch_bounds__(256)
k23_sort_extent(Species23 s, double dx, double dy, long long* __restrict__ ext) {
    __shared__ long long red[4][256];
    long long a = LLONG_MAX, b = LLONG_MIN, c = LLONG_MAX, e = LLONG_MIN;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < s.P; p += (int64_t)gridDim.x * blockDim.x) {
        const long long cx = ceil_index(s.x[p] / dx), cy = ceil_index(s.y[p] / dy);
        a = cx < a ? cx : a;
        b = cx > b ? cx : b;
        c = cy < c ? cy : c;
        e = cy > e ? cy : e;
    }
    const int t = threadIdx.x;
    red[0][t] = a;
    red[1][t] = b;
    red[2][t] = c;
    red[3][t] = e;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (t < h) {
            red[0][t] = red[0][t + h] < red[0][t] ? red[0][t + h] : red[0][t];
            red[1][t] = red[1][t + h] > red[1][t] ? red[1][t + h] : red[1][t];
            red[2][t] = red[2][t + h] < red[2][t] ? red[2][t + h] : red[2][t];
            red[3][t] = red[3][t + h] > red[3][t] ? red[3][t + h] : red[3][t];
        }
        __syncthreads();
    }
    if (t == 0) {
        atomicMin(ext + 0, red[0][0]);
        atomicMax(ext + 1, red[1][0]);
        atomicMin(ext + 2, red[2][0]);
        atomicMax(ext + 3, red[3][0]);
    }
}

__global__ void __launch_bounds__(256)
k23_sort_count(Species23 s, double dx, double dy, const long long* __restrict__ ext, int64_t ry,
               uint32_t* __restrict__ bucket, u64* __restrict__ vkey, uint32_t* __restrict__ count) {
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < s.P; p += (int64_t)gridDim.x * blockDim.x) {
        const int64_t cx = ceil_index(s.x[p] / dx) - ext[0], cy = ceil_index(s.y[p] / dy) - ext[2];
        const uint32_t b = (uint32_t)(cx * ry + cy);
        bucket[p] = b;
        vkey[p] = order_key(s.vx[p]);
        atomicAdd(count + b, 1u);
    }
}

// Exclusive scan of count[0..nb) into start[0..nb] and the largest count, by
// one block in tiles of 4096; clears the scatter cursors.
__global__ void __launch_bounds__(1024)
k23_sort_scan(const uint32_t* __restrict__ count, uint32_t* __restrict__ start, uint32_t* __restrict__ cursor,
              int64_t nb, uint32_t* __restrict__ maxcount) {
    __shared__ uint32_t sums[1024];
    __shared__ uint32_t carry_s;
    const int t = threadIdx.x;
    uint32_t carry = 0, mx = 0;
    for (int64_t base = 0; base < nb; base += 4096) {
        uint32_t v[4], loc = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int64_t i = base + 4 * t + k;
            v[k] = i < nb ? count[i] : 0u;
            mx = v[k] > mx ? v[k] : mx;
            loc += v[k];
        }
        sums[t] = loc;
        __syncthreads();
        for (int h = 1; h < 1024; h <<= 1) {  // inclusive Hillis-Steele scan of the thread sums
            const uint32_t a = t >= h ? sums[t - h] : 0u;
            __syncthreads();
            sums[t] += a;
            __syncthreads();
        }
        uint32_t run = carry + (t > 0 ? sums[t - 1] : 0u);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int64_t i = base + 4 * t + k;
            if (i < nb) {
                start[i] = run;
                cursor[i] = 0u;
            }
            run += v[k];
        }
        if (t == 1023) carry_s = carry + sums[1023];
        __syncthreads();
        carry = carry_s;
        __syncthreads();
    }
    if (t == 0) start[nb] = carry;
    sums[t] = mx;
    __syncthreads();
    for (int h = 512; h > 0; h >>= 1) {
        if (t < h) sums[t] = sums[t + h] > sums[t] ? sums[t + h] : sums[t];
        __syncthreads();
    }
    if (t == 0) *maxcount = sums[0];
}

__global__ void __launch_bounds__(256)
k23_sort_scatter(int64_t P, const uint32_t* __restrict__ bucket, const u64* __restrict__ vkey,
                 const uint32_t* __restrict__ start, uint32_t* __restrict__ cursor, u64* __restrict__ skey,
                 uint32_t* __restrict__ sidx, uint32_t* __restrict__ sbucket) {
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < P; p += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t b = bucket[p];
        const uint32_t slot = start[b] + atomicAdd(cursor + b, 1u);
        skey[slot] = vkey[p];
        sidx[slot] = (uint32_t)p;
        sbucket[slot] = b;
    }
}

__global__ void __launch_bounds__(256)
k23_sort_rank(int64_t P, const uint32_t* __restrict__ start, const u64* __restrict__ skey,
              const uint32_t* __restrict__ sidx, const uint32_t* __restrict__ sbucket, uint32_t* __restrict__ perm) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < P; e += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t b = sbucket[e];
        const uint32_t s0 = start[b], s1 = start[b + 1];
        const u64 k = skey[e];
        const uint32_t i = sidx[e];
        uint32_t rank = 0;
        for (uint32_t f = s0; f < s1; ++f) {
            const u64 kf = skey[f];
            const uint32_t jf = sidx[f];
            rank += (kf < k || (kf == k && jf < i)) ? 1u : 0u;
        }
        perm[s0 + rank] = i;
    }
}

struct Dst23 {
    double* a[5];
};

__global__ void __launch_bounds__(256) k23_sort_apply(Species23 s, const uint32_t* __restrict__ perm, Dst23 d) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < s.P; k += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t p = perm[k];
        d.a[0][k] = s.x[p];
        d.a[1][k] = s.y[p];
        d.a[2][k] = s.vx[p];
        d.a[3][k] = s.vy[p];
        d.a[4][k] = s.vz[p];
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static unsigned blocks_for(int64_t n, int threads) {
    const int64_t b = (n + threads - 1) / threads;
    return (unsigned)(b < 16384 ? (b > 0 ? b : 1) : 16384);
}

#define PIC23_SHAPES(X) X(0) X(1) X(2) X(3) X(4) X(5) X(6) X(7)

void launch23_push(const Grid23& g, const Species23& s, const double2* E, unsigned long long* racc, hipStream_t st) {
    const unsigned nb = (unsigned)((s.P + kTile23Chunk - 1) / kTile23Chunk);
#define PIC23_PUSH(SH)                                                                                 \
    case SH:                                                                                           \
        if (g.boris)                                                                                   \
            k23_push<SH, 1><<<nb, kTile23Threads, 0, st>>>(g, s, E, racc);                             \
        else                                                                                           \
            k23_push<SH, 0><<<nb, kTile23Threads, 0, st>>>(g, s, E, racc);                             \
        break;
    switch (s.shape) { PIC23_SHAPES(PIC23_PUSH) }
#undef PIC23_PUSH
}

void launch23_deposit_op(const Grid23& g, int shape, const double* x, const double* y, int64_t n, double w,
                         unsigned long long* racc, hipStream_t st) {
    const unsigned nb = blocks_for(n, kPush23Threads);
#define PIC23_DEP(SH)                                                                \
    case SH:                                                                         \
        k23_deposit_op<SH><<<nb, kPush23Threads, 0, st>>>(g, x, y, n, w, racc);      \
        break;
    switch (shape) { PIC23_SHAPES(PIC23_DEP) }
#undef PIC23_DEP
}

void launch23_limbs_to_rho(const Grid23& g, unsigned long long* racc, double* rho, hipStream_t st) {
    const int64_t G = (int64_t)g.NX * g.NY;
    k23_limbs_to_rho<<<blocks_for(G, 256), 256, 0, st>>>(G, g.q, racc, rho);
}

void launch23_gather_op(const Grid23& g, int shape, const double2* E, const double* x, const double* y, int64_t n,
                        double* ex, double* ey, hipStream_t st) {
    const unsigned nb = blocks_for(n, kPush23Threads);
#define PIC23_GAT(SH)                                                                \
    case SH:                                                                         \
        k23_gather_op<SH><<<nb, kPush23Threads, 0, st>>>(g, E, x, y, n, ex, ey);     \
        break;
    switch (shape) { PIC23_SHAPES(PIC23_GAT) }
#undef PIC23_GAT
}

void launch23_boris_op(const Grid23& g, double q_m, double* vx, double* vy, double* vz, const double* ex,
                       const double* ey, int64_t n, hipStream_t st) {
    const unsigned nb = blocks_for(n, kPush23Threads);
    if (g.boris)
        k23_boris_op<1><<<nb, kPush23Threads, 0, st>>>(g, q_m, vx, vy, vz, ex, ey, n);
    else
        k23_boris_op<0><<<nb, kPush23Threads, 0, st>>>(g, q_m, vx, vy, vz, ex, ey, n);
}

void launch23_update(const Grid23& g, const Diag23& d, const Buffers2D& b, double2* E, int accumulate,
                     hipStream_t st) {
    launch23_update32(g, d, b, E, nullptr, accumulate, st);
}

void launch23_update32(const Grid23& g, const Diag23& d, const Buffers2D& b, double2* E, float2* E32,
                       int accumulate, hipStream_t st) {
    const int64_t G = (int64_t)g.NX * g.NY;
    k23_update<<<blocks_for(G, 256), 256, 0, st>>>(g, d, b.Ex, b.Ey, b.phi, E, E32, accumulate);
}

void launch23_tick(int64_t* t, hipStream_t st) { k23_tick<<<1, 64, 0, st>>>(t); }

void launch23_diag_species(const Species23& s, double* kpart, double* mpart, hipStream_t st) {
    const int64_t nk = (3 * s.P + kTreeChunk - 1) / kTreeChunk, nm = (s.P + kTreeChunk - 1) / kTreeChunk;
    k23_diag_kinetic<<<(unsigned)nk, 256, 0, st>>>(s, kpart);
    k23_diag_momentum<<<(unsigned)nm, 256, 0, st>>>(s, mpart);
}

void launch23_diag(const Grid23& g, const Diag23& d, const double2* E, const Species23* sp, int ns, hipStream_t st) {
    const int64_t nf = 2 * (int64_t)(g.NX + 2 * d.buffer) * (g.NY + 2 * d.buffer);
    DiagFinal23 f{};
    f.ns = ns;
    f.nf = nf;
    f.nfc = (nf + kTreeChunk - 1) / kTreeChunk;
    k23_diag_field<<<(unsigned)f.nfc, 256, 0, st>>>(g, d, E, nf);
    for (int s = 0; s < ns; ++s) {
        f.s[s].nk = (3 * sp[s].P + kTreeChunk - 1) / kTreeChunk;
        f.s[s].nm = (sp[s].P + kTreeChunk - 1) / kTreeChunk;
        f.s[s].mass = sp[s].mass;
        f.s[s].weight = sp[s].weight;
        if (sp[s].mixed) {
            f.s[s].nk = f.s[s].nm = launch23m_diag_species(sp[s], d.kpart[s], d.mpart[s], st);
            continue;
        }
        k23_diag_kinetic<<<(unsigned)f.s[s].nk, 256, 0, st>>>(sp[s], d.kpart[s]);
        k23_diag_momentum<<<(unsigned)f.s[s].nm, 256, 0, st>>>(sp[s], d.mpart[s]);
    }
    k23_diag_final<<<1, 256, 0, st>>>(d, f);
}

void launch23_unpack(const double* xyv, const Species23& s, hipStream_t st) {
    k23_unpack<<<blocks_for(s.P, 256), 256, 0, st>>>(xyv, s);
}

void launch23_pack(const Species23& s, double* xyv, hipStream_t st) {
    k23_pack<<<blocks_for(s.P, 256), 256, 0, st>>>(s, xyv);
}

void launch23_regroup(const Grid23& g, const Tiles23& T, const Species23& s, const Species23& o, const uint32_t* pid,
                      uint32_t* pido, hipStream_t st) {
    if (s.P < 1) return;
    const size_t lds = sizeof(int32_t) * (size_t)T.ntiles;
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(k23_regroup_count),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(k23_regroup_scatter),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    const int64_t per = 16384;
    const unsigned nb = (unsigned)((s.P + per - 1) / per);
    k23_regroup_count<<<nb, 256, lds, st>>>(g, T, s, per);
    launch23_regroup_scan(T, st);
    k23_regroup_scatter<<<nb, 256, lds, st>>>(g, T, s, o, pid, pido, per);
}

void launch23_regroup_scan(const Tiles23& T, hipStream_t st) { k23_regroup_scan<<<1, 1024, 0, st>>>(T); }

void launch23_to_caller(const Species23& s, const Species23& o, hipStream_t st) {
    if (s.P > 0) k23_to_caller<<<blocks_for(s.P, 256), 256, 0, st>>>(s, o);
}

void launch23_identity(uint32_t* pid, uint32_t* slot, int64_t P, hipStream_t st) {
    if (P > 0) k23_identity<<<blocks_for(P, 256), 256, 0, st>>>(pid, slot, P);
}

void launch23_pair_field(const double* Ex, const double* Ey, double2* E, int64_t G, hipStream_t st) {
    k23_pair<<<blocks_for(G, 256), 256, 0, st>>>(Ex, Ey, E, G);
}

void launch23_split_field(const double2* E, double* Ex, double* Ey, int64_t G, hipStream_t st) {
    k23_split<<<blocks_for(G, 256), 256, 0, st>>>(E, Ex, Ey, G);
}

void launch23_sort_extent(const Species23& s, double dx, double dy, Sort23& w, hipStream_t st) {
    k23_sort_extent<<<blocks_for(s.P, 256), 256, 0, st>>>(s, dx, dy, reinterpret_cast<long long*>(w.ext));
}

void launch23_sort_buckets(const Species23& s, double dx, double dy, int64_t ry, int64_t nb, Sort23& w,
                           hipStream_t st) {
    k23_sort_count<<<blocks_for(s.P, 256), 256, 0, st>>>(s, dx, dy, reinterpret_cast<const long long*>(w.ext), ry,
                                                         w.bucket, w.vkey, w.count);
    k23_sort_scan<<<1, 1024, 0, st>>>(w.count, w.start, w.cursor, nb, w.maxcount);
}

void launch23_sort_order(const Species23& s, int64_t nb, Sort23& w, hipStream_t st) {
    (void)nb;
    k23_sort_scatter<<<blocks_for(s.P, 256), 256, 0, st>>>(s.P, w.bucket, w.vkey, w.start, w.cursor, w.skey, w.sidx,
                                                           w.sbucket);
    k23_sort_rank<<<blocks_for(s.P, 256), 256, 0, st>>>(s.P, w.start, w.skey, w.sidx, w.sbucket, w.perm);
}

void launch23_sort_apply(const Species23& s, const uint32_t* perm, double* const* dst, hipStream_t st) {
    Dst23 d;
    for (int k = 0; k < 5; ++k) d.a[k] = dst[k];
    k23_sort_apply<<<blocks_for(s.P, 256), 256, 0, st>>>(s, perm, d);
}

}  // namespace pic
