// kernels_2d3v_mixed.hip -- the mixed-precision electrostatic particle loop of
// the structured 2D3V module (loop! of src/PIC2D3V.jl:537-555) and the moves
// of its state.
//
// State: positions as 32-bit fractions of the box (x = xq*Lx/2^32, xq = 0
// meaning x = Lx, so the periodic wrap is integer overflow), velocities as
// FP32 hi + lo pairs.  The field is gathered from an FP32 copy of the FP64
// Exy with FP32 shape weights; the Boris update is computed in FP32 as a
// change added to the pairs; the move is rint(((v0 + v)/2)*dt*2^32/L) quanta.
// Charge: each axis' shape weights are rounded to integer quanta of 2^-22 of
// the particle (summing exactly to 2^22), their products are added per cell
// into 64-bit LDS words and then into the species' 128-bit per-cell counters,
// so the deposit is exact and independent of the particle order.  The same
// quanta weight the gather.
#include <climits>

#include "pic_device.h"
#include "pic2d3v_internal.h"
#include "shapes23.h"

// FP32 arithmetic here may contract into fmas (the FP64 helpers of the
// included header switch contraction off for their own expressions)
#pragma clang fp contract(fast)

namespace pic {

// ---------------------------------------------------------------------------
// shapes on fixed-point coordinates
// ---------------------------------------------------------------------------

// Uniform B-spline weights of degree n at the fraction u: W[m] = M_n(u + m),
// m = 0..n, with M_n the B-spline on [0, n+1], by the Cox-de Boor recursion
//   M_d(t) = (t M_{d-1}(t) + (d+1-t) M_{d-1}(t-1)) / d,
// whose terms are all products of nonnegative numbers (no cancellation in
// FP32).  W[m] = B_n(u + m - (n+1)/2) for the centred B-spline B_n, which is
// the reference's weight m at x = u (bspline, :1124-1162).
template <int n>
__device__ __forceinline__ void bspline_uf(float u, float* W) {
    W[0] = 1.0f;
#pragma unroll
    for (int d = 1; d <= n; ++d) {
        const float inv = 1.0f / (float)d;
        W[d] = ((1.0f - u) * W[d - 1]) * inv;
#pragma unroll
        for (int m = d - 1; m >= 1; --m) W[m] = ((u + (float)m) * W[m] + ((float)(d + 1 - m) - u) * W[m - 1]) * inv;
        W[0] = (u * W[0]) * inv;
    }
}

// Weights in quanta of 2^-kQ23 that sum to exactly 2^kQ23: the rounded
// prefix sums, kept nondecreasing, differenced; the last weight takes the
// remainder.
template <int N>
__device__ __forceinline__ void quantize23(const float* w, uint32_t* q) {
    constexpr int32_t one = 1 << kQ23;
    float acc = 0.0f;
    int32_t prev = 0;
#pragma unroll
    for (int k = 0; k < N - 1; ++k) {
        acc += w[k];
        int32_t c = (int32_t)rintf(acc * (float)one);
        c = c < prev ? prev : (c > one ? one : c);
        q[k] = (uint32_t)(c - prev);
        prev = c;
    }
    q[N - 1] = (uint32_t)(one - prev);
}

// the product of two numbers below 2^24 (the masks let the compiler use the
// full-rate 24-bit multiplies)
__device__ __forceinline__ u64 quanta_product(uint32_t a, uint32_t b) {
    return (u64)(a & 0xFFFFFFu) * (b & 0xFFFFFFu);
}

// depositindicesfractions + the shape (:1106-1187) on zq: zN = zq/2^sh, the
// base cell c = ceil(zN) - 1 = (zq - 1) >> sh and r = ceil(zN) - zN =
// ((c + 1) << sh) - zq in units of 2^-sh (both exact in 32-bit wrapping
// arithmetic, zq = 0 giving c = N - 1, r = 0).  Returns the first stencil
// cell (unwrapped: the caller masks it) and the n weights in quanta.
template <int SH>
__device__ __forceinline__ int32_t shape_q23(uint32_t zq, int sh, uint32_t* q) {
    const uint32_t c = (zq - 1u) >> sh;
    const uint32_t r = ((c + 1u) << sh) - zq;
    if constexpr (SH == 0) {
        q[0] = 1u << kQ23;
        return (int32_t)c;
    } else if constexpr (SH == 1) {
        const int d = sh - kQ23;
        const uint32_t rq = (r + ((1u << d) >> 1)) >> d;
        q[0] = (1u << kQ23) - rq;
        q[1] = rq;
        return (int32_t)c;
    } else if constexpr (SH == 4) {
        // the quadratic B-spline in integers: x = (up + 1/2) - r on 24 bits
        // (u), weights x^2/2, 1 - x^2/2 - (1-x)^2/2, (1-x)^2/2 rounded to
        // quanta from the exact 48-bit squares (1 - x = (v + 1)/2^24 with
        // v = 2^24 - 1 - u, so that v fits the 24-bit multiplier)
        const uint32_t up = r > (1u << (sh - 1)) ? 1u : 0u;
        const uint32_t xi = ((1u + 2u * up) << (sh - 1)) - r;  // x in units of 2^-sh, [0, 2^sh)
        const uint32_t u = sh >= 24 ? xi >> (sh - 24) : xi << (24 - sh);
        const uint32_t v = 0xFFFFFFu - u;
        const u64 uu = quanta_product(u, u);
        const u64 vv = quanta_product(v, v) + 2u * (u64)v + 1u;
        constexpr u64 half = 1ull << (48 + 1 - kQ23 - 1);
        const uint32_t q0 = (uint32_t)((uu + half) >> (48 + 1 - kQ23));
        const uint32_t q2 = (uint32_t)((vv + half) >> (48 + 1 - kQ23));
        q[0] = q0;
        q[1] = (1u << kQ23) - q0 - q2;
        q[2] = q2;
        return (int32_t)c + (int32_t)up - 1;
    } else {
        constexpr int n = SH - 2;
        const float sc = __int_as_float((127 - sh) << 23);  // 2^-sh
        float u;
        int32_t first;
        if constexpr (n & 1) {
            u = (float)((1u << sh) - r) * sc;  // x = 1 - r
            first = (int32_t)c - (n - 1) / 2;
        } else {
            const uint32_t up = r > (1u << (sh - 1)) ? 1u : 0u;  // r > 1/2
            u = (float)(((1u + 2u * up) << (sh - 1)) - r) * sc;  // x = (up + 1/2) - r
            first = (int32_t)c + (int32_t)up - n / 2;
        }
        float w[n + 1];
        bspline_uf<n>(u, w);
        quantize23<n + 1>(w, q);
        return first;
    }
}

// v += d on a hi + lo pair: the two-sum of hi and d, its error folded into
// lo, renormalised.
__device__ __forceinline__ void hilo_add23(float& hi, float& lo, float d) {
    const float s = hi + d;
    const float bb = s - hi;
    const float err = (hi - (s - bb)) + (d - bb);
    const float t = lo + err;
    hi = s + t;
    lo = t - (hi - s);
}

// rint((hi + lo + d) * step) as a 32-bit wrapping move, step = st[0] + st[1]:
// hi*st[0] is split exactly (an fma gives its rounding error) and its integer
// part taken first, so the rounding of the fraction sees every term.
__device__ __forceinline__ uint32_t move_q23(float hi, float lo, float d, const float* st) {
    const float p = hi * st[0];
    const float pe = fmaf(hi, st[0], -p);
    const float ai = rintf(p);
    const float rest = (p - ai) + (pe + (hi * st[1] + (lo + d) * st[0]));
    return (uint32_t)(int64_t)ai + (uint32_t)(int32_t)rintf(rest);
}

// an exact 128-bit per-cell sum: lo words at k, hi words at G + k; every add
// that wraps the lo word carries one into the hi word
__device__ __forceinline__ void add_q128(u64* __restrict__ cnt, int64_t G, int64_t k, u64 v) {
    const u64 old = atomicAdd(cnt + k, v);
    if (old + v < old) atomicAdd(cnt + G + k, 1ull);
}

// ---------------------------------------------------------------------------
// the particle loop for one species.  As the FP64 loop, a block takes a chunk
// of kTile23Chunk consecutive held particles, stages the field of the cells
// about them in LDS and sums their charge there; a particle whose stencil
// leaves the window (every particle, when the window would not fit) reads and
// adds through global memory instead, with the same result.
// ---------------------------------------------------------------------------
// Occupancy: the quadratic B-spline's loop holds 84 registers by default (5
// waves per SIMD); capped at 76 it runs 6 waves without spilling.
template <int SH>
__global__ void __launch_bounds__(kTile23Threads) __attribute__((amdgpu_waves_per_eu(SH == 4 ? 6 : 1)))
k23m_push(Grid23 g, Species23 s, const float2* __restrict__ E32) {
    constexpr int n = ShapeCount<SH>::value;
    constexpr int LO = ShapeExtent<SH>::lo, HI = ShapeExtent<SH>::hi;
    constexpr float kW = 1.0f / (float)(1u << kQ23);
    __shared__ float2 Ew[kTile23Cells];
    __shared__ u64 Rw[kTile23Cells];
    __shared__ uint32_t Xs[kTile23Chunk], Ys[kTile23Chunk];  // the chunk's positions, read once
    __shared__ int red[kTile23Threads / 64][4];
    __shared__ int win[5];
    const int t = threadIdx.x;
    const int NX = g.NX, NY = g.NY;
    const int64_t G = (int64_t)NX * NY;
    u64* __restrict__ cnt = s.cnt;
    const int64_t c0 = (int64_t)blockIdx.x * kTile23Chunk;
    const int ax = (int)((s.xq[c0] - 1u) >> g.sx), ay = (int)((s.yq[c0] - 1u) >> g.sy);
    int mnx = INT_MAX, mxx = INT_MIN, mny = INT_MAX, mxy = INT_MIN;
#pragma unroll
    for (int k = 0; k < kTile23PPT; ++k) {
        const int64_t p = c0 + t + (int64_t)k * kTile23Threads;
        if (p < s.P) {
            const uint32_t xq = s.xq[p], yq = s.yq[p];
            Xs[t + k * kTile23Threads] = xq;
            Ys[t + k * kTile23Threads] = yq;
            const int dx = (((int)((xq - 1u) >> g.sx) - ax + NX / 2) & (NX - 1)) - NX / 2;
            const int dy = (((int)((yq - 1u) >> g.sy) - ay + NY / 2) & (NY - 1)) - NY / 2;
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
            Ew[e] = E32[((ox + rx) & (NX - 1)) + (int64_t)NX * ((oy + ry) & (NY - 1))];
            Rw[e] = 0;
        }
    }
    __syncthreads();
    const float t1 = s.bt[0], t2 = s.bt[1], t3 = s.bt[2];
    int fb = 0;  // stencils read or added through global memory
#pragma unroll 1
    for (int k = 0; k < kTile23PPT; ++k) {
        const int64_t p = c0 + t + (int64_t)k * kTile23Threads;
        if (p >= s.P) break;
        uint32_t xq = Xs[t + k * kTile23Threads], yq = Ys[t + k * kTile23Threads];
        float vx = s.vxh[p], vy = s.vyh[p], vz = s.vzh[p];
        float lx = s.vxl[p], ly = s.vyl[p], lz = s.vzl[p];
        uint32_t qx[n], qy[n];
        // gather at the old position.  The stencil's cells are consecutive
        // (wrapped), so it lies in the window when its first cell's window
        // offset leaves room for the rest; the window and the fallback read
        // through separate loads (LDS, global), not one generic pointer.
        int32_t fx = shape_q23<SH>(xq, g.sx, qx), fy = shape_q23<SH>(yq, g.sy, qy);
        int32_t rx = (fx - ox) & (NX - 1), ry = (fy - oy) & (NY - 1);
        bool in = tile && rx + n <= wx && ry + n <= wy;
        float wgx[n], wgy[n];
#pragma unroll
        for (int q = 0; q < n; ++q) {
            wgx[q] = (float)qx[q] * kW;
            wgy[q] = (float)qy[q] * kW;
        }
        float f1 = 0.0f, f2 = 0.0f;
        fb += in ? 0 : 1;
        if (in) {
            const float2* ew = Ew + rx + __umul24(wx, ry);
#pragma unroll
            for (int jj = 0; jj < n; ++jj)
#pragma unroll
                for (int ii = 0; ii < n; ++ii) {
                    const float w = wgx[ii] * wgy[jj];
                    const float2 e = ew[ii + wx * jj];
                    f1 = fmaf(e.x, w, f1);
                    f2 = fmaf(e.y, w, f2);
                }
        } else {
#pragma unroll
            for (int jj = 0; jj < n; ++jj) {
                const int64_t row = (int64_t)NX * ((fy + jj) & (NY - 1));
#pragma unroll
                for (int ii = 0; ii < n; ++ii) {
                    const float w = wgx[ii] * wgy[jj];
                    const float2 e = E32[row + ((fx + ii) & (NX - 1))];
                    f1 = fmaf(e.x, w, f1);
                    f2 = fmaf(e.y, w, f2);
                }
            }
        }
        // Boris as a change: half kick e, rotation of m = v + e, half kick
        const float e1 = f1 * s.eh, e2 = f2 * s.eh;
        const float m1 = vx + e1, m2 = vy + e2, m3 = vz;
        const float u1 = m1 + (m2 * t3 - m3 * t2), u2 = m2 + (m3 * t1 - m1 * t3), u3 = m3 + (m1 * t2 - m2 * t1);
        const float dv1 = (e1 + e1) + (u2 * t3 - u3 * t2) * s.bs;
        const float dv2 = (e2 + e2) + (u3 * t1 - u1 * t3) * s.bs;
        const float dv3 = (u1 * t2 - u2 * t1) * s.bs;
        // x += ((v0 + v)/2)*dt with the mean velocity v0 + dv/2
        xq += move_q23(vx, lx, 0.5f * dv1, g.stepx);
        yq += move_q23(vy, ly, 0.5f * dv2, g.stepy);
        hilo_add23(vx, lx, dv1);
        hilo_add23(vy, ly, dv2);
        hilo_add23(vz, lz, dv3);
        s.xq[p] = xq;
        s.yq[p] = yq;
        s.vxh[p] = vx;
        s.vyh[p] = vy;
        s.vzh[p] = vz;
        s.vxl[p] = lx;
        s.vyl[p] = ly;
        s.vzl[p] = lz;
        // deposit at the new position
        fx = shape_q23<SH>(xq, g.sx, qx);
        fy = shape_q23<SH>(yq, g.sy, qy);
        rx = (fx - ox) & (NX - 1);
        ry = (fy - oy) & (NY - 1);
        in = tile && rx + n <= wx && ry + n <= wy;
        fb += in ? 0 : 1;
        if (in) {
            u64* rw = Rw + rx + __umul24(wx, ry);
#pragma unroll
            for (int jj = 0; jj < n; ++jj)
#pragma unroll
                for (int ii = 0; ii < n; ++ii) {
                    const u64 v = quanta_product(qx[ii], qy[jj]);
                    if (v) atomicAdd(rw + ii + wx * jj, v);
                }
        } else {
#pragma unroll
            for (int jj = 0; jj < n; ++jj) {
                const int64_t row = (int64_t)NX * ((fy + jj) & (NY - 1));
#pragma unroll
                for (int ii = 0; ii < n; ++ii) {
                    const u64 v = quanta_product(qx[ii], qy[jj]);
                    if (v) add_q128(cnt, G, row + ((fx + ii) & (NX - 1)), v);
                }
            }
        }
    }
    count_fallback(g.fallback, fb);
    __syncthreads();
    if (tile) {
        for (int e = t; e < wx * wy; e += kTile23Threads) {
            const u64 v = Rw[e];
            if (!v) continue;
            const int ry0 = e / wx, rx0 = e - ry0 * wx;
            add_q128(cnt, G, ((ox + rx0) & (NX - 1)) + (int64_t)NX * ((oy + ry0) & (NY - 1)), v);
        }
    }
}

// rho = sum over species (in order) of count * qscale; the counters cleared
// for the next step
struct Rho23m {
    int ns;
    u64* cnt[kMaxSpecies23];
    double scale[kMaxSpecies23];
};

__global__ void __launch_bounds__(256) k23m_rho(int64_t G, Rho23m r, double* __restrict__ rho) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < G; k += (int64_t)gridDim.x * blockDim.x) {
        double v = 0.0;
        for (int s = 0; s < r.ns; ++s) {
            u64* c = r.cnt[s];
            const u64 lo = c[k], hi = c[G + k];
            c[k] = 0;
            c[G + k] = 0;
            const double q = ((double)hi * 18446744073709551616.0 + (double)lo) * r.scale[s];
            v = s == 0 ? q : v + q;
        }
        rho[k] = v;
    }
}

// ---------------------------------------------------------------------------
// conversions: xyv (5 x P, column-major, caller order) <-> the mixed state
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t to_q23(double z, double L) {
    const double a = rint(z / L * 4294967296.0);
    return (a > -9.0e18 && a < 9.0e18) ? (uint32_t)(int64_t)a : 0u;  // mod 2^32: the periodic image
}
__device__ __forceinline__ double from_q23(uint32_t zq, double L) {
    return zq == 0u ? L : (double)zq * L * 2.3283064365386963e-10;  // 2^-32
}
__device__ __forceinline__ void split23(double v, float& hi, float& lo) {
    hi = (float)v;
    lo = (float)(v - (double)hi);
}

__global__ void __launch_bounds__(256) k23m_unpack(Grid23 g, const double* __restrict__ xyv, Species23 s) {
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < s.P; p += (int64_t)gridDim.x * blockDim.x) {
        const double* q = xyv + 5 * p;
        s.xq[p] = to_q23(q[0], g.Lx);
        s.yq[p] = to_q23(q[1], g.Ly);
        split23(q[2], s.vxh[p], s.vxl[p]);
        split23(q[3], s.vyh[p], s.vyl[p]);
        split23(q[4], s.vzh[p], s.vzl[p]);
    }
}

// caller index p from its held slot
__global__ void __launch_bounds__(256) k23m_pack(Grid23 g, Species23 s, double* __restrict__ xyv) {
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < s.P; p += (int64_t)gridDim.x * blockDim.x) {
        double* q = xyv + 5 * p;
        const int64_t h = s.slot ? (int64_t)s.slot[p] : p;
        q[0] = from_q23(s.xq[h], g.Lx);
        q[1] = from_q23(s.yq[h], g.Ly);
        q[2] = (double)s.vxh[h] + (double)s.vxl[h];
        q[3] = (double)s.vyh[h] + (double)s.vyl[h];
        q[4] = (double)s.vzh[h] + (double)s.vzl[h];
    }
}

// the FP64 arrays of s from its mixed state, in the held order
__global__ void __launch_bounds__(256) k23m_to64(Grid23 g, Species23 s) {
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < s.P; p += (int64_t)gridDim.x * blockDim.x) {
        s.x[p] = from_q23(s.xq[p], g.Lx);
        s.y[p] = from_q23(s.yq[p], g.Ly);
        s.vx[p] = (double)s.vxh[p] + (double)s.vxl[p];
        s.vy[p] = (double)s.vyh[p] + (double)s.vyl[p];
        s.vz[p] = (double)s.vzh[p] + (double)s.vzl[p];
    }
}

__device__ __forceinline__ void copy_m23(const Species23& s, int64_t from, const Species23& o, int64_t to) {
    o.xq[to] = s.xq[from];
    o.yq[to] = s.yq[from];
    o.vxh[to] = s.vxh[from];
    o.vyh[to] = s.vyh[from];
    o.vzh[to] = s.vzh[from];
    o.vxl[to] = s.vxl[from];
    o.vyl[to] = s.vyl[from];
    o.vzl[to] = s.vzl[from];
}

// ---------------------------------------------------------------------------
// regrouping by tile (as the FP64 state's, on the mixed arrays)
// ---------------------------------------------------------------------------
__device__ __forceinline__ int32_t tile23m(const Grid23& g, const Tiles23& T, uint32_t xq, uint32_t yq) {
    return (int32_t)((xq - 1u) >> g.sx) / T.TX + T.ntx * ((int32_t)((yq - 1u) >> g.sy) / T.TY);
}

__global__ void __launch_bounds__(256) k23m_regroup_count(Grid23 g, Tiles23 T, Species23 s, int64_t per_block) {
    extern __shared__ int32_t hist[];
    for (int i = threadIdx.x; i < T.ntiles; i += blockDim.x) hist[i] = 0;
    __syncthreads();
    const int64_t b0 = (int64_t)blockIdx.x * per_block;
    const int64_t b1 = b0 + per_block < s.P ? b0 + per_block : s.P;
    for (int64_t p = b0 + threadIdx.x; p < b1; p += blockDim.x) atomicAdd(&hist[tile23m(g, T, s.xq[p], s.yq[p])], 1);
    __syncthreads();
    for (int i = threadIdx.x; i < T.ntiles; i += blockDim.x)
        if (hist[i]) atomicAdd(&T.count[i], hist[i]);
}

__global__ void __launch_bounds__(256)
k23m_regroup_scatter(Grid23 g, Tiles23 T, Species23 s, Species23 o, const uint32_t* __restrict__ pid,
                     uint32_t* __restrict__ pido, int64_t per_block) {
    extern __shared__ int32_t hist[];
    for (int i = threadIdx.x; i < T.ntiles; i += blockDim.x) hist[i] = 0;
    __syncthreads();
    const int64_t b0 = (int64_t)blockIdx.x * per_block;
    const int64_t b1 = b0 + per_block < s.P ? b0 + per_block : s.P;
    for (int64_t p = b0 + threadIdx.x; p < b1; p += blockDim.x) atomicAdd(&hist[tile23m(g, T, s.xq[p], s.yq[p])], 1);
    __syncthreads();
    for (int i = threadIdx.x; i < T.ntiles; i += blockDim.x)
        if (hist[i]) hist[i] = atomicAdd(&T.cursor[i], hist[i]);
    __syncthreads();
    for (int64_t p = b0 + threadIdx.x; p < b1; p += blockDim.x) {
        const int32_t d = atomicAdd(&hist[tile23m(g, T, s.xq[p], s.yq[p])], 1);
        const uint32_t id = pid[p];
        copy_m23(s, p, o, d);
        pido[d] = id;
        s.slot[id] = (uint32_t)d;
    }
}

// the held particles of s in the caller's order into o
__global__ void __launch_bounds__(256) k23m_to_caller(Species23 s, Species23 o) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < s.P; i += (int64_t)gridDim.x * blockDim.x)
        copy_m23(s, (int64_t)s.slot[i], o, i);
}

// o[k] = s[perm[k]] (sort!'s permutation applied to the mixed state)
__global__ void __launch_bounds__(256) k23m_permute(Species23 s, Species23 o, const uint32_t* __restrict__ perm) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < s.P; k += (int64_t)gridDim.x * blockDim.x)
        copy_m23(s, (int64_t)perm[k], o, k);
}

// ---------------------------------------------------------------------------
// diagnostics: the kinetic sum sum(|v|^2), the momentum sums (vx, vy, vz)
// and the characteristic ones (|vx|, |vy|, |vz|) of a species in one pass
// over the caller's order (each particle read once, through slot): a block
// takes kTreeChunk particles, its lanes sum them strided and the fixed tree
// folds the lanes, so the partials do not depend on the held order.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k23m_diag_species(Species23 s, double* __restrict__ kpart, double* __restrict__ mpart) {
    __shared__ double red[3 * kTreeLanes + 8];
    const int64_t b0 = (int64_t)blockIdx.x * kTreeChunk;
    const int64_t n = s.P - b0 < kTreeChunk ? s.P - b0 : kTreeChunk;
    double ke = 0.0, a = 0.0, b = 0.0, c = 0.0, e = 0.0, f = 0.0, h = 0.0;
    if (threadIdx.x < kTreeLanes)
        for (int64_t i = threadIdx.x; i < n; i += kTreeLanes) {
            const int64_t p = s.slot[b0 + i];
            const double vx = (double)s.vxh[p] + (double)s.vxl[p];
            const double vy = (double)s.vyh[p] + (double)s.vyl[p];
            const double vz = (double)s.vzh[p] + (double)s.vzl[p];
            ke = ke + ((vx * vx + vy * vy) + vz * vz);
            a = a + vx;
            b = b + vy;
            c = c + vz;
            e = e + fabs(vx);
            f = f + fabs(vy);
            h = h + fabs(vz);
        }
    tree256x3<false>(red, a, b, c);
    tree256x3<false>(red, e, f, h);
    ke = tree256<false>(red, ke);
    if (threadIdx.x == 0) {
        kpart[blockIdx.x] = ke;
        double* o = mpart + 6 * blockIdx.x;
        o[0] = a;
        o[1] = b;
        o[2] = c;
        o[3] = e;
        o[4] = f;
        o[5] = h;
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static unsigned blocks23m(int64_t n, int threads) {
    const int64_t b = (n + threads - 1) / threads;
    return (unsigned)(b < 16384 ? (b > 0 ? b : 1) : 16384);
}

void launch23m_push(const Grid23& g, const Species23& s, const float2* E32, hipStream_t st) {
    const unsigned nb = (unsigned)((s.P + kTile23Chunk - 1) / kTile23Chunk);
#define PIC23M_PUSH(SH)                                             \
    case SH:                                                        \
        k23m_push<SH><<<nb, kTile23Threads, 0, st>>>(g, s, E32);    \
        break;
    switch (s.shape) {
        PIC23M_PUSH(0) PIC23M_PUSH(1) PIC23M_PUSH(2) PIC23M_PUSH(3) PIC23M_PUSH(4) PIC23M_PUSH(5) PIC23M_PUSH(6)
        PIC23M_PUSH(7)
    }
#undef PIC23M_PUSH
}

void launch23m_rho(const Grid23& g, const Species23* sp, int ns, double* rho, hipStream_t st) {
    const int64_t G = (int64_t)g.NX * g.NY;
    Rho23m r{};
    r.ns = ns;
    for (int s = 0; s < ns; ++s) {
        r.cnt[s] = sp[s].cnt;
        r.scale[s] = sp[s].qscale;
    }
    k23m_rho<<<blocks23m(G, 256), 256, 0, st>>>(G, r, rho);
}

void launch23m_unpack(const Grid23& g, const double* xyv, const Species23& s, hipStream_t st) {
    k23m_unpack<<<blocks23m(s.P, 256), 256, 0, st>>>(g, xyv, s);
}

void launch23m_pack(const Grid23& g, const Species23& s, double* xyv, hipStream_t st) {
    k23m_pack<<<blocks23m(s.P, 256), 256, 0, st>>>(g, s, xyv);
}

void launch23m_to64(const Grid23& g, const Species23& s, hipStream_t st) {
    k23m_to64<<<blocks23m(s.P, 256), 256, 0, st>>>(g, s);
}

void launch23m_regroup(const Grid23& g, const Tiles23& T, const Species23& s, const Species23& o,
                       const uint32_t* pid, uint32_t* pido, hipStream_t st) {
    if (s.P < 1) return;
    const size_t lds = sizeof(int32_t) * (size_t)T.ntiles;
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(k23m_regroup_count),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(k23m_regroup_scatter),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    const int64_t per = 16384;
    const unsigned nb = (unsigned)((s.P + per - 1) / per);
    k23m_regroup_count<<<nb, 256, lds, st>>>(g, T, s, per);
    launch23_regroup_scan(T, st);
    k23m_regroup_scatter<<<nb, 256, lds, st>>>(g, T, s, o, pid, pido, per);
}

void launch23m_to_caller(const Species23& s, const Species23& o, hipStream_t st) {
    if (s.P > 0) k23m_to_caller<<<blocks23m(s.P, 256), 256, 0, st>>>(s, o);
}

int64_t launch23m_diag_species(const Species23& s, double* kpart, double* mpart, hipStream_t st) {
    const int64_t nb = (s.P + kTreeChunk - 1) / kTreeChunk;
    k23m_diag_species<<<(unsigned)nb, 256, 0, st>>>(s, kpart, mpart);
    return nb;
}

void launch23m_permute(const Species23& s, const Species23& o, const uint32_t* perm, hipStream_t st) {
    if (s.P > 0) k23m_permute<<<blocks23m(s.P, 256), 256, 0, st>>>(s, o, perm);
}

}  // namespace pic
