// kernels_2d.hip -- electrostatic 2D3V step (src/Electrostatic2D3V.jl):
// particle loop (gather, Boris about B = B0 x, move, area deposit), exact
// per-cell charge sums, the 2D spectral Poisson solve and the diagnostics.
//
// FP64 kernels evaluate the script's expressions in its order (contraction
// off; the gather's @muladd is an explicit fma) and sum charge exactly on a
// 2^-q fixed-point grid (three 42-bit limbs per cell, device-scope integer
// atomics), so every result is independent of the schedule.
//
// Mixed kernels keep positions as 32-bit fixed-point fractions of the box
// (periodic wrap for free), velocities as FP32 hi + lo pairs (the step's
// change, computed in FP32, is added to the pair exactly, so no FP32 rounding
// of the velocity itself accumulates over a run), and use exact integer area
// weights of 2^-20 of a cell per axis: a particle deposits exactly 2^40 quanta,
// summed in u64 per cell.  The solve and the diagnostics stay FP64.
//
// The transform is the radix-2 decimation-in-time FFT of the CPU restatement
// applied to lines of the grid: forward along x then y, inverse along y then x.
#include "fft_lines.h"
#include "pic_device.h"
#include "pic2d_internal.h"

#pragma clang fp contract(off)

namespace pic {

// ---------------------------------------------------------------------------
// scalar pieces
// ---------------------------------------------------------------------------

// unimod(x, n) = 0 < x <= n ? x : x > n ? x - n : x + n   (:83)
__device__ __forceinline__ int64_t unimodi(int64_t x, int64_t n) {
    return (0 < x && x <= n) ? x : (x > n ? x - n : x + n);
}
__device__ __forceinline__ double unimod1(double x) {
    return (0.0 < x && x <= 1.0) ? x : (x > 1.0 ? x - 1.0 : x + 1.0);
}

// g(z, NZ) (:84-92): cells (0-based) and weights (1 - r, r)
struct Area64 {
    int32_t c0, c1;
    double w0, w1;
};
__device__ __forceinline__ Area64 area64(double z, int32_t NZ, int32_t pairing) {
    const double zN = z * (double)NZ;
    const int64_t i = unimodi((int64_t)ceil(zN), NZ);
    const double r = (double)i - zN;
    const int64_t i2 = pairing ? unimodi(i - 1, NZ) : unimodi(i + 1, NZ);
    // z in (0, 1] gives i, i2 in 1..NZ; the mask only keeps a non-finite or
    // out-of-range z (which the host rejects) from addressing outside the grid
    return {(int32_t)((i - 1) & (NZ - 1)), (int32_t)((i2 - 1) & (NZ - 1)), 1.0 - r, r};
}

// The same pairing from a fixed-point coordinate zq (z = zq/2^32, 0 meaning
// z == 1): i = ceil(zN) in 1..NZ, r = i - zN truncated to 2^-20 of a cell;
// weights in quanta q0 = 2^20 - r, q1 = r.
struct AreaQ {
    int32_t c0, c1;
    uint32_t q0, q1;
};
__device__ __forceinline__ AreaQ areaq(uint32_t zq, int32_t NZ, int sh, int32_t pairing) {
    uint32_t i = (uint32_t)(((uint64_t)zq + ((1u << sh) - 1)) >> sh);
    if (i == 0) i = (uint32_t)NZ;
    const uint32_t r = (uint32_t)(((uint64_t)i << sh) - zq);
    const uint32_t rq = r >> (sh - kQuanta2D);
    const int32_t c0 = (int32_t)i - 1;
    const int32_t c1 = pairing ? (i == 1 ? NZ - 1 : (int32_t)i - 2) : (i == (uint32_t)NZ ? 0 : (int32_t)i);
    return {c0, c1, (1u << kQuanta2D) - rq, rq};
}

// boris (:32-41) with tvec = (t1, 0, 0), every product of both cross products
// written out as StaticArrays evaluates them.
__device__ __forceinline__ void boris64(double& vx, double& vy, double& vz, double Ex, double Ey, double dt,
                                        double t1, double tscale) {
    const double dt_2 = dt / 2.0;
    const double e1 = Ex * dt_2, e2 = Ey * dt_2, e3 = 0.0;
    const double m1 = vx + e1, m2 = vy + e2, m3 = vz + e3;
    const double c1 = m2 * 0.0 - m3 * 0.0, c2 = m3 * t1 - m1 * 0.0, c3 = m1 * 0.0 - m2 * t1;
    const double u1 = m1 + c1, u2 = m2 + c2, u3 = m3 + c3;
    const double d1 = u2 * 0.0 - u3 * 0.0, d2 = u3 * t1 - u1 * 0.0, d3 = u1 * 0.0 - u2 * t1;
    const double p1 = m1 + d1 * tscale, p2 = m2 + d2 * tscale, p3 = m3 + d3 * tscale;
    vx = p1 + e1;
    vy = p2 + e2;
    vz = p3 + e3;
}

__device__ __forceinline__ void boris_consts(double dt, double B0, double& t1, double& tscale) {
    t1 = B0 * dt / 2.0;
    tscale = 2.0 / (1.0 + (t1 * t1 + 0.0 * 0.0 + 0.0 * 0.0));
}

// eval (:94-103): j outer, i inner, F = fma(real(F[i,j]), wx*wy, F)
__device__ __forceinline__ void gather64(const double* __restrict__ Ex, const double* __restrict__ Ey, int32_t NX,
                                         const Area64& ax, const Area64& ay, double& f1, double& f2) {
    f1 = 0.0;
    f2 = 0.0;
    const int32_t cj[2] = {ay.c0, ay.c1};
    const double wj[2] = {ay.w0, ay.w1};
    const int32_t ci[2] = {ax.c0, ax.c1};
    const double wi[2] = {ax.w0, ax.w1};
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            const double wxy = wi[a] * wj[b];
            const int64_t k = (int64_t)ci[a] + (int64_t)NX * cj[b];
            f1 = fma(Ex[k], wxy, f1);
            f2 = fma(Ey[k], wxy, f2);
        }
}

constexpr u64 kLimb2DMask = (1ull << kLimb2D) - 1;

__device__ __forceinline__ void add_limbs42(u64* __restrict__ racc, int64_t G, int64_t k, u128 v) {
    const u64 l0 = (u64)v & kLimb2DMask;
    const u64 l1 = (u64)(v >> kLimb2D) & kLimb2DMask;
    const u64 l2 = (u64)(v >> (2 * kLimb2D));
    if (l0) atomicAdd(racc + k, l0);
    if (l1) atomicAdd(racc + G + k, l1);
    if (l2) atomicAdd(racc + 2 * G + k, l2);
}

// depositcharge! (:105-109): j outer, i inner, contribution (wx*wy)*w
__device__ __forceinline__ void deposit64(u64* __restrict__ racc, const Desc2D& D, int64_t G, double x, double y) {
    const Area64 ax = area64(x, D.NX, D.pairing), ay = area64(y, D.NY, D.pairing);
    const int32_t cj[2] = {ay.c0, ay.c1};
    const double wj[2] = {ay.w0, ay.w1};
    const int32_t ci[2] = {ax.c0, ax.c1};
    const double wi[2] = {ax.w0, ax.w1};
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            const double c = wi[a] * wj[b] * D.w;
            add_limbs42(racc, G, (int64_t)ci[a] + (int64_t)D.NX * cj[b], fxp_from(c, D.q));
        }
}

// ---------------------------------------------------------------------------
// particle kernels
// MODE 0: gather, Boris, move, deposit (the particle loop, :128-135)
// MODE 1: gather, Boris, move
// MODE 2: deposit at the current positions
// ---------------------------------------------------------------------------
template <int MODE>
__global__ void __launch_bounds__(kPush2DThreads)
k2d_push64(Desc2D D, double* __restrict__ X, double* __restrict__ Y, double* __restrict__ VX,
           double* __restrict__ VY, double* __restrict__ VZ, const double* __restrict__ Ex,
           const double* __restrict__ Ey, u64* __restrict__ racc, double dt, double B0) {
    const int64_t G = (int64_t)D.NX * D.NY;
    double t1, tscale;
    boris_consts(dt, B0, t1, tscale);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < D.P; p += stride) {
        double x = X[p], y = Y[p];
        if (MODE != 2) {
            double vx = VX[p], vy = VY[p], vz = VZ[p];
            double f1, f2;
            gather64(Ex, Ey, D.NX, area64(x, D.NX, D.pairing), area64(y, D.NY, D.pairing), f1, f2);
            boris64(vx, vy, vz, f1, f2, dt, t1, tscale);
            x = unimod1(x + vx * dt);
            y = unimod1(y + vy * dt);
            X[p] = x;
            Y[p] = y;
            VX[p] = vx;
            VY[p] = vy;
            VZ[p] = vz;
        }
        if (MODE != 1) deposit64(racc, D, G, x, y);
    }
}

__global__ void __launch_bounds__(kPush2DThreads)
k2d_gather64(Desc2D D, const double* __restrict__ Ex, const double* __restrict__ Ey, const double* __restrict__ X,
             const double* __restrict__ Y, double* __restrict__ ex, double* __restrict__ ey, int64_t P) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < P; p += stride) {
        double f1, f2;
        gather64(Ex, Ey, D.NX, area64(X[p], D.NX, D.pairing), area64(Y[p], D.NY, D.pairing), f1, f2);
        ex[p] = f1;
        ey[p] = f2;
    }
}

__global__ void __launch_bounds__(kPush2DThreads)
k2d_boris64(double* __restrict__ VX, double* __restrict__ VY, double* __restrict__ VZ, const double* __restrict__ ex,
            const double* __restrict__ ey, int64_t P, double dt, double B0) {
    double t1, tscale;
    boris_consts(dt, B0, t1, tscale);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < P; p += stride) {
        double vx = VX[p], vy = VY[p], vz = VZ[p];
        boris64(vx, vy, vz, ex[p], ey[p], dt, t1, tscale);
        VX[p] = vx;
        VY[p] = vy;
        VZ[p] = vz;
    }
}

// ---------------------------------------------------------------------------
// v += d on a hi + lo pair: two-sum of hi and d, the error folded into lo,
// renormalised (the error terms are exact for |d| well below |hi|, as here).
__device__ __forceinline__ void add_hilo(float& hi, float& lo, float d) {
    const float s = hi + d;
    const float bb = s - hi;
    const float err = (hi - (s - bb)) + (d - bb);
    const float t = lo + err;
    hi = s + t;
    lo = t - (hi - s);
}

// rint((hi + lo) * step): the integer part of hi*step is taken first, so the
// rounding sees both words (hi*step and lo*step are exact for a power-of-two
// step, dt = 2^-k, and within half an FP32 ulp otherwise).
__device__ __forceinline__ int32_t step_of(float hi, float lo, float step) {
    const float a = hi * step, ai = rintf(a);
    return (int32_t)ai + (int32_t)rintf((a - ai) + lo * step);
}

// Mixed particle loop: FP32 gather with exact quanta weights, the Boris update
// about x as FP32 velocity changes added to the hi + lo pairs, fixed-point move
// (x += rint(vx*dt*2^32) mod 2^32), exact quanta deposit.
// Particles are held by tile.  A block handles part of one tile's
// particles: it stages the field of the tile and kHalo2D cells around it in
// LDS, gathers and deposits there, and adds the staged charge to the grid
// accumulator at the end.  A particle whose cells fall outside the staged
// window (it drifted further than the halo since the last sort) reads and adds
// through global memory instead, so every result is the same as without
// tiles: the charge sums are exact integers, each particle's update depends
// only on its own state and the field.
// ---------------------------------------------------------------------------
struct TileWin {
    int32_t ox, oy;  // global cell of window slot (0, 0) (may be negative: periodic)
    int32_t wx, wy;  // window size in cells
};

__device__ __forceinline__ int32_t win_rel(int32_t c, int32_t o, int32_t n) {
    int32_t r = c - o;
    r += (r < 0) ? n : 0;
    r -= (r >= n) ? n : 0;
    return r;
}

__global__ void __launch_bounds__(kTileThreads)
k2d_push32t(Desc2D D, Tiles2D T, uint32_t* __restrict__ Xq, uint32_t* __restrict__ Yq, float* __restrict__ VX,
            float* __restrict__ VY, float* __restrict__ VZ, float* __restrict__ VXl, float* __restrict__ VYl,
            float* __restrict__ VZl, const float2* __restrict__ E32, u64* __restrict__ racc, float dt_2, float t1,
            float tscale, float xstep) {
    __shared__ float2 ew[(kTile2D + 2 * kHalo2D) * (kTile2D + 2 * kHalo2D)];
    __shared__ u64 aw[(kTile2D + 2 * kHalo2D) * (kTile2D + 2 * kHalo2D)];
    constexpr float kQ = 1.0f / (float)(1u << kQuanta2D);
    const int tile = blockIdx.x / T.splits, part = blockIdx.x % T.splits;
    const int tx = tile % T.ntx, ty = tile / T.ntx;
    const int wx = T.TX + 2 * kHalo2D, wy = T.TY + 2 * kHalo2D, nw = wx * wy;
    const int ox = tx * T.TX - kHalo2D, oy = ty * T.TY - kHalo2D;
    for (int i = threadIdx.x; i < nw; i += blockDim.x) {
        const int gx = (ox + i % wx + D.NX) % D.NX, gy = (oy + i / wx + D.NY) % D.NY;
        ew[i] = E32[gx + D.NX * gy];
        aw[i] = 0;
    }
    __syncthreads();
    const int64_t a0 = T.start[tile], a1 = T.start[tile + 1];
    const int64_t len = a1 - a0;
    const int64_t p0 = a0 + len * part / T.splits, p1 = a0 + len * (part + 1) / T.splits;
    for (int64_t p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
        uint32_t xq = Xq[p], yq = Yq[p];
        float vx = VX[p], vy = VY[p], vz = VZ[p];
        float lx = VXl[p], ly = VYl[p], lz = VZl[p];
        {
            const AreaQ ax = areaq(xq, D.NX, D.sx, D.pairing), ay = areaq(yq, D.NY, D.sy, D.pairing);
            const float wx0 = (float)ax.q0 * kQ, wx1 = (float)ax.q1 * kQ;
            const float wy0 = (float)ay.q0 * kQ, wy1 = (float)ay.q1 * kQ;
            const int32_t rx0 = win_rel(ax.c0, ox, D.NX), rx1 = win_rel(ax.c1, ox, D.NX);
            const int32_t ry0 = win_rel(ay.c0, oy, D.NY), ry1 = win_rel(ay.c1, oy, D.NY);
            float2 e00, e10, e01, e11;
            if (rx0 < wx && rx1 < wx && ry0 < wy && ry1 < wy) {
                e00 = ew[rx0 + wx * ry0];
                e10 = ew[rx1 + wx * ry0];
                e01 = ew[rx0 + wx * ry1];
                e11 = ew[rx1 + wx * ry1];
            } else {
                e00 = E32[ax.c0 + D.NX * ay.c0];
                e10 = E32[ax.c1 + D.NX * ay.c0];
                e01 = E32[ax.c0 + D.NX * ay.c1];
                e11 = E32[ax.c1 + D.NX * ay.c1];
            }
            float f1 = 0.0f, f2 = 0.0f;
            float wxy = wx0 * wy0;
            f1 = fmaf(e00.x, wxy, f1);
            f2 = fmaf(e00.y, wxy, f2);
            wxy = wx1 * wy0;
            f1 = fmaf(e10.x, wxy, f1);
            f2 = fmaf(e10.y, wxy, f2);
            wxy = wx0 * wy1;
            f1 = fmaf(e01.x, wxy, f1);
            f2 = fmaf(e01.y, wxy, f2);
            wxy = wx1 * wy1;
            f1 = fmaf(e11.x, wxy, f1);
            f2 = fmaf(e11.y, wxy, f2);
            // the Boris update as changes: half kick, rotation about x, half kick
            const float e1 = f1 * dt_2, e2 = f2 * dt_2;
            const float m2 = vy + e2, m3 = vz;
            const float u2 = m2 + m3 * t1, u3 = m3 - m2 * t1;
            add_hilo(vx, lx, e1 + e1);
            add_hilo(vy, ly, (e2 + e2) + (u3 * t1) * tscale);
            add_hilo(vz, lz, -((u2 * t1) * tscale));
        }
        xq += (uint32_t)step_of(vx, lx, xstep);
        yq += (uint32_t)step_of(vy, ly, xstep);
        Xq[p] = xq;
        Yq[p] = yq;
        VX[p] = vx;
        VY[p] = vy;
        VZ[p] = vz;
        VXl[p] = lx;
        VYl[p] = ly;
        VZl[p] = lz;
        const AreaQ bx = areaq(xq, D.NX, D.sx, D.pairing), by = areaq(yq, D.NY, D.sy, D.pairing);
        const u64 q00 = (u64)bx.q0 * by.q0, q10 = (u64)bx.q1 * by.q0, q01 = (u64)bx.q0 * by.q1,
                  q11 = (u64)bx.q1 * by.q1;
        const int32_t rx0 = win_rel(bx.c0, ox, D.NX), rx1 = win_rel(bx.c1, ox, D.NX);
        const int32_t ry0 = win_rel(by.c0, oy, D.NY), ry1 = win_rel(by.c1, oy, D.NY);
        if (rx0 < wx && rx1 < wx && ry0 < wy && ry1 < wy) {
            if (q00) atomicAdd(&aw[rx0 + wx * ry0], q00);
            if (q10) atomicAdd(&aw[rx1 + wx * ry0], q10);
            if (q01) atomicAdd(&aw[rx0 + wx * ry1], q01);
            if (q11) atomicAdd(&aw[rx1 + wx * ry1], q11);
        } else {
            const int64_t r0 = (int64_t)D.NX * by.c0, r1 = (int64_t)D.NX * by.c1;
            if (q00) atomicAdd(racc + r0 + bx.c0, q00);
            if (q10) atomicAdd(racc + r0 + bx.c1, q10);
            if (q01) atomicAdd(racc + r1 + bx.c0, q01);
            if (q11) atomicAdd(racc + r1 + bx.c1, q11);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nw; i += blockDim.x) {
        const u64 v = aw[i];
        if (v) {
            const int gx = (ox + i % wx + D.NX) % D.NX, gy = (oy + i / wx + D.NY) % D.NY;
            atomicAdd(racc + gx + (int64_t)D.NX * gy, v);
        }
    }
}

// Tile of a held particle: the tile holding its cell.
__device__ __forceinline__ int32_t tile_of(const Desc2D& D, const Tiles2D& T, uint32_t xq, uint32_t yq) {
    const int32_t cx = (int32_t)(xq >> D.sx), cy = (int32_t)(yq >> D.sy);
    return (cx / T.TX) + T.ntx * (cy / T.TY);
}

// Sort, pass 1: particles per tile (block histogram in LDS, then one global
// add per tile).
__global__ void __launch_bounds__(256)
k2d_tile_count(Desc2D D, Tiles2D T, const uint32_t* __restrict__ Xq, const uint32_t* __restrict__ Yq) {
    extern __shared__ int32_t hist[];
    for (int i = threadIdx.x; i < T.ntiles; i += blockDim.x) hist[i] = 0;
    __syncthreads();
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < D.P; p += stride)
        atomicAdd(&hist[tile_of(D, T, Xq[p], Yq[p])], 1);
    __syncthreads();
    for (int i = threadIdx.x; i < T.ntiles; i += blockDim.x)
        if (hist[i]) atomicAdd(&T.count[i], hist[i]);
}

// Sort, pass 2 (one block): exclusive scan of the counts into start and the
// cursors; counts cleared for the next sort.
__global__ void __launch_bounds__(1024) k2d_tile_scan(Tiles2D T) {
    __shared__ int32_t part[1024];
    const int per = (T.ntiles + blockDim.x - 1) / blockDim.x;
    const int i0 = threadIdx.x * per;
    int32_t s = 0;
    for (int i = i0; i < i0 + per && i < T.ntiles; ++i) s += T.count[i];
    part[threadIdx.x] = s;
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

// Sort, pass 3: move every particle to a free slot of its tile's range in the
// other buffers.  A block reserves one contiguous run per tile it holds, then
// hands out slots from LDS; the order inside a tile is arbitrary (results do
// not depend on it).
__global__ void __launch_bounds__(256)
k2d_tile_scatter(Desc2D D, Tiles2D T, const uint32_t* __restrict__ Xq, const uint32_t* __restrict__ Yq,
                 const float* __restrict__ VX, const float* __restrict__ VY, const float* __restrict__ VZ,
                 const float* __restrict__ VXl, const float* __restrict__ VYl, const float* __restrict__ VZl,
                 const uint32_t* __restrict__ pid, uint32_t* __restrict__ Xo, uint32_t* __restrict__ Yo,
                 float* __restrict__ VXo, float* __restrict__ VYo, float* __restrict__ VZo,
                 float* __restrict__ VXlo, float* __restrict__ VYlo, float* __restrict__ VZlo,
                 uint32_t* __restrict__ pido, int64_t per_block) {
    extern __shared__ int32_t hist[];
    for (int i = threadIdx.x; i < T.ntiles; i += blockDim.x) hist[i] = 0;
    __syncthreads();
    const int64_t b0 = (int64_t)blockIdx.x * per_block;
    const int64_t b1 = b0 + per_block < D.P ? b0 + per_block : D.P;
    for (int64_t p = b0 + threadIdx.x; p < b1; p += blockDim.x) atomicAdd(&hist[tile_of(D, T, Xq[p], Yq[p])], 1);
    __syncthreads();
    for (int i = threadIdx.x; i < T.ntiles; i += blockDim.x)
        if (hist[i]) hist[i] = atomicAdd(&T.cursor[i], hist[i]);
    __syncthreads();
    for (int64_t p = b0 + threadIdx.x; p < b1; p += blockDim.x) {
        const uint32_t xq = Xq[p], yq = Yq[p];
        const int32_t d = atomicAdd(&hist[tile_of(D, T, xq, yq)], 1);
        Xo[d] = xq;
        Yo[d] = yq;
        VXo[d] = VX[p];
        VYo[d] = VY[p];
        VZo[d] = VZ[p];
        VXlo[d] = VXl[p];
        VYlo[d] = VYl[p];
        VZlo[d] = VZl[p];
        pido[d] = pid[p];
    }
}

// ---------------------------------------------------------------------------
// FP64 particles held by tile.  The tile of a particle is that of its area64
// cell c0 on each axis.  k2d_push64t is k2d_push64<0>'s arithmetic with the
// tile's field (and kHalo2D cells around it) read from LDS and the charge's
// three limbs summed in LDS, then added to the grid's limbs; a particle whose
// cells fall outside the window reads and adds through global memory.  Every
// particle's update depends only on its own state and the field, and the limb
// sums are exact, so the results are k2d_push64<0>'s.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int32_t tile_of64(const Desc2D& D, const Tiles2D& T, double x, double y) {
    const int32_t cx = area64(x, D.NX, D.pairing).c0, cy = area64(y, D.NY, D.pairing).c0;
    return (cx / T.TX) + T.ntx * (cy / T.TY);
}

__global__ void __launch_bounds__(256)
k2d_tile_count64(Desc2D D, Tiles2D T, const double* __restrict__ X, const double* __restrict__ Y) {
    extern __shared__ int32_t hist[];
    for (int i = threadIdx.x; i < T.ntiles; i += blockDim.x) hist[i] = 0;
    __syncthreads();
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < D.P; p += stride)
        atomicAdd(&hist[tile_of64(D, T, X[p], Y[p])], 1);
    __syncthreads();
    for (int i = threadIdx.x; i < T.ntiles; i += blockDim.x)
        if (hist[i]) atomicAdd(&T.count[i], hist[i]);
}

// as k2d_tile_scatter; slot[i] gets the new slot of caller index i
__global__ void __launch_bounds__(256)
k2d_tile_scatter64(Desc2D D, Tiles2D T, const double* __restrict__ X, const double* __restrict__ Y,
                   const double* __restrict__ VX, const double* __restrict__ VY, const double* __restrict__ VZ,
                   const uint32_t* __restrict__ pid, double* __restrict__ Xo, double* __restrict__ Yo,
                   double* __restrict__ VXo, double* __restrict__ VYo, double* __restrict__ VZo,
                   uint32_t* __restrict__ pido, uint32_t* __restrict__ slot, int64_t per_block) {
    extern __shared__ int32_t hist[];
    for (int i = threadIdx.x; i < T.ntiles; i += blockDim.x) hist[i] = 0;
    __syncthreads();
    const int64_t b0 = (int64_t)blockIdx.x * per_block;
    const int64_t b1 = b0 + per_block < D.P ? b0 + per_block : D.P;
    for (int64_t p = b0 + threadIdx.x; p < b1; p += blockDim.x) atomicAdd(&hist[tile_of64(D, T, X[p], Y[p])], 1);
    __syncthreads();
    for (int i = threadIdx.x; i < T.ntiles; i += blockDim.x)
        if (hist[i]) hist[i] = atomicAdd(&T.cursor[i], hist[i]);
    __syncthreads();
    for (int64_t p = b0 + threadIdx.x; p < b1; p += blockDim.x) {
        const double x = X[p], y = Y[p];
        const int32_t d = atomicAdd(&hist[tile_of64(D, T, x, y)], 1);
        const uint32_t id = pid[p];
        Xo[d] = x;
        Yo[d] = y;
        VXo[d] = VX[p];
        VYo[d] = VY[p];
        VZo[d] = VZ[p];
        pido[d] = id;
        slot[id] = (uint32_t)d;
    }
}

__global__ void __launch_bounds__(kTileThreads)
k2d_push64t(Desc2D D, Tiles2D T, double* __restrict__ X, double* __restrict__ Y, double* __restrict__ VX,
            double* __restrict__ VY, double* __restrict__ VZ, const double* __restrict__ Ex,
            const double* __restrict__ Ey, u64* __restrict__ racc, double dt, double B0) {
    constexpr int W = kTile2D + 2 * kHalo2D;
    __shared__ double exw[W * W], eyw[W * W];
    __shared__ u64 aw[3 * W * W];
    const int tile = blockIdx.x / T.splits, part = blockIdx.x % T.splits;
    const int tx = tile % T.ntx, ty = tile / T.ntx;
    const int wx = T.TX + 2 * kHalo2D, wy = T.TY + 2 * kHalo2D, nw = wx * wy;
    const int ox = tx * T.TX - kHalo2D, oy = ty * T.TY - kHalo2D;
    const int64_t G = (int64_t)D.NX * D.NY;
    for (int i = threadIdx.x; i < nw; i += blockDim.x) {
        const int gx = (ox + i % wx + D.NX) % D.NX, gy = (oy + i / wx + D.NY) % D.NY;
        exw[i] = Ex[gx + (int64_t)D.NX * gy];
        eyw[i] = Ey[gx + (int64_t)D.NX * gy];
        aw[i] = 0;
        aw[nw + i] = 0;
        aw[2 * nw + i] = 0;
    }
    __syncthreads();
    double t1, tscale;
    boris_consts(dt, B0, t1, tscale);
    const int64_t a0 = T.start[tile], a1 = T.start[tile + 1];
    const int64_t len = a1 - a0;
    const int64_t p0 = a0 + len * part / T.splits, p1 = a0 + len * (part + 1) / T.splits;
    for (int64_t p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
        double x = X[p], y = Y[p], vx = VX[p], vy = VY[p], vz = VZ[p];
        {
            const Area64 ax = area64(x, D.NX, D.pairing), ay = area64(y, D.NY, D.pairing);
            const int32_t ci[2] = {ax.c0, ax.c1}, cj[2] = {ay.c0, ay.c1};
            const double wi[2] = {ax.w0, ax.w1}, wj[2] = {ay.w0, ay.w1};
            const int32_t ri[2] = {win_rel(ax.c0, ox, D.NX), win_rel(ax.c1, ox, D.NX)};
            const int32_t rj[2] = {win_rel(ay.c0, oy, D.NY), win_rel(ay.c1, oy, D.NY)};
            const bool in = ri[0] < wx && ri[1] < wx && rj[0] < wy && rj[1] < wy;
            double f1 = 0.0, f2 = 0.0;  // gather64's order: j outer, i inner
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int a = 0; a < 2; ++a) {
                    const double wxy = wi[a] * wj[b];
                    double e1, e2;
                    if (in) {
                        e1 = exw[ri[a] + wx * rj[b]];
                        e2 = eyw[ri[a] + wx * rj[b]];
                    } else {
                        const int64_t k = (int64_t)ci[a] + (int64_t)D.NX * cj[b];
                        e1 = Ex[k];
                        e2 = Ey[k];
                    }
                    f1 = fma(e1, wxy, f1);
                    f2 = fma(e2, wxy, f2);
                }
            boris64(vx, vy, vz, f1, f2, dt, t1, tscale);
        }
        x = unimod1(x + vx * dt);
        y = unimod1(y + vy * dt);
        X[p] = x;
        Y[p] = y;
        VX[p] = vx;
        VY[p] = vy;
        VZ[p] = vz;
        // deposit64's contributions (wx*wy)*w, j outer, i inner
        const Area64 bx = area64(x, D.NX, D.pairing), by = area64(y, D.NY, D.pairing);
        const int32_t ci[2] = {bx.c0, bx.c1}, cj[2] = {by.c0, by.c1};
        const double wi[2] = {bx.w0, bx.w1}, wj[2] = {by.w0, by.w1};
        const int32_t ri[2] = {win_rel(bx.c0, ox, D.NX), win_rel(bx.c1, ox, D.NX)};
        const int32_t rj[2] = {win_rel(by.c0, oy, D.NY), win_rel(by.c1, oy, D.NY)};
        const bool in = ri[0] < wx && ri[1] < wx && rj[0] < wy && rj[1] < wy;
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                const double c = wi[a] * wj[b] * D.w;
                const u128 v = fxp_from(c, D.q);
                if (in) {
                    const int e = ri[a] + wx * rj[b];
                    const u64 l0 = (u64)v & kLimb2DMask;
                    const u64 l1 = (u64)(v >> kLimb2D) & kLimb2DMask;
                    const u64 l2 = (u64)(v >> (2 * kLimb2D));
                    if (l0) atomicAdd(&aw[e], l0);
                    if (l1) atomicAdd(&aw[nw + e], l1);
                    if (l2) atomicAdd(&aw[2 * nw + e], l2);
                } else {
                    add_limbs42(racc, G, (int64_t)ci[a] + (int64_t)D.NX * cj[b], v);
                }
            }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nw; i += blockDim.x) {
        const int gx = (ox + i % wx + D.NX) % D.NX, gy = (oy + i / wx + D.NY) % D.NY;
        const int64_t k = gx + (int64_t)D.NX * gy;
#pragma unroll
        for (int l = 0; l < 3; ++l) {
            const u64 v = aw[l * nw + i];
            if (v) atomicAdd(racc + l * G + k, v);
        }
    }
}


// Mixed kinetic diagnostics, independent of the particle order: per chunk of
// 4096 held particles the exact integer sums of (vx^2 + vy^2) on a 2^-50 grid
// and of vx, vy on a 2^-40 grid.
__global__ void __launch_bounds__(256)
k2d_diag_parts32x(const float* __restrict__ VX, const float* __restrict__ VY, const float* __restrict__ VXl,
                  const float* __restrict__ VYl, int64_t P, i128* __restrict__ ppart) {
    __shared__ i128 red[3 * 256];
    const int64_t s = (int64_t)blockIdx.x * kTreeChunk;
    const int64_t n = P - s < kTreeChunk ? P - s : kTreeChunk;
    i128 a = 0, b = 0, c = 0;
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
        const double vx = (double)VX[s + i] + (double)VXl[s + i], vy = (double)VY[s + i] + (double)VYl[s + i];
        a += fix_round128(vx * vx + vy * vy, 50);
        b += fix_round128(vx, 40);
        c += fix_round128(vy, 40);
    }
    red[threadIdx.x] = a;
    red[256 + threadIdx.x] = b;
    red[512 + threadIdx.x] = c;
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
        ppart[3 * blockIdx.x] = red[0];
        ppart[3 * blockIdx.x + 1] = red[256];
        ppart[3 * blockIdx.x + 2] = red[512];
    }
}

__global__ void __launch_bounds__(256)
k2d_diag_final32x(double* __restrict__ gpart, int64_t ng, const i128* __restrict__ ppart, int64_t np, int64_t G,
                  int64_t P, double w, double* __restrict__ out, int64_t* __restrict__ drow, int64_t cap) {
    __shared__ double redd[kTreeLanes + 8];
    __shared__ i128 red[3 * 256];
    const double k1 = tree_strided(redd, gpart, 1, ng) / (double)G;
    i128 a = 0, b = 0, c = 0;
    for (int64_t i = threadIdx.x; i < np; i += blockDim.x) {
        a += ppart[3 * i];
        b += ppart[3 * i + 1];
        c += ppart[3 * i + 2];
    }
    red[threadIdx.x] = a;
    red[256 + threadIdx.x] = b;
    red[512 + threadIdx.x] = c;
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
        const int64_t row = drow ? *drow : 0;
        if (row < cap) {
            double* o = out + 5 * row;
            const double k2 = i128_to_double(red[0], 50) * w;
            o[0] = k1;
            o[1] = k2;
            o[2] = k1 + k2;
            o[3] = i128_to_double(red[256], 40) / (double)P;
            o[4] = i128_to_double(red[512], 40) / (double)P;
        }
        if (drow) *drow = row + 1;
    }
}

// Signed exact sum held in three 38-bit two's-complement limbs (modulo 2^114),
// rounded once on the 2^-q grid.
__device__ __forceinline__ double limbs_s_to_double(u64 a, u64 b, u64 c, int q) {
    u128 v = (u128)a + ((u128)b << kLimbS) + ((u128)c << (2 * kLimbS));
    v &= ((u128)1 << (3 * kLimbS)) - 1;
    const i128 s = (v >> (3 * kLimbS - 1)) ? (i128)v - ((i128)1 << (3 * kLimbS)) : (i128)v;
    return i128_to_double(s, q);
}

// Forward transform along x of R grid rows.  SRC 0: FP64 limbs, 1: mixed
// quanta, 3: signed limbs (each read and cleared for the next step), 2: a
// rho array.
template <int SRC>
__global__ void __launch_bounds__(kFFT2DThreads)
k2d_rows_fwd(Desc2D D, u64* __restrict__ racc, const double* __restrict__ rho, double2* __restrict__ S,
             const double2* __restrict__ tw) {
    extern __shared__ double lds2[];
    const int NX = D.NX, lx = D.lx;
    const int R = rows_per_block2d(NX, D.NY);
    const int st = line_stride(NX);
    double* re = lds2;
    double* im = lds2 + R * st;
    const int64_t G = (int64_t)NX * D.NY;
    const int64_t k0 = (int64_t)blockIdx.x * R * NX;
    for (int e = threadIdx.x; e < R * NX; e += blockDim.x) {
        const int64_t k = k0 + e;
        double v;
        if (SRC == 0) {
            const u64 a = racc[k], b = racc[G + k], c = racc[2 * G + k];
            racc[k] = 0;
            racc[G + k] = 0;
            racc[2 * G + k] = 0;
            v = fxp_to((u128)a + ((u128)b << kLimb2D) + ((u128)c << (2 * kLimb2D)), D.q);
        } else if (SRC == 1) {
            const u64 a = racc[k];
            racc[k] = 0;
            v = (double)a * D.rho_scale;
        } else if (SRC == 3) {
            const u64 a = racc[k], b = racc[G + k], c = racc[2 * G + k];
            racc[k] = 0;
            racc[G + k] = 0;
            racc[2 * G + k] = 0;
            v = limbs_s_to_double(a, b, c, D.q);
        } else {
            v = rho[k];
        }
        const int r = e >> lx, i = e & (NX - 1);
        const int pos = r * st + padi((int)brev((uint32_t)i, lx));
        re[pos] = v;
        im[pos] = 0.0;
    }
    __syncthreads();
    fft_lines(re, im, tw, lx, R, st, false);
    for (int e = threadIdx.x; e < R * NX; e += blockDim.x) {
        const int r = e >> lx, i = e & (NX - 1);
        S[k0 + e] = make_double2(re[r * st + padi(i)], im[r * st + padi(i)]);
    }
}

// Forward transform along y of C columns, the spectral field (:142-152), and
// the inverse transforms along y of Ex^, Ey^ (and phi^ when want_phi), one
// after the other through one work group of lines, so that the block holds
// two groups of C lines in LDS (the spectrum and the work lines).
__global__ void __launch_bounds__(kFFT2DThreads)
k2d_cols(Desc2D D, const double2* __restrict__ S, double2* __restrict__ T1, double2* __restrict__ T2,
         double2* __restrict__ T3, const double2* __restrict__ tw, int32_t* __restrict__ flag, int want_phi) {
    extern __shared__ double lds2[];
    const int NX = D.NX, NY = D.NY, ly = D.ly;
    const int C = cols_per_block2d(NX, NY);
    const int st = line_stride(NY);
    const int lc = __ffs(C) - 1;
    // re/im planes: line c of group a (0: spectrum, 1: work) at (a*C + c)*st
    double* re = lds2;
    double* im = lds2 + 2 * C * st;
    // blocks go round-robin to the 8 XCDs, each with its own L2: give every
    // XCD a contiguous range of columns, so the neighbouring columns that
    // share a cache line are read through one L2
    const int nb = gridDim.x;
    const int bi = (nb & 7) == 0 ? (blockIdx.x & 7) * (nb >> 3) + (blockIdx.x >> 3) : blockIdx.x;
    const int i0 = bi * C;
    for (int e = threadIdx.x; e < C * NY; e += blockDim.x) {
        const int c = e & (C - 1), j = e >> lc;
        const double2 v = S[(int64_t)j * NX + i0 + c];
        const int pos = c * st + padi((int)brev((uint32_t)j, ly));
        re[pos] = v.x;
        im[pos] = v.y;
    }
    __syncthreads();
    fft_lines(re, im, tw, ly, C, st, false);
    // the spectral field: phi^ = rho^ * minvkk in place in the spectrum
    // group (bit-reversed for the inverse), phi[1, 1] = 0
    bool bad = false;
    for (int e = threadIdx.x; e < C * NY; e += blockDim.x) {
        const int c = e & (C - 1), j = e >> lc;
        const int i = i0 + c;
        double a = re[c * st + padi(j)], b = im[c * st + padi(j)];
        const double kx = D.kxs * (double)(i < NX / 2 ? i : i - NX);
        const double ky = D.kys * (double)(j < NY / 2 ? j : j - NY);
        double cc = 0.0, d = 0.0;
        if (i == 0 && j == 0) {
            a = 0.0;  // phi[1, 1] = 0
            b = 0.0;
        } else {
            const double s = kx * kx + ky * ky;  // minvkk = -im ./ (kx^2 + ky^2) = (0/s, -1/s)
            cc = 0.0 / s;
            d = -1.0 / s;
        }
        const double tr = a * cc - b * d;
        const double ti = a * d + b * cc;
        bad |= !(isfinite(tr) && isfinite(ti));
        // the work group takes the pre-product spectrum (a, b) in bit-reversed
        // order for phi^; (tr, ti) replace the spectrum in natural order
        const int pb = padi((int)brev((uint32_t)j, ly));
        re[(C + c) * st + pb] = a;
        im[(C + c) * st + pb] = b;
        re[c * st + padi(j)] = tr;
        im[c * st + padi(j)] = ti;
    }
    if (bad) atomicOr(flag, 1);
    __syncthreads();
    // phi^ first (its lines are in place), then Ex^ = tr*kx and Ey^ = tr*ky
    // from the products kept in the spectrum group
    for (int a = want_phi ? 0 : 1; a < 3; ++a) {
        double2* T = a == 0 ? T3 : (a == 1 ? T1 : T2);
        if (a > 0) {
            for (int e = threadIdx.x; e < C * NY; e += blockDim.x) {
                const int c = e & (C - 1), j = e >> lc;
                const int i = i0 + c;
                const double k = a == 1 ? D.kxs * (double)(i < NX / 2 ? i : i - NX)
                                        : D.kys * (double)(j < NY / 2 ? j : j - NY);
                const int pb = padi((int)brev((uint32_t)j, ly));
                re[(C + c) * st + pb] = re[c * st + padi(j)] * k;
                im[(C + c) * st + pb] = im[c * st + padi(j)] * k;
            }
            __syncthreads();
        }
        fft_lines(re + C * st, im + C * st, tw, ly, C, st, true);
        for (int e = threadIdx.x; e < C * NY; e += blockDim.x) {
            const int c = e & (C - 1), j = e >> lc;
            const int pj = padi(j);
            T[(int64_t)j * NX + i0 + c] = make_double2(re[(C + c) * st + pj], im[(C + c) * st + pj]);
        }
        __syncthreads();
    }
}

// Inverse transform along x of R rows of Ex^, Ey^ (and phi^), scale
// 1/(NX*NY) (ifft, :155-156), real parts out (FP64, and FP32 pairs for the
// mixed gather).
__global__ void __launch_bounds__(kFFT2DThreads)
k2d_rows_inv(Desc2D D, const double2* __restrict__ T1, const double2* __restrict__ T2, const double2* __restrict__ T3,
             double* __restrict__ Ex, double* __restrict__ Ey, double* __restrict__ phi, float2* __restrict__ E32,
             const double2* __restrict__ tw, int want_phi) {
    extern __shared__ double lds2[];
    const int NX = D.NX, lx = D.lx;
    const int R = rows_per_block2d(NX, D.NY);
    const int st = line_stride(NX);
    const int na = want_phi ? 3 : 2;
    double* re = lds2;
    double* im = lds2 + 3 * R * st;
    const int64_t k0 = (int64_t)blockIdx.x * R * NX;
    for (int e = threadIdx.x; e < R * NX; e += blockDim.x) {
        const int r = e >> lx, i = e & (NX - 1);
        const int pos = r * st + padi((int)brev((uint32_t)i, lx));
        const double2 a = T1[k0 + e], b = T2[k0 + e];
        re[pos] = a.x;
        im[pos] = a.y;
        re[R * st + pos] = b.x;
        im[R * st + pos] = b.y;
        if (want_phi) {
            const double2 c = T3[k0 + e];
            re[2 * R * st + pos] = c.x;
            im[2 * R * st + pos] = c.y;
        }
    }
    __syncthreads();
    fft_lines(re, im, tw, lx, na * R, st, true);
    const double scale = 1.0 / (double)((int64_t)NX * D.NY);
    for (int e = threadIdx.x; e < R * NX; e += blockDim.x) {
        const int r = e >> lx, i = e & (NX - 1);
        const int pos = r * st + padi(i);
        const double ex = re[pos] * scale, ey = re[R * st + pos] * scale;
        Ex[k0 + e] = ex;
        Ey[k0 + e] = ey;
        if (E32) E32[k0 + e] = make_float2((float)ex, (float)ey);
        if (want_phi) phi[k0 + e] = re[2 * R * st + pos] * scale;
    }
}

// ---------------------------------------------------------------------------
// diagnostics (:166-170) with the canonical tree
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k2d_diag_grid(const double* __restrict__ Ex, const double* __restrict__ Ey, int64_t G, double* __restrict__ gpart) {
    __shared__ double red[kTreeLanes + 8];
    const int64_t s = (int64_t)blockIdx.x * kTreeChunk;
    const int64_t n = G - s < kTreeChunk ? G - s : kTreeChunk;
    const double r = block_chunk_sum(red, n, [&](int64_t i) {
        const double a = Ex[s + i], b = Ey[s + i];
        return a * a + b * b;
    });
    if (threadIdx.x == 0) gpart[blockIdx.x] = r;
}

// slot (may be NULL): where each caller index's particle is held, so that the
// sums run in the caller's order
template <typename T>
__global__ void __launch_bounds__(256)
k2d_diag_parts(const T* __restrict__ VX, const T* __restrict__ VY, int64_t P, double w, double* __restrict__ ppart,
               const uint32_t* __restrict__ slot) {
    __shared__ double red[3 * kTreeLanes + 8];
    const int64_t s = (int64_t)blockIdx.x * kTreeChunk;
    const int64_t n = P - s < kTreeChunk ? P - s : kTreeChunk;
    double a = 0.0, b = 0.0, c = 0.0;
    if (threadIdx.x < kTreeLanes)
        for (int64_t i = threadIdx.x; i < n; i += kTreeLanes) {
            const int64_t q = slot ? (int64_t)slot[s + i] : s + i;
            const double vx = (double)VX[q], vy = (double)VY[q];
            a = a + (vx * vx + vy * vy) * w;
            b = b + vx;
            c = c + vy;
        }
    tree256x3<false>(red, a, b, c);
    if (threadIdx.x == 0) {
        ppart[3 * blockIdx.x] = a;
        ppart[3 * blockIdx.x + 1] = b;
        ppart[3 * blockIdx.x + 2] = c;
    }
}

__global__ void __launch_bounds__(256)
k2d_diag_final(double* __restrict__ gpart, int64_t ng, double* __restrict__ ppart, int64_t np, int64_t G, int64_t P,
               double* __restrict__ out, int64_t* __restrict__ drow, int64_t cap) {
    __shared__ double red[kTreeLanes + 8];
    const double k1 = tree_strided(red, gpart, 1, ng) / (double)G;
    const double k2 = tree_strided(red, ppart, 3, np);
    const double s4 = tree_strided(red, ppart + 1, 3, np);
    const double s5 = tree_strided(red, ppart + 2, 3, np);
    if (threadIdx.x == 0) {
        const int64_t row = drow ? *drow : 0;
        if (row < cap) {
            double* o = out + 5 * row;
            o[0] = k1;
            o[1] = k2;
            o[2] = k1 + k2;
            o[3] = s4 / (double)P;
            o[4] = s5 / (double)P;
        }
        if (drow) *drow = row + 1;
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static int push_blocks(int64_t P) {
    const int64_t b = (P + kPush2DThreads - 1) / kPush2DThreads;
    return (int)(b < 8192 ? (b > 0 ? b : 1) : 8192);
}

void launch2d_push(const Buffers2D& b, hipStream_t s) {
    if (b.precision == 1) {
        const double t1 = b.B0 * b.dt / 2.0;
        const double tscale = 2.0 / (1.0 + t1 * t1);
        k2d_push32t<<<b.T.ntiles * b.T.splits, kTileThreads, 0, s>>>(
            b.D, b.T, b.Xq, b.Yq, b.VXf, b.VYf, b.VZf, b.VXl, b.VYl, b.VZl, b.E32, b.racc, (float)(b.dt / 2.0), (float)t1,
            (float)tscale, (float)(b.dt * 4294967296.0));
    } else {
        k2d_push64t<<<b.T.ntiles * b.T.splits, kTileThreads, 0, s>>>(b.D, b.T, b.X, b.Y, b.VX, b.VY, b.VZ, b.Ex, b.Ey,
                                                                     b.racc, b.dt, b.B0);
    }
}

void launch2d_tile_sort(const Buffers2D& b, hipStream_t s) {
    const size_t lds = sizeof(int32_t) * (size_t)b.T.ntiles;
    const int64_t per = 16384;
    const int nb = (int)((b.D.P + per - 1) / per);
    if (b.precision == 1) {
        k2d_tile_count<<<nb, 256, lds, s>>>(b.D, b.T, b.Xq, b.Yq);
        k2d_tile_scan<<<1, 1024, 0, s>>>(b.T);
        k2d_tile_scatter<<<nb, 256, lds, s>>>(b.D, b.T, b.Xq, b.Yq, b.VXf, b.VYf, b.VZf, b.VXl, b.VYl, b.VZl, b.pid,
                                              b.Xq_alt, b.Yq_alt, b.VXf_alt, b.VYf_alt, b.VZf_alt, b.VXl_alt,
                                              b.VYl_alt, b.VZl_alt, b.pid_alt, per);
    } else {
        k2d_tile_count64<<<nb, 256, lds, s>>>(b.D, b.T, b.X, b.Y);
        k2d_tile_scan<<<1, 1024, 0, s>>>(b.T);
        k2d_tile_scatter64<<<nb, 256, lds, s>>>(b.D, b.T, b.X, b.Y, b.VX, b.VY, b.VZ, b.pid, b.X_alt, b.Y_alt,
                                                b.VX_alt, b.VY_alt, b.VZ_alt, b.pid_alt, b.slot, per);
    }
}

void launch2d_solve(const Buffers2D& b, int src, const double* rho_in, bool want_phi, hipStream_t s) {
    const int NX = b.D.NX, NY = b.D.NY;
    const int nr = NY / rows_per_block2d(NX, NY);
    const int nc = NX / cols_per_block2d(NX, NY);
    const size_t lr1 = rows_lds_bytes(NX, NY, 1), lr3 = rows_lds_bytes(NX, NY, 3), lc = cols_lds_bytes(NX, NY);
    if (src == 0)
        k2d_rows_fwd<0><<<nr, kFFT2DThreads, lr1, s>>>(b.D, b.racc, nullptr, b.S, b.tw);
    else if (src == 1)
        k2d_rows_fwd<1><<<nr, kFFT2DThreads, lr1, s>>>(b.D, b.racc, nullptr, b.S, b.tw);
    else if (src == 3)
        k2d_rows_fwd<3><<<nr, kFFT2DThreads, lr1, s>>>(b.D, b.racc, nullptr, b.S, b.tw);
    else
        k2d_rows_fwd<2><<<nr, kFFT2DThreads, lr1, s>>>(b.D, nullptr, rho_in, b.S, b.tw);
    k2d_cols<<<nc, kFFT2DThreads, lc, s>>>(b.D, b.S, b.T1, b.T2, b.T3, b.tw, b.flag, want_phi ? 1 : 0);
    k2d_rows_inv<<<nr, kFFT2DThreads, lr3, s>>>(b.D, b.T1, b.T2, b.T3, b.Ex, b.Ey, b.phi, b.E32, b.tw,
                                                want_phi ? 1 : 0);
}

void launch2d_diag_op(const double* Ex, const double* Ey, int64_t G, const double* vx, const double* vy, int64_t P,
                      double w, double* gpart, double* ppart, double* out, int64_t* drow, hipStream_t s) {
    const int64_t ng = (G + kTreeChunk - 1) / kTreeChunk, np = (P + kTreeChunk - 1) / kTreeChunk;
    k2d_diag_grid<<<(unsigned)ng, 256, 0, s>>>(Ex, Ey, G, gpart);
    k2d_diag_parts<double><<<(unsigned)np, 256, 0, s>>>(vx, vy, P, w, ppart, nullptr);
    k2d_diag_final<<<1, 256, 0, s>>>(gpart, ng, ppart, np, G, P, out, drow, drow ? INT64_MAX : 1);
}

void launch2d_diag(const Buffers2D& b, hipStream_t s) {
    const int64_t G = (int64_t)b.D.NX * b.D.NY, P = b.D.P;
    const int64_t ng = (G + kTreeChunk - 1) / kTreeChunk, np = (P + kTreeChunk - 1) / kTreeChunk;
    k2d_diag_grid<<<(unsigned)ng, 256, 0, s>>>(b.Ex, b.Ey, G, b.gpart);
    if (b.precision == 1) {
        // exact sums: the held particle order changes with every tile sort
        i128* pp = reinterpret_cast<i128*>(b.ppart);
        k2d_diag_parts32x<<<(unsigned)np, 256, 0, s>>>(b.VXf, b.VYf, b.VXl, b.VYl, P, pp);
        k2d_diag_final32x<<<1, 256, 0, s>>>(b.gpart, ng, pp, np, G, P, b.D.w, b.diag, b.drow, b.diag_cap);
        return;
    }
    k2d_diag_parts<double><<<(unsigned)np, 256, 0, s>>>(b.VX, b.VY, P, b.D.w, b.ppart, b.slot);
    k2d_diag_final<<<1, 256, 0, s>>>(b.gpart, ng, b.ppart, np, G, P, b.diag, b.drow, b.diag_cap);
}

__global__ void __launch_bounds__(256) k2d_limbs_to_rho(Desc2D D, u64* __restrict__ racc, double* __restrict__ rho) {
    const int64_t G = (int64_t)D.NX * D.NY;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < G; k += (int64_t)gridDim.x * blockDim.x) {
        const u64 a = racc[k], b = racc[G + k], c = racc[2 * G + k];
        racc[k] = 0;
        racc[G + k] = 0;
        racc[2 * G + k] = 0;
        rho[k] = fxp_to((u128)a + ((u128)b << kLimb2D) + ((u128)c << (2 * kLimb2D)), D.q);
    }
}

// depositcharge! of P particles at (x, y) into rho (clears the accumulator).
// D.q must be the fraction bits for these P and w.
void launch2d_deposit_op(const Buffers2D& b, const double* x, const double* y, int64_t P, double* rho,
                         hipStream_t s) {
    Desc2D D = b.D;
    D.P = P;
    k2d_push64<2><<<push_blocks(P), kPush2DThreads, 0, s>>>(D, const_cast<double*>(x), const_cast<double*>(y),
                                                         nullptr, nullptr, nullptr, nullptr, nullptr, b.racc, 0.0,
                                                         0.0);
    const int64_t G = (int64_t)D.NX * D.NY;
    k2d_limbs_to_rho<<<(unsigned)((G + 255) / 256), 256, 0, s>>>(D, b.racc, rho);
}

void launch2d_gather_op(const Buffers2D& b, const double* Ex, const double* Ey, const double* x, const double* y,
                        int64_t P, double* ex, double* ey, hipStream_t s) {
    k2d_gather64<<<push_blocks(P), kPush2DThreads, 0, s>>>(b.D, Ex, Ey, x, y, ex, ey, P);
}

void launch2d_boris_op(double* vx, double* vy, double* vz, const double* ex, const double* ey, int64_t P, double dt,
                       double B0, hipStream_t s) {
    k2d_boris64<<<push_blocks(P), kPush2DThreads, 0, s>>>(vx, vy, vz, ex, ey, P, dt, B0);
}

void launch2d_push_op(const Buffers2D& b, double* x, double* y, double* vx, double* vy, double* vz, int64_t P,
                      const double* Ex, const double* Ey, double dt, double B0, hipStream_t s) {
    Desc2D D = b.D;
    D.P = P;
    k2d_push64<1><<<push_blocks(P), kPush2DThreads, 0, s>>>(D, x, y, vx, vy, vz, Ex, Ey, nullptr, dt, B0);
}

int configure_kernels_2d() {
    const int bytes = (int)cols_lds_bytes(kMax2D, kMax2D);  // the largest: one 1024-point column per block
    const void* fns[] = {reinterpret_cast<const void*>(k2d_cols),
                         reinterpret_cast<const void*>(k2d_rows_inv),
                         reinterpret_cast<const void*>(k2d_rows_fwd<0>),
                         reinterpret_cast<const void*>(k2d_rows_fwd<1>),
                         reinterpret_cast<const void*>(k2d_rows_fwd<2>),
                         reinterpret_cast<const void*>(k2d_rows_fwd<3>)};
    for (const void* f : fns) (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    (void)hipGetLastError();
    return 0;
}

}  // namespace pic
