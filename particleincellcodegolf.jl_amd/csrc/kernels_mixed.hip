// kernels_mixed.hip -- mixed-precision kernels of the implicit-midpoint step
// (src/GaussianFixedPoint.jl:7-9, src/GaussianFixedPointQuiet.jl:8-10).
//
// Particle state: position as a 32-bit fixed-point fraction of the periodic
// box (2^-32 resolution, wrap-around for free, cell index = top log2 N bits,
// sub-cell offset = the rest), velocity as an FP32 hi + lo pair (the lo word
// is read and written once per step to keep the time integration compensated),
// Picard iterate in FP32.  Shape weights are FP32 polynomials in the sub-cell
// offset (weights32.h); the gather sums in FP32; the charge is accumulated
// exactly as integer multiples of 2^-30 of a particle's charge, so the per-cell
// total is exact and independent of scheduling.  The FFT, the convergence
// test and the diagnostics stay FP64 (kernels_1d.hip).
#include "pic_device.h"
#include "pic_internal.h"
#include "weights32.h"

namespace pic {

constexpr float kQuantum = 1073741824.0f;  // 2^30 weight quanta per particle

// Stencil of a fixed-point centre: i0 = nearest cell centre (Julia index,
// N == 0), d = offset of the centre from it in cells, in [-1/2, 1/2).
struct Stencil32 {
    int32_t i0;
    float w[2 * kStencil32 + 1];
};

__device__ __forceinline__ void stencil32(uint32_t cq, int shift, Stencil32& s) {
    const uint64_t t = (uint64_t)cq + (1ull << (shift - 1));
    const uint32_t i0 = (uint32_t)(t >> shift);
    const int32_t dint = (int32_t)(cq - (i0 << shift));
    const float d = ldexpf((float)dint, -shift);
    const float u = d * d;
    s.i0 = (int32_t)i0;
    s.w[kStencil32] = wA0(u);
#define PIC_PAIR(J)                                \
    {                                              \
        const float a = wA##J(u), b = d * wB##J(u); \
        s.w[kStencil32 + J] = a + b;               \
        s.w[kStencil32 - J] = a - b;               \
    }
    PIC_PAIR(1)
    PIC_PAIR(2)
    PIC_PAIR(3)
    PIC_PAIR(4)
    PIC_PAIR(5)
#undef PIC_PAIR
}

__device__ __forceinline__ uint32_t advance(uint32_t xq, float v_sum, float scale) {
    return xq + (uint32_t)(int32_t)rintf(v_sum * scale);
}

template <int MODE>
__global__ void __launch_bounds__(1024)
k_mid_push32(const InstDesc* __restrict__ insts, InstState* __restrict__ st,
             const TileDesc* __restrict__ tiles, uint32_t* Xq0, uint32_t* Xq1, float* Vh0,
             float* Vh1, float* Vl0, float* Vl1, float* vf0, float* vf1,
             const float* __restrict__ E32, u64* __restrict__ part, int k) {
    extern __shared__ u64 lds[];
    const TileDesc T = tiles[blockIdx.x];
    const InstDesc I = insts[T.inst];
    InstState S = st[T.inst];
    if (MODE == 0) {
        if (T.local == 0 && threadIdx.x == 0) st[T.inst].last_sweep = -1;
        S.last_sweep = -1;
    }
    if (MODE == 1 && S.last_sweep >= 0 && S.last_sweep < k) return;
    const bool last = (MODE == 1) && (S.last_sweep == k);
    const int32_t N = I.N, nmask = I.nmask;
    u64* acc = lds;
    float* Eld = reinterpret_cast<float*>(lds + N);
    if (MODE == 1)
        for (int i = threadIdx.x; i < N; i += blockDim.x) Eld[i] = E32[I.goff + i];
    if (!last)
        for (int i = threadIdx.x; i < N; i += blockDim.x) acc[i] = 0;
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
    const float dtf = (float)I.dt;
    const float cscale = (float)(I.dt * 1073741824.0);  // (v + V)/4*dt in 2^-32 units
    const float xscale = (float)(I.dt * 2147483648.0);  // (v + V)/2*dt in 2^-32 units
    const int64_t base = I.poff + T.start;
    for (int64_t p = threadIdx.x; p < T.count; p += blockDim.x) {
        const int64_t g = base + p;
        const uint32_t xq = Xq[g];
        const float vh = Vh[g];
        float vn = vh;
        Stencil32 s;
        if (MODE == 1) {
            const float vk = (k == 0) ? vh : vin[g];
            stencil32(advance(xq, vk + vh, cscale), shift, s);
            float acc_e = 0.0f;
#pragma unroll
            for (int j = -kStencil32; j <= kStencil32; ++j)
                acc_e = fmaf(Eld[(s.i0 + j - 1) & nmask], s.w[j + kStencil32], acc_e);
            const float kick = acc_e * dtf;
            vn = vh + kick;
            if (last) {
                // V_new = (V_hi + V_lo) + kick, renormalised (two-sum)
                const float sum = vh + kick;
                const float bb = sum - vh;
                const float err = (vh - (sum - bb)) + (kick - bb);
                const float t = Vl[g] + err;
                const float hi = sum + t;
                Xo[g] = advance(xq, vk + vh, xscale);
                Vho[g] = hi;
                Vlo[g] = t - (hi - sum);
                continue;
            }
            vout[g] = vn;
        }
        stencil32(advance(xq, vn + vh, cscale), shift, s);
#pragma unroll
        for (int j = -kStencil32; j <= kStencil32; ++j) {
            const uint32_t qv = (uint32_t)rintf(fmaxf(s.w[j + kStencil32], 0.0f) * kQuantum);
            if (qv) atomicAdd(&acc[(s.i0 + j - 1) & nmask], (u64)qv);
        }
    }
    if (last) return;
    __syncthreads();
    u64* out = part + T.part_off;
    for (int c = threadIdx.x; c < N; c += blockDim.x) out[c] = acc[c];
}

// Charge reduction, stage 1 (see k_reduce64): u64 sums of weight quanta.
__global__ void __launch_bounds__(256)
k_reduce32(const InstDesc* __restrict__ insts, const InstState* __restrict__ st,
           const RedDesc* __restrict__ reds, const u64* __restrict__ part, u64* __restrict__ part2, int k) {
    const RedDesc D = reds[blockIdx.x];
    const InstDesc I = insts[D.inst];
    const InstState S = st[D.inst];
    if (S.last_sweep >= 0 && S.last_sweep < k) return;
    const int c = D.c0 + threadIdx.x;
    if (c >= I.N) return;
    const int t0 = D.g * I.tpg;
    const int t1 = t0 + I.tpg < I.ntiles ? t0 + I.tpg : I.ntiles;
    u64 s = 0;
    const u64* p = part + I.part0 + c;
    for (int t = t0; t < t1; ++t) s += p[(int64_t)t * I.N];
    part2[I.red0 + (int64_t)D.g * I.N + c] = s;
}

__global__ void __launch_bounds__(256)
k_diag_chunks32(const InstDesc* __restrict__ insts, const InstState* __restrict__ st,
                const ChunkDesc* __restrict__ chunks, const float* Vh0, const float* Vh1,
                const float* Vl0, const float* Vl1, double* __restrict__ dpart) {
    __shared__ double red[kTreeLanes + 1];
    const ChunkDesc C = chunks[blockIdx.x];
    const InstDesc I = insts[C.inst];
    const InstState S = st[C.inst];
    const int nxt = (S.step + 1) & 1;
    const float* vh = (nxt ? Vh1 : Vh0) + I.poff + C.start;
    const float* vl = (nxt ? Vl1 : Vl0) + I.poff + C.start;
    const double Pd = (double)I.P;
    const double s2 = block_chunk_sum(red, C.count, [&](int64_t i) {
        const double a = (double)vh[i] + (double)vl[i];
        return a * a;
    });
    const double s1 = block_chunk_sum(red, C.count, [&](int64_t i) {
        return ((double)vh[i] + (double)vl[i]) / Pd;
    });
    if (threadIdx.x == 0) {
        dpart[2 * blockIdx.x] = s2;
        dpart[2 * blockIdx.x + 1] = s1;
    }
}

void launch_mid_push32(const RunBuffers& b, int mode, int k, hipStream_t s) {
    const dim3 g(b.ntiles), t(b.block_push);
    u64* part = (u64*)b.part;
    if (mode == 0)
        k_mid_push32<0><<<g, t, b.lds_push, s>>>(b.insts, b.st, b.tiles, b.Xq[0], b.Xq[1], b.Vh[0], b.Vh[1],
                                                 b.Vl[0], b.Vl[1], b.vf[0], b.vf[1], b.E32, part, 0);
    else
        k_mid_push32<1><<<g, t, b.lds_push, s>>>(b.insts, b.st, b.tiles, b.Xq[0], b.Xq[1], b.Vh[0], b.Vh[1],
                                                 b.Vl[0], b.Vl[1], b.vf[0], b.vf[1], b.E32, part, k);
}

void launch_reduce32(const RunBuffers& b, int k, hipStream_t s) {
    k_reduce32<<<b.nreds, 256, 0, s>>>(b.insts, b.st, b.reds, (const u64*)b.part, (u64*)b.part2, k);
}

void launch_diag_chunks32(const RunBuffers& b, hipStream_t s) {
    k_diag_chunks32<<<b.nchunks, 256, 0, s>>>(b.insts, b.st, b.chunks, b.Vh[0], b.Vh[1], b.Vl[0], b.Vl[1],
                                              b.dpart);
}

int configure_kernels_mixed() {
    const int bytes = 12 * kMaxN;
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(k_mid_push32<0>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(k_mid_push32<1>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    (void)hipGetLastError();
    return 0;
}

}  // namespace pic
