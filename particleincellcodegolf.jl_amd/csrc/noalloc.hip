// noalloc.hip -- pic_deposit_noalloc: the deposit r(x, y) of pic_deposit on
// device-resident arrays, enqueued on the caller's stream, allocating nothing.
//
// Its one piece of scratch memory, the per-cell exact charge accumulator (three
// 43-bit limbs per cell), is the file-scope device array `scratch` below: each
// GPU holds its own copy for the life of the process.  The host side takes its
// address once per device with hipGetSymbolAddress and passes that address to
// the kernels.  A call is three kernels -- clear the limbs, deposit (LDS limbs
// per workgroup, added into the array by device-scope integer atomics),
// convert to FP64 -- and no memset, synchronisation or allocation, so it can
// be captured in a graph.  The result is bit for bit that of pic_deposit.
//
// Calls that overlap on one GPU (two streams, or two threads) share the array
// and race: the caller serialises them.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <string>

#include "../../include/pic.h"
#include "pic_device.h"
#include "pic_internal.h"
#include "stencil1d.h"
#include "trace.h"

#pragma clang fp contract(off)

namespace pic {

constexpr int kNoallocMaxN = kMaxN;                           // cells
constexpr size_t kScratchWords = 3 * (size_t)kNoallocMaxN;    // 96 KiB of u64 limbs
constexpr int kNoallocTileP = 16384;                          // particles per workgroup
constexpr int kMaxDevices = 64;

static __device__ __align__(256) u64 scratch[kScratchWords];

__global__ void k_fill_u64(u64* __restrict__ p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// Deposit of one tile of particles into LDS limbs, then into the device limbs.
template <int R>
__global__ void __launch_bounds__(256)
k_deposit_noalloc(const double* __restrict__ x, const double* __restrict__ y, int64_t P, double w, int q, int N,
                  u64* __restrict__ acc) {
    extern __shared__ u64 lds[];
    zero_lds(lds, R == 0 ? N : 3 * N);
    __syncthreads();
    const double Nd = (double)N;
    const int32_t nmask = N - 1;
    const int64_t start = (int64_t)blockIdx.x * kNoallocTileP;
    const int64_t end = start + kNoallocTileP < P ? start + kNoallocTileP : P;
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
    if (R == 0) {
        const u128 one = fxp_from(w, q);
        for (int c = threadIdx.x; c < N; c += blockDim.x)
            if (lds[c]) flush_charge64((u128)lds[c] * one, acc + 3 * c);
    } else {
        flush_limbs(lds, N, acc);
    }
}

__global__ void k_limbs_to_rho(const u64* __restrict__ acc, int N, int q, double* __restrict__ rho) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < N) rho[c] = fxp_to(limbs_to_u128(acc[3 * c], acc[3 * c + 1], acc[3 * c + 2]), q);
}

// Device address of `scratch` per device, taken once with that device current.
static std::once_flag g_once[kMaxDevices];
static void* g_scratch[kMaxDevices];
static hipError_t g_scratch_err[kMaxDevices];

// Set-up per device: the address of `scratch`, and the dynamic LDS limit of
// the deposit kernels (3 N limbs of 8 bytes at N = 4096 exceed the default).
static hipError_t scratch_address(int dev, void** out) {
    std::call_once(g_once[dev], [dev] {
        void* ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        const int lds = (int)(3 * sizeof(u64) * kNoallocMaxN);
        for (const void* f : {reinterpret_cast<const void*>(k_deposit_noalloc<0>),
                              reinterpret_cast<const void*>(k_deposit_noalloc<1>),
                              reinterpret_cast<const void*>(k_deposit_noalloc<6>),
                              reinterpret_cast<const void*>(k_deposit_noalloc<7>)}) {
            const hipError_t a = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
            if (e == hipSuccess) e = a;
        }
        g_scratch_err[dev] = e;
        g_scratch[dev] = ptr;
    });
    *out = g_scratch[dev];
    return g_scratch_err[dev];
}

static thread_local std::string g_noalloc_err;

static int fail(int code, const char* msg) {
    g_noalloc_err = msg;
    return code;
}

}  // namespace pic

using namespace pic;

extern "C" const char* pic_noalloc_last_error(void) { return g_noalloc_err.c_str(); }

extern "C" int pic_deposit_noalloc(const double* x, const double* y, int64_t P, double w, int32_t N, int32_t shape,
                                   double* rho, void* stream) {
    PIC_TRACE("pic_deposit_noalloc");
    if (!x || !rho || P < 1) return fail(PIC_EINVAL, "pic_deposit_noalloc: null array or P < 1");
    if (N < 8 || N > kNoallocMaxN || (N & (N - 1)) != 0)
        return fail(PIC_EINVAL, "pic_deposit_noalloc: N must be a power of two in [8, 4096] (the scratch holds "
                                "3 x 4096 limbs)");
    if (shape != PIC_SHAPE_NGP && shape != PIC_SHAPE_AREA && shape != PIC_SHAPE_GAUSS6 && shape != PIC_SHAPE_GAUSS7)
        return fail(PIC_EINVAL, "pic_deposit_noalloc: unknown shape");
    if (!(w >= 0.0) || !std::isfinite(w)) return fail(PIC_EINVAL, "pic_deposit_noalloc: w must be finite and >= 0");
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess || dev < 0 || dev >= kMaxDevices) return fail(PIC_EHIP, "pic_deposit_noalloc: no device");
    void* sp = nullptr;
    if ((e = scratch_address(dev, &sp)) != hipSuccess || !sp) {
        g_noalloc_err = std::string("pic_deposit_noalloc: hipGetSymbolAddress failed: ") + hipGetErrorString(e);
        return PIC_EHIP;
    }
    u64* acc = static_cast<u64*>(sp);
    hipStream_t s = static_cast<hipStream_t>(stream);
    // q as pic_deposit chooses it: P*w*2^q < 2^125
    int q = 64;
    const double total = (double)P * w;
    if (total > 0.0 && std::isfinite(total)) {
        q = 124 - std::ilogb(total);
        if (q > 1100) q = 1100;
    }
    const int64_t words = 3 * (int64_t)N;
    k_fill_u64<<<(unsigned)((words + 255) / 256), 256, 0, s>>>(acc, words);
    const int64_t tiles = (P + kNoallocTileP - 1) / kNoallocTileP;
    if (tiles > 0) {
        const int R = shape == PIC_SHAPE_NGP ? 0 : shape;
        const size_t lds = sizeof(u64) * (R == 0 ? (size_t)N : 3 * (size_t)N);
        const dim3 g((unsigned)tiles), t(256);
        if (R == 0) k_deposit_noalloc<0><<<g, t, lds, s>>>(x, y, P, w, q, N, acc);
        else if (R == 1) k_deposit_noalloc<1><<<g, t, lds, s>>>(x, y, P, w, q, N, acc);
        else if (R == 6) k_deposit_noalloc<6><<<g, t, lds, s>>>(x, y, P, w, q, N, acc);
        else k_deposit_noalloc<7><<<g, t, lds, s>>>(x, y, P, w, q, N, acc);
    }
    k_limbs_to_rho<<<(unsigned)((N + 255) / 256), 256, 0, s>>>(acc, N, q, rho);
    if ((e = hipGetLastError()) != hipSuccess) {
        g_noalloc_err = std::string("pic_deposit_noalloc: launch failed: ") + hipGetErrorString(e);
        return PIC_EHIP;
    }
    return PIC_OK;
}
