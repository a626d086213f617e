// picem1d_internal.h -- descriptors and launchers of the 1D electromagnetic
// fixed-point and explicit steps, shared by the host runtime (capi_em1d.cpp) and the kernels
// (kernels_em1d.hip).
#pragma once
#include "pic_internal.h"

namespace pic {

constexpr int kMaxNEM = 1024;
constexpr int64_t kMaxPEM = int64_t(1) << 26;
constexpr int kEMThreads = 256;
// the charge-and-currents deposit holds a whole grid of limbs in LDS per
// block (9N u64, 72 KB at N = 1024: two blocks per CU) and 146 registers per
// lane (three waves per SIMD): 768-thread blocks run three waves per SIMD
// where two 256-thread blocks ran two
constexpr int kEMDepThreads = 768;
constexpr int kEMSolveThreads = 256;
constexpr int kEMMomentGroups = 4;  // moment grids per group: 2, 2, 2, 1
constexpr int kEMScanTile = 4096;
constexpr int kEMSmallSort = 4096;  // particles the one-block sort takes (the bucket sort beyond)
constexpr int kLimbEM = 38;         // exact sums: three 38-bit two's-complement limbs per cell
constexpr int kEMMaxSweeps = 8;     // ctl[s] = sweep s active; ctl[kEMMaxSweeps] = sweeps run this step

// Device state of one context (all pointers device memory).
struct EM1D {
    int32_t N, lN, max_sweeps, tile;
    int64_t P;
    double eps, dt, w, rtol;
    double *x, *v1, *v2, *v3;   // particles
    double *X, *V1, *V2, *V3;   // their values at the start of the step
    double* c;                  // the sweep's midpoints (X + x)/2
    double *Ex, *Ey, *Ez, *By, *Bz;
    double *Fx, *Fy, *Fz, *Cy, *Cz;
    double *ex, *ey, *ez, *by, *bz;
    unsigned long long* racc;   // 10 grids x 3 limbs x N: rho, Jy, Jz, then the 7 moments
    unsigned long long* umax;   // [8]: max |uy|, |uz| of the sweep; max |v1|, |v2|, |v3| of the step's end
    int32_t* ctl;               // [kEMMaxSweeps + 1]: sweep s active, sweeps run this step
    int64_t* t;                 // [2]: steps since upload, step of the run call's first output slot
    double* out;                // out_cap slots of 12 x N
    int32_t* sweeps_out;        // out_cap
    int64_t out_cap;
    const double2* tw;
    // the per-step sort
    int32_t nb;                 // buckets (power of two)
    uint32_t *count, *cursor, *start, *bsum, *bucket, *sidx, *sbucket, *perm;
    unsigned long long *vkey, *skey;
    double* tmp;                // 4 x P
};

void em_launch_sweep(const EM1D& s, int sweep, hipStream_t st);
void em_launch_end(const EM1D& s, bool sort, bool moments, hipStream_t st);
// one step of the explicit scheme
void emx_launch_step(const EM1D& s, bool sort, bool moments, hipStream_t st);
// operators
void em_launch_deposit_op(const EM1D& s, const double* y, const double* u, int64_t n, double w, double* rho,
                          unsigned long long* scratch_max, hipStream_t st);
// one sweep's field updates from rho, Jy, Jz arrays (s points at scratch copies)
void em_launch_fields_op(const EM1D& s, const double* rho, const double* Jy, const double* Jz, hipStream_t st);
void em_launch_push_op(const EM1D& s, const double* c, const double* u, int64_t n, double* v, hipStream_t st);
// the explicit scheme's field updates (Gauss, Faraday, then Ampere) from rho,
// Jy, Jz arrays, and its velocity update at positions x
void emx_launch_fields_op(const EM1D& s, const double* rho, const double* Jy, const double* Jz, hipStream_t st);
void emx_launch_push_op(const EM1D& s, const double* x, const double* u, int64_t n, double* v, hipStream_t st);
int configure_kernels_em1d();
size_t em_solve_lds_bytes(int N);

}  // namespace pic
