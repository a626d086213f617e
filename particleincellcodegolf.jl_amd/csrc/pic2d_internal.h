// pic2d_internal.h -- descriptors and launchers of the 2D3V step shared by
// the host runtime (capi2d.cpp) and the kernels (kernels_2d.hip).
#pragma once
#include "pic_internal.h"

namespace pic {

constexpr int kMax2D = 1024;        // largest NX, NY
constexpr int kLine2D = 512;        // transform points per block (rows*NX or cols*NY, >= one line)
constexpr int kFFT2DThreads = 256;
constexpr int kPush2DThreads = 256;
constexpr int kLimb2D = 42;         // FP64 exact deposit: 3 limbs of 42 bits (P <= 2^22)
constexpr int kLimbS = 38;          // signed exact deposit: 3 limbs of 38 bits, two's complement
                                    // (114 bits; up to 2^26 contributions per cell)
constexpr int kQuanta2D = 20;       // mixed: area weight quanta per axis, 2^-20 of a cell
constexpr int64_t kMaxP2D64 = int64_t(1) << 22;
constexpr int64_t kMaxP2D32 = int64_t(1) << 24;
// particles held in tiles of kTile2D x kTile2D cells; a push
// block stages its tile's field and charge with kHalo2D cells around it
constexpr int kTile2D = 16;
constexpr int kHalo2D = 4;
constexpr int kTileBlockP = 4096;   // particles per push block (target)
constexpr int kTileThreads = 256;
constexpr int kSortEvery2D = 8;     // default steps between tile sorts

struct Tiles2D {
    int32_t TX, TY, ntx, nty, ntiles, splits;
    int32_t* count;   // [ntiles] particles per tile (sort scratch, zero between sorts)
    int32_t* start;   // [ntiles + 1] first held particle of each tile
    int32_t* cursor;  // [ntiles] scatter cursors
};

struct Desc2D {
    int32_t NX, NY, lx, ly;  // grid and log2 sizes
    int64_t P;
    int32_t q;               // FP64: fraction bits of the exact per-cell sums
    int32_t pairing;         // 0: cell i with i+1 (reference), 1: with i-1
    int32_t sx, sy;          // mixed: 32 - lx, 32 - ly (fixed-point cell shift)
    double w;                // particle weight
    double rho_scale;        // mixed: charge of one weight quantum, w * 2^-40
    double kxs, kys;         // wavenumber scales: kx = kxs*m (2pi/Lx), ky = kys*m
};

// Device buffers of one 2D context.  Grids are NX*NY with x fastest.
struct Buffers2D {
    Desc2D D;
    int precision;
    int blocks_push;
    double dt, B0;
    // particles (FP64 or mixed, by precision)
    double *X, *Y, *VX, *VY, *VZ;
    uint32_t *Xq, *Yq;
    float *VXf, *VYf, *VZf;  // mixed: velocity hi words ...
    float *VXl, *VYl, *VZl;  // ... and lo words (v = hi + lo, |lo| <= ulp(hi)/2)
    uint32_t* pid;           // caller's index of each held particle
    // the other buffer of each particle array (tile sorts move the particles
    // there)
    uint32_t *Xq_alt, *Yq_alt, *pid_alt;
    float *VXf_alt, *VYf_alt, *VZf_alt;
    float *VXl_alt, *VYl_alt, *VZl_alt;
    double *X_alt, *Y_alt, *VX_alt, *VY_alt, *VZ_alt;
    uint32_t* slot;          // FP64: the slot holding caller index i (diagnostics run in the caller's order)
    Tiles2D T;
    unsigned long long* racc;  // FP64: 3 limbs x NX*NY; mixed: NX*NY quanta
    double2 *S, *T1, *T2, *T3;
    double *Ex, *Ey, *phi;
    float2* E32;             // (Ex, Ey) in FP32 for the mixed gather
    const double2* tw;       // per-pass twiddles
    double *gpart, *ppart;   // diagnostics partials (grid chunks, 3 per particle chunk)
    double* diag;            // rows of 5
    int64_t* drow;           // next diagnostics row (device counter)
    int64_t diag_cap;
    int32_t* flag;           // bit 0: non-finite field
};

// Line transforms: a block holds R = kLine2D/NX rows (or C = kLine2D/NY
// columns), at least one, at most the grid; lines are padded by one slot per
// 16 doubles in LDS.
__host__ __device__ inline int line_stride(int L) { return L + (L >> 4); }
__host__ __device__ inline int rows_per_block2d(int NX, int NY) {
    const int r = kLine2D / NX < 1 ? 1 : kLine2D / NX;
    return r < NY ? r : NY;
}
__host__ __device__ inline int cols_per_block2d(int NX, int NY) {
    const int c = kLine2D / NY < 1 ? 1 : kLine2D / NY;
    return c < NX ? c : NX;
}
// re and im planes of `arrays` groups of R rows / of the column kernel's two
// groups (spectrum, work) of C columns
inline size_t rows_lds_bytes(int NX, int NY, int arrays) {
    return sizeof(double) * 2 * (size_t)arrays * rows_per_block2d(NX, NY) * line_stride(NX);
}
inline size_t cols_lds_bytes(int NX, int NY) {
    return sizeof(double) * 2 * 2 * (size_t)cols_per_block2d(NX, NY) * line_stride(NY);
}

// one time step: push + deposit, then the field solve
void launch2d_push(const Buffers2D& b, hipStream_t s);
// mixed: regroup the held particles by tile into the _alt buffers
void launch2d_tile_sort(const Buffers2D& b, hipStream_t s);
// src: 0 = FP64 limbs, 1 = mixed quanta, 2 = rho array, 3 = signed 38-bit limbs
void launch2d_solve(const Buffers2D& b, int src, const double* rho_in, bool want_phi, hipStream_t s);
// diagnostics from b.Ex/b.Ey and the particle velocities, appended at *drow
void launch2d_diag(const Buffers2D& b, hipStream_t s);
// operators (FP64)
void launch2d_deposit_op(const Buffers2D& b, const double* x, const double* y, int64_t P, double* rho,
                         hipStream_t s);
void launch2d_gather_op(const Buffers2D& b, const double* Ex, const double* Ey, const double* x, const double* y,
                        int64_t P, double* ex, double* ey, hipStream_t s);
void launch2d_boris_op(double* vx, double* vy, double* vz, const double* ex, const double* ey, int64_t P,
                       double dt, double B0, hipStream_t s);
void launch2d_push_op(const Buffers2D& b, double* x, double* y, double* vx, double* vy, double* vz, int64_t P,
                      const double* Ex, const double* Ey, double dt, double B0, hipStream_t s);
void launch2d_diag_op(const double* Ex, const double* Ey, int64_t G, const double* vx, const double* vy, int64_t P,
                      double w, double* gpart, double* ppart, double* out, int64_t* drow, hipStream_t s);
int configure_kernels_2d();

}  // namespace pic
