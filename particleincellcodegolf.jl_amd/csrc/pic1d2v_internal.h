// pic1d2v_internal.h -- device buffers and launchers of the magnetised 1D2V
// step (src/NGP1D2V.jl, src/NGP1D2V2S.jl).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pic_internal.h"

namespace pic {

// Device-side loop state of a 1D2V context.
struct State12 {
    int64_t step;  // steps completed since the upload
    double se;     // sum(E.^2) of the latest field (canonical tree)
    int32_t err;   // bit 0: non-finite charge or field
    int32_t pad;
};

// A run of at most `count` particles of one species from `start` in the
// packed arrays (species s particle j at s*P + j), pushed by one workgroup.
struct Tile12 {
    int32_t species;
    int32_t pad;
    int64_t start, count;
};

struct Buffers12 {
    int32_t N, log2N, R, ns, bin, qbits;
    int64_t P;
    double q[2], q_m[2], mass[2];
    double w, n0, B0, dt;
    double *x, *vx, *vy;           // ns*P each
    unsigned long long* racc;      // 3 limbs per cell of the positive charge, then of the negative
    double* E;                     // N
    double* Es;                    // cap_bins x N
    double* D;                     // cap_bins x 5
    int64_t cap_bins;
    double* dpart;                 // 3 per chunk of 4096 particle indices
    int64_t nchunks;
    State12* st;
    const double2* tw;             // twiddle table (pass_twiddle_table)
    Tile12* tiles;
    int32_t ntiles, block;
    // mixed (precision 1): positions as 32-bit fractions of the box,
    // velocities as FP32 hi + lo pairs in steps of that grid per time step
    // (v * dt * 2^32), FP32 copy of E in the gather's units (E * dt^2 * 4),
    // per-species u64 charge quanta (2^-30 of a particle) and 6 exact 128-bit
    // diagnostics partials per chunk
    int32_t precision;
    uint32_t* xq;
    float *vxh, *vxl, *vyh, *vyl;
    float* E32;
    unsigned long long* racc32;
    void* dpart32;
    double vstep, inv_vstep, escale2;
    float kick_half[2];  // q_m / 2: the half kick per gathered step-unit field
    float t3[2];         // B0 * (dt / 2) * q_m
};

int push12_lds_bytes(int N);
void launch12_deposit(const Buffers12& b, hipStream_t s);  // the charge of the current positions
void launch12_solve(const Buffers12& b, hipStream_t s);    // E, binned E, sum(E.^2)
void launch12_push(const Buffers12& b, hipStream_t s);     // gather, Boris, move, next charge
void launch12_diag(const Buffers12& b, hipStream_t s);     // D row when t % bin == 0; step += 1
void launch12_rho(const Buffers12& b, double* rho, hipStream_t s);  // charge -> rho (clears it)
void launch12_op_gather(const Buffers12& b, const double* E, const double* x, int64_t n, double* Ej,
                        hipStream_t s);
void launch12_op_boris(double* vx, double* vy, const double* Ej, int64_t n, double B, double dt, double q_m,
                       hipStream_t s);
int configure_kernels_1d2v();
// kernels_mixed_aux.hip: the mixed forms of the deposit, push and diagnostics
void launch12_deposit32(const Buffers12& b, hipStream_t s);
void launch12_push32(const Buffers12& b, hipStream_t s);
void launch12_diag32(const Buffers12& b, hipStream_t s);

}  // namespace pic
