/*
 * pic2d.h -- C ABI of the electrostatic 2D3V step on AMD MI355X (gfx950).
 *
 * Replaces the particle loop, charge reduction and field solve of
 * src/Electrostatic2D3V.jl (jwscook/ParticleInCellCodeGolf.jl):
 *   g(z, NZ)            area weights              src/Electrostatic2D3V.jl:84-92
 *   eval(F1, F2, x, y)  field gather              :94-103
 *   boris(...)          Boris push, B = B0 x      :32-41
 *   depositcharge!      area deposit              :105-109
 *   particle loop       gather, Boris, move, deposit   :125-138
 *   phi .= sum(ns, 3)   per-thread grid reduction :139-141 (exact per-cell sums here)
 *   field invert/solve  2D spectral Poisson solve :70-81,142-157
 *   K[ti, 1:5]          diagnostics every NS steps :163-170
 *
 * Grids are column-major NX x NY doubles (x fastest), Julia's layout: value
 * [i, j] (1-based) is element (i-1) + NX*(j-1).  Particle positions live in
 * (0, 1] (the script's unimod convention).  Return codes, ownership, threading
 * and error reporting follow pic.h: every call returns PIC_OK or a negative
 * PIC_E* code, copies host arrays in and out and keeps no host pointer, and a
 * context is driven by one thread at a time.
 */
#ifndef PIC2D_H
#define PIC2D_H
#include <stdint.h>

#include "pic.h"

#ifdef __cplusplus
extern "C" {
#endif

#define PIC2D_PAIR_REFERENCE 0 /* g(): cell i with i+1 (src/Electrostatic2D3V.jl:89) */
#define PIC2D_PAIR_TEXTBOOK 1  /* cloud in cell: cell i with i-1                      */

typedef struct pic2d_config {
    int32_t NX, NY;     /* grid, powers of two in [8, 1024]                            */
    int64_t P;          /* particles (FP64: <= 2^22, mixed: <= 2^24)                   */
    int32_t precision;  /* PIC_PREC_FP64, or PIC_PREC_MIXED: FP32 push with
                           32-bit fixed-point positions and exact integer area weights
                           (2^-20 of a cell per axis); FP64 solve and diagnostics     */
    int32_t pairing;    /* PIC2D_PAIR_REFERENCE or PIC2D_PAIR_TEXTBOOK                 */
    double dt;          /* time step                                                   */
    double w;           /* particle weight, n0/P/(dx*dy) in the script (:25)           */
    double B0;          /* magnetic field along x (:24)                                */
    int32_t ns;         /* diagnostics every ns steps (the script's NS = 2, :25)       */
    int32_t device;     /* HIP device ordinal                                          */
    int32_t sort_every; /* steps between regroupings of the particles by 16x16 cell
                           tile (0 = 8); results do not depend on it               */
} pic2d_config;

typedef struct pic2d_ctx pic2d_ctx;

/* Context lifetime.  pic2d_last_error(NULL) returns the message of the last
 * failed pic2d_create on this thread. */
int pic2d_create(pic2d_ctx** ctx, const pic2d_config* cfg);
int pic2d_destroy(pic2d_ctx* ctx);
const char* pic2d_last_error(const pic2d_ctx* ctx);

/* Device-resident state: particles (P each; x, y in (0, 1]) and the field the
 * next step gathers from (Ex, Ey: NX*NY each, NULL = zeros as in the script's
 * start, :68-69).  Upload resets the step counter to 0. */
int pic2d_upload(pic2d_ctx* ctx, const double* x, const double* y, const double* vx,
                 const double* vy, const double* vz, const double* Ex, const double* Ey);
/* Any output may be NULL. */
int pic2d_download(pic2d_ctx* ctx, double* x, double* y, double* vx, double* vy, double* vz,
                   double* Ex, double* Ey);

/* nsteps of the time loop (:120-176).  After every step t (counted from the
 * upload) with t % ns == 0 one row of K (5 doubles: mean(Ex^2+Ey^2),
 * sum((vx^2+vy^2)*w), their sum, sum(vx)/P, sum(vy)/P) is appended; K may be
 * NULL and must otherwise hold (t1/ns - t0/ns) rows, t0 and t1 the step count
 * before and after the call.  Returns PIC_ENONFINITE if the field became
 * non-finite (the script's `@assert isfinite(tmp)`, :148). */
int pic2d_run(pic2d_ctx* ctx, int64_t nsteps, double* K);

/* Operators on host arrays (FP64; the context supplies NX, NY, the pairing,
 * the device and the stream). */
/* depositcharge! for P particles: rho = sum of wx*wy*w (:105-109, :139-141). */
int pic2d_deposit(pic2d_ctx* ctx, const double* x, const double* y, int64_t P, double w,
                  double* rho);
/* Ex, Ey = real(ifft(fft(rho)*minvkk*k)) with phi[1,1] = 0 (:142-157); phi
 * (may be NULL) receives real(ifft(phi)) as stored in `phis` (:173). */
int pic2d_fieldsolve(pic2d_ctx* ctx, const double* rho, double* Ex, double* Ey, double* phi);
/* eval: ex[p], ey[p] = field at (x[p], y[p]) (:94-103). */
int pic2d_gather(pic2d_ctx* ctx, const double* Ex, const double* Ey, const double* x,
                 const double* y, int64_t P, double* ex, double* ey);
/* boris in place on vx, vy, vz with per-particle ex, ey (:32-41). */
int pic2d_boris(pic2d_ctx* ctx, double* vx, double* vy, double* vz, const double* ex,
                const double* ey, int64_t P, double dt, double B0);
/* The particle loop without the deposit, in place: gather, Boris,
 * x = unimod(x + vx*dt, 1), y = unimod(y + vy*dt, 1) (:128-134). */
int pic2d_push(pic2d_ctx* ctx, double* x, double* y, double* vx, double* vy, double* vz,
               int64_t P, const double* Ex, const double* Ey, double dt, double B0);
/* K[0..4] of the state (:166-170). */
int pic2d_diagnostics(pic2d_ctx* ctx, const double* Ex, const double* Ey, const double* vx,
                      const double* vy, int64_t P, double w, double* K);

#ifdef __cplusplus
}
#endif
#endif
