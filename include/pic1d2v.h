/*
 * pic1d2v.h -- C ABI of the magnetised 1D2V step on AMD MI355X (gfx950).
 *
 * Replaces the step loop of the Bernstein-wave scripts of
 * jwscook/ParticleInCellCodeGolf.jl, one species (src/NGP1D2V.jl) or two
 * (src/NGP1D2V2S.jl):
 *   rho() / rho(x)        Gaussian +-7 deposit, summed over species with their
 *                         charges                 src/NGP1D2V2S.jl:28-29, src/NGP1D2V.jl:38
 *   E = real(ifft(...))   spectral solve          src/NGP1D2V2S.jl:33, src/NGP1D2V.jl:40
 *   Ej = sum(...d(x[j]))  gather at x             src/NGP1D2V2S.jl:35, src/NGP1D2V.jl:42
 *   boris(...)            Boris push, B = B0 z    src/NGP1D2V2S.jl:5-11, src/NGP1D2V.jl:5-10
 *   x += vx*dt; mod(x,1)  move and wrap           src/NGP1D2V2S.jl:37,46-47, src/NGP1D2V.jl:44,55
 *   Es[:, ti] .+= E       time-binned field for omega-k spectra   :49, :57
 *   D[ti, 1:5]            diagnostics every bin steps             :52-56, :58-62
 *
 * FP64: every result is bit-identical to the CPU restatement (per-cell
 * charge sums are exact signed 128-bit integer sums, every other sum uses the
 * canonical tree).  MIXED (precision = PIC_PREC_MIXED) runs the loop with
 * fixed-point positions, FP32 hi + lo velocities, FP32 Gaussian weights over
 * +-5 cells and exact integer charge quanta per species, with the solve and
 * the diagnostics in FP64; the operators stay FP64.  Return codes,
 * ownership, threading and error reporting follow pic.h.
 */
#ifndef PIC1D2V_H
#define PIC1D2V_H
#include <stdint.h>

#include "pic.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pic1d2v_config {
    int32_t N;         /* cells, a power of two in [8, 4096]                          */
    int32_t shape;     /* PIC_SHAPE_GAUSS7 (the scripts) or PIC_SHAPE_GAUSS6          */
    int32_t nspecies;  /* 1 (src/NGP1D2V.jl) or 2 (src/NGP1D2V2S.jl)                  */
    int32_t bin;       /* steps per time bin of Es and per diagnostics row (T/TO)     */
    int64_t P;         /* particles per species                                      */
    double q[2];       /* charge per species: (1) or (-1, 1)                          */
    double q_m[2];     /* charge-to-mass ratio per species: (1) or (-1, 1/M)          */
    double mass[2];    /* species weight in the kinetic sums of D: (1) or (1, M)      */
    double w;          /* particle weight: n0/P, or n0/2P                             */
    double n0;         /* the scripts' n0, which normalises D                         */
    double B0;         /* magnetic field along z                                      */
    double dt;         /* time step                                                   */
    int32_t device;    /* HIP device ordinal                                          */
    int32_t precision; /* PIC_PREC_FP64 (0, the default) or PIC_PREC_MIXED            */
} pic1d2v_config;

typedef struct pic1d2v_ctx pic1d2v_ctx;

/* Context lifetime.  pic1d2v_last_error(NULL) returns the message of the last
 * failed pic1d2v_create on this thread. */
int pic1d2v_create(pic1d2v_ctx** ctx, const pic1d2v_config* cfg);
int pic1d2v_destroy(pic1d2v_ctx* ctx);
const char* pic1d2v_last_error(const pic1d2v_ctx* ctx);

/* Device-resident particles of one species (P each).  An upload restarts the
 * step count t at 0 and clears the binned field and the diagnostics; upload
 * every species before running. */
int pic1d2v_upload(pic1d2v_ctx* ctx, int32_t species, const double* x, const double* vx, const double* vy);
int pic1d2v_download(pic1d2v_ctx* ctx, int32_t species, double* x, double* vx, double* vy);

/* nsteps steps of the loop.  D (may be NULL) receives one row of 5 for every
 * t in this call with t % bin == 0: the scripts' D[t/bin, 1:5] as computed in
 * the loop (they divide D by bin after it). */
int pic1d2v_run(pic1d2v_ctx* ctx, int64_t nsteps, double* D);

/* The binned field: Es[b*N + i] = sum of E[i] over the steps t with
 * ceil(t/bin) == b+1, for the first nbins bins (the scripts divide by bin).
 * nbins may not exceed ceil(t/bin). */
int pic1d2v_fields(pic1d2v_ctx* ctx, double* Es, int64_t nbins);

/* Operators (bitwise equal to the CPU restatement):
 * rho(): the deposit of ns*P positions (species back to back) with the charges
 * of the context (src/NGP1D2V2S.jl:28-29); the gather sum(k->E[k[1]]*k[2], d(x[j]))
 * at n positions; and boris(vx, vy, E, B, dt, q_m) of src/NGP1D2V2S.jl:5-11
 * on n particles in place. */
int pic1d2v_deposit(pic1d2v_ctx* ctx, const double* x, double* rho);
int pic1d2v_gather(pic1d2v_ctx* ctx, const double* E, const double* x, int64_t n, double* Ej);
int pic1d2v_boris(pic1d2v_ctx* ctx, double* vx, double* vy, const double* Ej, int64_t n, double B, double dt,
                  double q_m);

#ifdef __cplusplus
}
#endif
#endif
