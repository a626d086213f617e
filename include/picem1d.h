/*
 * picem1d.h -- C ABI of the 1D electromagnetic PIC steps of
 * jwscook/ParticleInCellCodeGolf.jl on AMD MI355X (gfx950): the fixed point
 * of src/GaussianEMFixedPoint.jl and the explicit step of
 * src/GaussianEMExplicit.jl.
 *
 * Replaces the fixed-point script's step loop and its closures:
 *   s(c, u), d(y, u)      Gaussian +-6 stencil centred on the wrapped cell
 *                         irn(c) = Int(mod1(round(c*N), N)); deposit      :8-11
 *   Ex = real(ifft(fft(d)./ik))      Gauss                                :20
 *   Ey, Ez = F + dt*rif(-+ik.*fft(b)/eps^2 .- fft(J))   Ampere            :22-23
 *   By, Bz = C - dt*rif(-+ik.*fft(e))                   Faraday           :25-26
 *   v = u + (a + cross(u + a*dt/2, b))*dt               velocity update   :28-29
 *   up to three sweeps while !isapprox(Fx, Ex, rtol)                      :14
 *   x = mod(x, 1); p = sortperm(x)                                        :41
 *   fields[:, t, 1:12]    Ex, Ey, Ez, By, Bz and the moments d(x, v^k)    :45-49
 *
 * and the explicit script's (src/GaussianEMExplicit.jl, same closures :7-11):
 *   X .= x; Ex = Gauss(d(X))                                            :13
 *   x = mod(X + v[:,1]*dt, 1); p = sortperm(x); x, v = x[p], v[p, :]      :14
 *   By, Bz .-= dt*rif(-+ik.*fft(Ez, Ey)); b = (B + C)/2   Faraday      :15-16
 *   v += a*dt/2; v += cross(v, b)*dt; v += a*dt/2 at s(x)  split kick  :17-22
 *   Ey, Ez .+= dt*rif(-+ik.*fft(Bz, By)/eps^2 .- fft(d((X.+x)/2, v)))  :23-24
 *   fields[:, t, 1:12]                                                  :25-29
 *
 * FP64 only; every result is bit-identical to the CPU restatement (charge,
 * current and moment sums are exact signed fixed-point sums; every other sum
 * is in the script's order).  The script's quirks are kept: the stencil's
 * centre is wrapped while the position is not (a particle with c*N < 1/2, or c
 * past 1 during a sweep, has zero weight); `Fz .*= Ez` leaves Fz zero, so Ez
 * restarts from the current term every step; the fixed point tests Ex only
 * and averages Ex against the previous sweep's.  The explicit script sorts x
 * and v but not X, so (X .+ x)/2 pairs one particle's old position with
 * another's new one; that is kept too.  Bx is the scripts' zero array.
 * Return codes, ownership, threading and error reporting follow pic.h.
 */
#ifndef PICEM1D_H
#define PICEM1D_H
#include <stdint.h>

#include "pic.h"

#ifdef __cplusplus
extern "C" {
#endif

#define PICEM1D_FIXED_POINT 0 /* src/GaussianEMFixedPoint.jl */
#define PICEM1D_EXPLICIT 1    /* src/GaussianEMExplicit.jl   */

typedef struct picem1d_config {
    int32_t N;           /* cells, a power of two in [8, 1024]                       */
    int32_t max_sweeps;  /* fixed-point sweeps per step (3: `for _∈0:2`); explicit: unused */
    int64_t P;           /* particles, at most 2^26                                  */
    double eps;          /* 1/c (the script's ϵ = 1/10)                              */
    double dt;           /* time step (ϵ^2/N/2)                                      */
    double w;            /* particle weight (W/P*N)                                  */
    double rtol;         /* isapprox tolerance of the fixed point (η = 1e-8); explicit: unused */
    int32_t sort;        /* 1: p = sortperm(x) after every step, as the script; 0: never */
    int32_t device;      /* HIP device ordinal                                       */
    int32_t scheme;      /* PICEM1D_FIXED_POINT or PICEM1D_EXPLICIT                  */
} picem1d_config;

typedef struct picem1d_ctx picem1d_ctx;

int picem1d_create(picem1d_ctx** ctx, const picem1d_config* cfg);
int picem1d_destroy(picem1d_ctx* ctx);
const char* picem1d_last_error(const picem1d_ctx* ctx);

/* State in: x (P), v (3*P: the columns v[:,1], v[:,2], v[:,3] back to back),
 * E (3*N: Ex, Ey, Ez; NULL = zeros) and B (2*N: By, Bz; NULL = the script's
 * zeros and ones).  Restarts the step count and clears the script's Fz. */
int picem1d_upload(picem1d_ctx* ctx, const double* x, const double* v, const double* E, const double* B);
/* State out (any NULL). */
int picem1d_download(picem1d_ctx* ctx, double* x, double* v, double* E, double* B);

/* nsteps of the loop.  fields (may be NULL) receives fields[:, t, 1:12] for
 * the steps of this call as an N x nsteps x 12 column-major array (Julia's
 * layout); sweeps (may be NULL) the sweeps each step ran (1 for the explicit
 * scheme). */
int picem1d_run(picem1d_ctx* ctx, int64_t nsteps, double* fields, int32_t* sweeps);

/* Operators (bitwise equal to the CPU restatement):
 * d(y, u): rho (N) = sum of (u[j]*ff)*w over the stencils at y (u NULL: 1);
 * one sweep's field updates (:20-27) from rho, Jy, Jz and the saved Fy, Fz,
 *   Cy, Cz, with E (3N) and B (2N: the iterate's By, Bz) updated in place and
 *   e3 (3N: ex, ey, ez, ex from the given Fx) and b2 (2N: by, bz) returned;
 * the velocity update of n particles at midpoints c with the fields
 *   ex, ey, ez, by, bz (N each): v (3n) = u + (a + cross(u + a*dt/2, b))*dt. */
int picem1d_deposit(picem1d_ctx* ctx, const double* y, const double* u, int64_t n, double w, double* rho);
int picem1d_fields(picem1d_ctx* ctx, const double* rho, const double* Jy, const double* Jz, const double* Fx,
                   const double* Fy, const double* Fz, const double* Cy, const double* Cz, double* E, double* B,
                   double* e3, double* b2);
int picem1d_push(picem1d_ctx* ctx, const double* c, const double* e3, const double* b2, const double* u,
                 int64_t n, double* v);

/* Operators of the explicit scheme (any context; bitwise equal to the CPU
 * restatement):
 * the step's field updates from its rho and currents Jy, Jz (N each): Ex =
 *   Gauss(rho) (:13), By, Bz from the given Ey, Ez (:15), b2 (2N: by, bz) =
 *   (B + old B)/2 (:16), then Ey, Ez from the new By, Bz and Jy, Jz (:23-24);
 *   E (3N) and B (2N) are updated in place;
 * the split-kick velocity update (:17-22) of n particles at positions x with
 *   the fields E (3N: Ex, Ey, Ez) and b2 (2N: by, bz): v (3n) from u (3n). */
int picem1d_explicit_fields(picem1d_ctx* ctx, const double* rho, const double* Jy, const double* Jz, double* E,
                            double* B, double* b2);
int picem1d_explicit_push(picem1d_ctx* ctx, const double* x, const double* E, const double* b2, const double* u,
                          int64_t n, double* v);

#ifdef __cplusplus
}
#endif
#endif
