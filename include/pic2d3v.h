/*
 * pic2d3v.h -- C ABI of the structured 2D3V module of
 * jwscook/ParticleInCellCodeGolf.jl (src/PIC2D3V.jl) on AMD MI355X (gfx950):
 * the electrostatic path and the Lorenz-gauge electromagnetic path.
 *
 * Replaces, for an ElectrostaticField and any number of Species (up to
 * PIC2D3V_MAX_SPECIES):
 *   Species(xyv 5 x P, charge, mass, weight, shape)          src/PIC2D3V.jl:148-156
 *   sort!(species, dx, dy)        per-cell particle sort       :210-215
 *   depositindicesfractions + the shapes NGP, area, B-spline 0..5
 *                                                              :1106-1187
 *   (field)(shape, x, y)          field gather                 :1217-1228
 *   ElectrostaticBoris            Boris push, uniform B        :39-54
 *   deposit!(rho, shape, x, y, ..., q*w/dV)                    :1249-1256
 *   loop!(plasma, field)          particle loop, reduction, spectral solve,
 *                                 update!                      :530-581, :294-297
 *   diagnose!(ElectrostaticDiagnostics, ...)                   :1301-1330
 * and, for a LorenzGaugeStaggeredField (field = PIC2D3V_LORENZ_STAGGERED, as
 * src/2D3V.jl:102-106 builds it), a LorenzGaugeField (PIC2D3V_LORENZ, the
 * centred loop! of :665-745, run with pic2d3v_loop only: the reference's
 * diagnose! reads field.rho0, which that type lacks) or a
 * LorenzGaugeSemiImplicitField (PIC2D3V_LORENZ_SEMI_IMPLICIT, :403-461, its
 * iterated loop! described at pic2d3v_lg_grid below):
 *   warmup!(field, plasma)        advect -dt/2, +dt/2, the currents   :826-859
 *   lorenzgauge!(imex, ...)       per-mode wave update, any ImEx theta :610-635
 *   E = -grad phi+ - dA/dt, B = curl (A+ + A0)/2, ifft, update!     :917-935
 *   (field)(shape, x, y)          six-component gather             :1231-1247
 *   ElectromagneticBoris                                           :56-69
 *   the two half moves, deposit!(rho) at the half step and
 *   deposit!(Jx, Jy, Jz) at the new positions                      :937-968
 *   diagnose!(LorenzGaugeDiagnostics, ...)  with its in-place inverse
 *                                 transforms of the state          :1332-1398
 *
 * FP64 by default.  Results are bit-identical to the CPU restatement: per-cell
 * charge is an exact signed fixed-point sum (independent of the particle
 * order and of the per-thread grids of :284-285), every other sum uses the
 * canonical reduction tree.
 *
 * precision = PIC_PREC_MIXED (electrostatic field only) holds positions as
 * 32-bit fractions of the box and velocities as FP32 hi + lo pairs, gathers
 * an FP32 copy of the field with FP32 shape weights, and deposits each
 * species' charge as exact integer quanta (2^-22 of a cell per axis, summed
 * in 128 bits per cell), so its results too are independent of the particle
 * order.  The solve, the gathered field Exy and the diagnostics stay FP64;
 * uploads and downloads convert.  The operator calls below are FP64 in
 * either precision.
 *
 * Layouts follow Julia's.  A species' particles are the 5 x P column-major
 * matrix xyv: particle p is (x, y, vx, vy, vz) at xyv[5p .. 5p+4].  Grids are
 * column-major NX x NY (x fastest).  Positions live in (0, Lx] x (0, Ly]
 * (unimod, :9).  The ghost layers of the reference's offset arrays hold the
 * periodic images of the NX x NY field; they are not stored here, but
 * `buffer` enters the field-energy diagnostic, which averages over them.
 *
 * Two behaviours of the reference are kept by default, each with an opt-in
 * alternative:
 *   - update! adds every step's field into the gathered field Exy and never
 *     clears it (:294-297 with applyperiodicity!'s +=, :20-27): the particles
 *     see the running sum of all fields so far.  PIC2D3V_FIELD_REPLACE uses
 *     the step's own field.
 *   - ElectrostaticBoris scales the rotation by q_m^2 (:49-54), which turns
 *     negative charges the wrong way and keeps |v| only for q_m^2 = 1.
 *     PIC2D3V_BORIS_TEXTBOOK uses ElectromagneticBoris's form (:61-69).
 * The shapes keep the reference's cell mirror: r = ceil(zN) - zN is taken as
 * the distance from the lower cell edge, so a particle's charge centroid is
 * 2 ceil(zN) - zN.
 *
 * The Lorenz-gauge path keeps the reference's data flow exactly:
 *   - the wave update reads rho0 and J0 as they are: the real-space
 *     reduction of the previous step, or its inverse transform once
 *     diagnose! has run (the forward transform and smoothing of :891-903 are
 *     overwritten by the reduction before anything reads them);
 *   - diagnose! inverse-transforms Ax0, Ay0, Az0, phi0, rho0, J0 in place
 *     (:1359-1368), which changes the next step; pic2d3v_loop runs loop!
 *     alone;
 *   - warmup!'s second deposit! clears the first one's charge, so rho0
 *     starts at zero.
 * Current sums are exact like the charge, with the fraction bits set from the
 * largest |v| of each component after the push.
 *
 * Return codes, ownership, threading and error reporting follow pic.h.
 */
#ifndef PIC2D3V_H
#define PIC2D3V_H
#include <stdint.h>

#include "pic.h"

#ifdef __cplusplus
extern "C" {
#endif

#define PIC2D3V_MAX_SPECIES 4

/* shapes (src/PIC2D3V.jl:141-145) */
#define PIC2D3V_NGP 0
#define PIC2D3V_AREA 1
#define PIC2D3V_BSPLINE0 2 /* BSplineWeighting{n}: PIC2D3V_BSPLINE0 + n, n = 0..5 */
#define PIC2D3V_BSPLINE5 7

#define PIC2D3V_FIELD_ACCUMULATE 0 /* update! as the reference: Exy += E     */
#define PIC2D3V_FIELD_REPLACE 1    /* Exy = E                                */
#define PIC2D3V_BORIS_REFERENCE 0  /* ElectrostaticBoris (q_m^2 rotation)    */
#define PIC2D3V_BORIS_TEXTBOOK 1   /* ElectromagneticBoris with Ez = 0       */

#define PIC2D3V_ELECTROSTATIC 0    /* ElectrostaticField (:270-297)          */
#define PIC2D3V_LORENZ_STAGGERED 1 /* LorenzGaugeStaggeredField (:347-392)   */
#define PIC2D3V_LORENZ 2           /* LorenzGaugeField (:307-345)            */
#define PIC2D3V_LORENZ_SEMI_IMPLICIT 3 /* LorenzGaugeSemiImplicitField (:403-461) */

/* Lorenz-gauge field state: complex NX x NY arrays (pic2d3v_lg_get/set).
 * For PIC2D3V_LORENZ the slots hold that type's arrays: PHI0 = phi+ (the
 * latest), PHIM = phi0 (the one before), AX0.. = A+, AXM.. = A0,
 * RHO0.. = rho+, J+ (in the spectrum, as loop! leaves them). */
#define PIC2D3V_LG_PHI0 0
#define PIC2D3V_LG_PHIM 1  /* phi at the previous step */
#define PIC2D3V_LG_AX0 2
#define PIC2D3V_LG_AY0 3
#define PIC2D3V_LG_AZ0 4
#define PIC2D3V_LG_AXM 5
#define PIC2D3V_LG_AYM 6
#define PIC2D3V_LG_AZM 7
#define PIC2D3V_LG_RHO0 8
#define PIC2D3V_LG_JX0 9
#define PIC2D3V_LG_JY0 10
#define PIC2D3V_LG_JZ0 11
#define PIC2D3V_LG_EX 12
#define PIC2D3V_LG_EY 13
#define PIC2D3V_LG_EZ 14
#define PIC2D3V_LG_BX 15
#define PIC2D3V_LG_BY 16
#define PIC2D3V_LG_BZ 17
#define PIC2D3V_LG_NSTATE 18
/* PIC2D3V_LORENZ_SEMI_IMPLICIT has four more: the sources at n-1 (rho-, J-),
 * which its warmup! deposits at -dt and nothing changes after; its RHO0, J0
 * are the sources at n (warmup! at +dt), changed only by diagnose!. */
#define PIC2D3V_LG_RHOM 18
#define PIC2D3V_LG_JXM 19
#define PIC2D3V_LG_JYM 20
#define PIC2D3V_LG_JZM 21
#define PIC2D3V_LG_NSTATE_SI 22
/* binned fields of LorenzGaugeDiagnostics, in its order (pic2d3v_lg_bins) */
#define PIC2D3V_LG_BIN_EX 0
#define PIC2D3V_LG_BIN_EY 1
#define PIC2D3V_LG_BIN_EZ 2
#define PIC2D3V_LG_BIN_BX 3
#define PIC2D3V_LG_BIN_BY 4
#define PIC2D3V_LG_BIN_BZ 5
#define PIC2D3V_LG_BIN_AX 6
#define PIC2D3V_LG_BIN_AY 7
#define PIC2D3V_LG_BIN_AZ 8
#define PIC2D3V_LG_BIN_PHI 9
#define PIC2D3V_LG_BIN_RHO 10
#define PIC2D3V_LG_BIN_JX 11
#define PIC2D3V_LG_BIN_JY 12
#define PIC2D3V_LG_BIN_JZ 13
#define PIC2D3V_LG_NBINS 14

typedef struct pic2d3v_species {
    int64_t P;        /* particles                                          */
    int32_t shape;    /* PIC2D3V_NGP, PIC2D3V_AREA or PIC2D3V_BSPLINE0 + n    */
    double charge;    /* Species.charge                                     */
    double mass;      /* Species.mass                                       */
    double weight;    /* Species.weight (n0*Lx*Ly/P in the reference, :189) */
} pic2d3v_species;

typedef struct pic2d3v_config {
    int32_t NX, NY;        /* grid, powers of two in [8, 1024]                       */
    double Lx, Ly;         /* box                                                    */
    double dt;             /* time step                                              */
    double B0[3];          /* uniform magnetic field (B0x, B0y, B0z)                 */
    int32_t buffer;        /* ghost layers of Exy (3 in the reference)               */
    int32_t ntskip;        /* diagnostics every ntskip steps, fields binned by them  */
    int32_t ngskip;        /* field subsampling of the binned fields (power of two)  */
    int32_t field_update;  /* PIC2D3V_FIELD_ACCUMULATE or PIC2D3V_FIELD_REPLACE      */
    int32_t boris;         /* PIC2D3V_BORIS_REFERENCE or PIC2D3V_BORIS_TEXTBOOK      */
    int32_t nspecies;      /* 1 .. PIC2D3V_MAX_SPECIES                               */
    pic2d3v_species species[PIC2D3V_MAX_SPECIES];
    int32_t device;        /* HIP device ordinal                                     */
    int32_t field;         /* PIC2D3V_ELECTROSTATIC, _LORENZ_STAGGERED, _LORENZ, ... */
    double theta;          /* Lorenz gauge: ImEx theta (0 Explicit(), 1 Implicit())   */
    double theta_source;   /* semi-implicit: sourceimex theta                          */
    double rtol;           /* semi-implicit: rtol of the convergence test (sqrt(eps))  */
    int32_t maxiters;      /* semi-implicit: at most maxiters + 1 iterations (10)      */
    int32_t precision;     /* PIC_PREC_FP64, or PIC_PREC_MIXED (electrostatic field)  */
} pic2d3v_config;

typedef struct pic2d3v_ctx pic2d3v_ctx;

/* Context lifetime: a new context holds a zero field, step t = 0 and no
 * diagnostics (a fresh ElectrostaticField and ElectrostaticDiagnostics).
 * pic2d3v_last_error(NULL) returns the message of the last failed
 * pic2d3v_create on this thread. */
int pic2d3v_create(pic2d3v_ctx** ctx, const pic2d3v_config* cfg);
int pic2d3v_destroy(pic2d3v_ctx* ctx);
const char* pic2d3v_last_error(const pic2d3v_ctx* ctx);

/* Species.xyv of one species (5*P doubles, column-major 5 x P) in and out.
 * Uploading leaves the field, t and the diagnostics alone. */
int pic2d3v_upload(pic2d3v_ctx* ctx, int32_t species, const double* xyv);
int pic2d3v_download(pic2d3v_ctx* ctx, int32_t species, double* xyv);

/* Back to a fresh field and diagnostics: Exy (or the Lorenz-gauge state) = 0,
 * t = 0, no rows, no bins. */
int pic2d3v_reset(pic2d3v_ctx* ctx);

/* sort!(species, dx, dy): reorders the species' particles by
 * (ceil(x/dx), ceil(y/dy), vx) under isless, ties by position (the stable
 * sortperm).  perm (may be NULL) receives the 0-based permutation: new
 * particle k was particle perm[k] (the reference's species.p minus 1).
 * Fails with PIC_EINVAL if the cells span more than 2^24 or one cell holds
 * more than 16384 particles.  The context itself keeps its particles grouped
 * by 8 x 8-cell tile for the LDS windows of its deposits (before the first
 * step after an upload; then, checked every 16 steps, when more than 1% of
 * the stencils fell outside their windows, or always for the Lorenz-gauge
 * fields); that order is internal:
 * pic2d3v_download, the diagnostics and this sort see the caller's order,
 * and no result depends on it. */
int pic2d3v_sort(pic2d3v_ctx* ctx, int32_t species, double dx, double dy, int64_t* perm);

/* nsteps of `loop!(plasma, field); diagnose!(d, field, plasma, t)` from the
 * current t.  rows (may be NULL) receives one row of 8 doubles for every t of
 * this call with t % ntskip == 0: kinetic energy, field energy, momentum
 * (x, y, z), characteristic momentum (x, y, z), i.e. kineticenergy[ti],
 * fieldenergy[ti], particlemomentum[ti], characteristicmomentum[ti].  The
 * Lorenz-gauge path writes rows of 11: kinetic energy, field energy,
 * particle momentum (3), field momentum (3), characteristic momentum (3).
 * Returns PIC_ENONFINITE if the solve met a non-finite value (the reference's
 * `@assert isfinite(tmp)`, :570); the state is then that of the last step. */
int pic2d3v_run(pic2d3v_ctx* ctx, int64_t nsteps, double* rows);

/* Lorenz gauge only (a LorenzGaugeField's EBxyz, which its loop! gathers
 * before it solves, follows the fields set with pic2d3v_lg_set; a fresh one
 * is zero): warmup!(field, plasma) (:826-859), which the reference's
 * driver calls once before its loop (src/2D3V.jl:123); nsteps of loop!
 * without diagnose!; the field state slot `which` (PIC2D3V_LG_*: 2*NX*NY
 * doubles, complex interleaved as Julia's ComplexF64) out and in; and the
 * binned field `which` (PIC2D3V_LG_BIN_*) of bins 0 .. nbins-1, laid out as
 * pic2d3v_fields lays out Exs.  pic2d3v_warmup on an electrostatic context
 * does nothing, as warmup!(::AbstractField) (:745). */
int pic2d3v_warmup(pic2d3v_ctx* ctx);
int pic2d3v_loop(pic2d3v_ctx* ctx, int64_t nsteps);
int pic2d3v_lg_get(pic2d3v_ctx* ctx, int32_t which, double* z);
int pic2d3v_lg_set(pic2d3v_ctx* ctx, int32_t which, const double* z);
int pic2d3v_lg_bins(pic2d3v_ctx* ctx, int32_t which, double* out, int64_t nbins);

/* LorenzGaugeSemiImplicitField (field = PIC2D3V_LORENZ_SEMI_IMPLICIT,
 * :403-461; src/2D3V.jl:104-105 builds one with fieldimex = ImEx(1.0),
 * sourceimex = ImEx(0.05), buffer = 10): warmup! (:864-871) and loop!
 * (:993-1100) through pic2d3v_warmup, pic2d3v_run and pic2d3v_loop.  Each
 * step iterates from its start state: push (xw = unimod(x + vx*dt) with the
 * old v, vw = ElectromagneticBoris), the exact current deposits into the
 * buffered grid rhoJs+, its reduction and forward transform, lorenzgauge!
 * with the source theta_source/2*s- + (1 - theta_source)*s0 +
 * theta_source/2*s+ and the field theta, E and B with theta/2 (x+ + x-) +
 * (1 - theta) x0, their inverse transforms and update!; it stops after
 * iteration k when k > maxiters or isapprox(grid before, grid after; rtol).
 * The reference keeps one grid per thread and its test compares them all,
 * so its iteration counts depend on nthreads(); this is the one-thread grid
 * (julia's default).  buffer must cover every species' stencil; a stencil
 * cell that still falls outside (a position rounding onto NX + 1, the
 * reference's BoundsError) fails the call with PIC_EINVAL.
 * pic2d3v_lg_grid / pic2d3v_lg_set_grid: the grid rhoJs+ (4*(NX+2*buffer)*
 * (NY+2*buffer) doubles, the reference's element order: component fastest,
 * then x, then y; axes -(buffer-1):NX+buffer), which the next step's first
 * test compares with.  pic2d3v_lg_iterations: the iteration counts of the
 * last n steps of the most recent pic2d3v_run or pic2d3v_loop call. */
int pic2d3v_lg_grid(pic2d3v_ctx* ctx, double* out);
int pic2d3v_lg_set_grid(pic2d3v_ctx* ctx, const double* in);
int pic2d3v_lg_iterations(pic2d3v_ctx* ctx, int64_t* out, int64_t n);

/* The time-binned fields of the diagnostics, Exs, Eys, phis[:, :, b] for
 * b < nbins, each (NX/ngskip) x (NY/ngskip) per bin, bins back to back
 * (Julia's Array{Float64,3} layout).  nbins may not exceed ceil(t/ntskip).
 * Any output may be NULL. */
int pic2d3v_fields(pic2d3v_ctx* ctx, double* Exs, double* Eys, double* phis, int64_t nbins);

/* The gathered field Exy: the interior NX x NY of each component (either
 * may be NULL); for the Lorenz gauge, EBxyz's Ex and Ey. */
int pic2d3v_field(pic2d3v_ctx* ctx, double* Ex, double* Ey);

/* Operators on host arrays, bit-identical to the CPU restatement:
 * deposit!: rho (NX x NY) = sum over n particles of (wx*wy)*w with the shape,
 * folded periodically (:1249-1256, :481-486); the gather (f)(shape, x, y) of
 * field (Ex, Ey) at n positions (:1217-1228); the context's Boris (variant,
 * B0, dt) on n particles in place with charge-to-mass ratio q_m (:39-54); and
 * the solve of rho in the context's box (:559-576), phi (may be NULL)
 * receiving real(ifft(phi)) as diagnose! stores it (:1326). */
int pic2d3v_deposit(pic2d3v_ctx* ctx, int32_t shape, const double* x, const double* y, int64_t n, double w,
                    double* rho);
int pic2d3v_gather(pic2d3v_ctx* ctx, int32_t shape, const double* Ex, const double* Ey, const double* x,
                   const double* y, int64_t n, double* ex, double* ey);
int pic2d3v_boris(pic2d3v_ctx* ctx, double* vx, double* vy, double* vz, const double* ex, const double* ey,
                  int64_t n, double q_m);
int pic2d3v_fieldsolve(pic2d3v_ctx* ctx, const double* rho, double* Ex, double* Ey, double* phi);

#ifdef __cplusplus
}
#endif
#endif
