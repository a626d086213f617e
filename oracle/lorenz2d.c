/*
 * lorenz2d.c -- CPU FP64 restatement of the Lorenz-gauge electromagnetic path
 * of the structured 2D3V module src/PIC2D3V.jl as src/2D3V.jl:102-106,123-126
 * drives it: LorenzGaugeStaggeredField (:347-392) with any ImEx theta,
 * warmup! (:826-859), loop! (:875-979), ElectromagneticBoris (:56-69), the
 * six-component gather (:1231-1247), the one- and three-weight deposits
 * (:1258-1284), update! (:468-479), reduction! (:504-522) and
 * diagnose!(::LorenzGaugeDiagnostics) (:1332-1398), whose in-place inverse
 * transforms of the potentials and sources are part of the dynamics; and
 * LorenzGaugeField's centred loop! (:307-345, :665-745) and
 * LorenzGaugeSemiImplicitField (:403-461, :864-871, :993-1100) at the end of
 * the file.
 *
 * TEST INFRASTRUCTURE ONLY (see picoracle.h): the oracle the GPU Lorenz-gauge
 * path is checked against.
 *
 * Field state: 18 complex NX x NY arrays (column-major, interleaved re, im as
 * Julia's ComplexF64), in the order of ORCLG_* in picoracle.h: phi0, phi-,
 * Ax0, Ay0, Az0, Ax-, Ay-, Az-, rho0, Jx0, Jy0, Jz0, Ex, Ey, Ez, Bx, By, Bz.
 * phi+ and A+ are temporaries: loop! writes them in full before it reads
 * them, so the forward transforms diagnose! applies to them (:1388-1391) are
 * not observable and are not restated; nor is the forward transform and
 * smoothing of rho0, J0 at :891-903, which the reduction at :964-966
 * overwrites before anything reads it.
 *
 * What loop! does, in the reference's order:
 *   1. lorenzgauge! (:629-635) for phi (source rho0) and Ax, Ay, Az (sources
 *      J0): x+ = (num*x0 + dt^2*s)/den - x-, num = 2 - dt^2*k^2*(1 - theta),
 *      den = 1 + dt^2*k^2*theta/2 (:610-615; Explicit is theta 0 and
 *      Implicit theta 1 bit for bit).  The sources are whatever rho0, J0
 *      hold: the real-space reduction of the previous step, or its inverse
 *      transform when diagnose! ran.
 *   2. E = -im*k*phi+ - (A+ - A0)/dt, B = curl((A+ + A0)/2) in the spectrum
 *      (:917-925), then ifft of the six (:927-934) and update! (:935):
 *      EBxyz = (0.0 + real(F)) (+ B0 for B), ghost layers periodic copies.
 *   3. the particle loop (:937-968): gather, Boris, half move, deposit rho,
 *      half move, deposit J; reduction into rho0, J0 (imaginary parts 0).
 *   4. phi- = phi0, phi0 = phi+, A- = A0, A0 = A+ (:969-978).
 *
 * Charge and current sums are exact (each contribution rounded to 2^-q,
 * summed exactly, rounded once): q from sum_s P_s |qw_s| for rho and
 * sum_s P_s (|qw_s| max|v_c|) for each current component, with the maxima of
 * the velocities the deposit uses.  The reference's `@muladd z += wxy*w` is
 * taken as the rounded product wxy*w added exactly.  Diagnostic sums over
 * particles and cells use the canonical tree (orc_tree_sum).
 *
 * Pinned by (tests/test_oracle_lorenz2d.py):
 *   orclg_wavenumbers  the FFTHelper closed form (kx = 2pi/Lx*[0:N/2-1, -N/2:-1]).
 *   orclg_boris        E = 0: |v| kept, turn angle 2 atan(|t|) about B; B = 0:
 *                      v += q_m E dt; E x B drift of the average velocity.
 *   orclg_gather       constant fields, linear reproduction of the B-splines.
 *   orclg_solve        single-mode closed forms of E = -grad phi - dA/dt and
 *                      B = curl A in an Lx x Ly box against numpy's FFT; the
 *                      per-mode recurrence of the vacuum wave (determinant 1,
 *                      |lambda| = 1 for theta >= 1/2) over many steps.
 *   orclg_warmup       rho0 stays zero, J0 is the deposit of the round-trip
 *                      advected particles.
 *   orclg_run          a numpy restatement of loop! and diagnose! (FFT in
 *                      numpy, deposits in floating point) over a few steps;
 *                      diagnose!'s in-place transforms; the binned averages
 *                      with ngskip = 2; two calls equal one.
 *   orclg_run_centred  (LorenzGaugeField) a numpy restatement of loop! over a
 *                      few steps; the vacuum recurrence; the centred E of a
 *                      single mode; uncharged particles stream freely in the
 *                      first step (EBxyz is still zero) and turn about B0 in
 *                      the second.
 * (tests/test_oracle_lorenz2d_si.py, LorenzGaugeSemiImplicitField):
 *   orclg_si_deposit   exact rational sums of the contributions on the 2^-q
 *                      grid per buffered cell and per periodic cell, every
 *                      shape, ghost cells included; the BoundsError.
 *   orclg_si_warmup    its three passes rebuilt from the pinned deposits.
 *   orclg_si_solve     single-mode closed forms of the theta-weighted E and B
 *                      and of the three source weights.
 *   orclg_si_run       a numpy restatement (floating-point deposits, numpy's
 *                      FFT, isapprox through numpy norms) with equal
 *                      iteration counts; the stopping rule; free streaming
 *                      before the first update!; diagnose!'s transforms of
 *                      rho0, J0; split calls.
 */
#include "picoracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef __int128 i128;
typedef unsigned __int128 u128;

static const double kTwoPiLG = 6.283185307179586; /* Julia 2pi = 2*Float64(pi) */

static inline int64_t wrap_cell(int64_t c, int32_t N) { return (c - 1) & (int64_t)(N - 1); }

/* FFTHelper (:251-267): kx = 2pi/Lx * vcat(0:NX/2-1, -NX/2:-1), ky likewise,
 * k2 = kx.^2 .+ ky'.^2. */
void orclg_wavenumbers(int32_t NX, int32_t NY, double Lx, double Ly, double* kx, double* ky, double* k2) {
    const double sx = kTwoPiLG / Lx, sy = kTwoPiLG / Ly;
    for (int32_t i = 0; i < NX; ++i) kx[i] = sx * (double)(i < NX / 2 ? i : i - NX);
    for (int32_t j = 0; j < NY; ++j) ky[j] = sy * (double)(j < NY / 2 ? j : j - NY);
    for (int32_t j = 0; j < NY; ++j)
        for (int32_t i = 0; i < NX; ++i) k2[i + (int64_t)NX * j] = kx[i] * kx[i] + ky[j] * ky[j];
}

/* ElectromagneticBoris (:56-69): theta = (dt/2)*q_m, t = B*theta,
 * s = 2/(1 + dot(t, t)), e = E*theta, m = v + e,
 * v+ = m + cross(m + cross(m, t), t)*s, v = v+ + e. */
void orclg_boris(double dt, double q_m, const double* E, const double* B, double* v) {
    const double th = dt / 2.0 * q_m;
    const double t1 = B[0] * th, t2 = B[1] * th, t3 = B[2] * th;
    const double sc = 2.0 / (1.0 + (t1 * t1 + t2 * t2 + t3 * t3));
    const double e1 = E[0] * th, e2 = E[1] * th, e3 = E[2] * th;
    const double m1 = v[0] + e1, m2 = v[1] + e2, m3 = v[2] + e3;
    const double c1 = m2 * t3 - m3 * t2, c2 = m3 * t1 - m1 * t3, c3 = m1 * t2 - m2 * t1;
    const double u1 = m1 + c1, u2 = m2 + c2, u3 = m3 + c3;
    const double d1 = u2 * t3 - u3 * t2, d2 = u3 * t1 - u1 * t3, d3 = u1 * t2 - u2 * t1;
    v[0] = (m1 + d1 * sc) + e1;
    v[1] = (m2 + d2 * sc) + e2;
    v[2] = (m3 + d3 * sc) + e3;
}

/* (f::AbstractLorenzGaugeField)(shape, x, y) (:1231-1247): j outer, i inner,
 * wxy = wx*wy, F = muladd(EBxyz[c, i, j], wxy, F) taken as a fused
 * multiply-add.  EB holds the six NX x NY components back to back. */
void orclg_gather(int32_t NX, int32_t NY, double NX_Lx, double NY_Ly, int32_t shape, const double* EB, double x,
                  double y, double* out) {
    int64_t cx[6], cy[6];
    double wx[6], wy[6];
    const int nx = orc23_weights(shape, x, NX_Lx, cx, wx);
    const int ny = orc23_weights(shape, y, NY_Ly, cy, wy);
    const int64_t G = (int64_t)NX * NY;
    double a[6] = {0, 0, 0, 0, 0, 0};
    for (int jj = 0; jj < ny; ++jj)
        for (int ii = 0; ii < nx; ++ii) {
            const double wxy = wx[ii] * wy[jj];
            const int64_t k = wrap_cell(cx[ii], NX) + (int64_t)NX * wrap_cell(cy[jj], NY);
            for (int c = 0; c < 6; ++c) a[c] = fma(EB[c * G + k], wxy, a[c]);
        }
    for (int c = 0; c < 6; ++c) out[c] = a[c];
}

/* ifft! in place of one complex array (FFTW's plan_ifft!: the unnormalised
 * inverse, then every element times 1/(NX*NY)). */
static void ifft2_inplace(int32_t NX, int32_t NY, double* z) {
    const int64_t G = (int64_t)NX * NY;
    double* re = (double*)malloc(sizeof(double) * (size_t)G);
    double* im = (double*)malloc(sizeof(double) * (size_t)G);
    for (int64_t k = 0; k < G; ++k) {
        re[k] = z[2 * k];
        im[k] = z[2 * k + 1];
    }
    orc2d_fft(NX, NY, re, im, 1);
    const double s = 1.0 / (double)G;
    for (int64_t k = 0; k < G; ++k) {
        z[2 * k] = re[k] * s;
        z[2 * k + 1] = im[k] * s;
    }
    free(re);
    free(im);
}

void orclg_ifft2(int32_t NX, int32_t NY, double* z) { ifft2_inplace(NX, NY, z); }

static inline void cmul(double ar, double ai, double br, double bi, double* cr, double* ci) {
    *cr = ar * br - ai * bi;
    *ci = ar * bi + ai * br;
}

/* Steps 1 and 2 of loop!: plus (4 complex arrays: phi+, Ax+, Ay+, Az+) from
 * the state, then Ex..Bz (state slots 12..17) in the spectrum, then their
 * inverse transforms.  -im*k is (0*k, -k) and im*k is (0*k, k), as Julia
 * multiplies Complex{Int}(0,-1) and Complex{Bool}(false,true) by a Float64. */
void orclg_solve(int32_t NX, int32_t NY, double Lx, double Ly, double dt, double theta, double* F, double* plus) {
    const int64_t G = (int64_t)NX * NY;
    double* kx = (double*)malloc(sizeof(double) * (size_t)NX);
    double* ky = (double*)malloc(sizeof(double) * (size_t)NY);
    double* k2 = (double*)malloc(sizeof(double) * (size_t)G);
    orclg_wavenumbers(NX, NY, Lx, Ly, kx, ky, k2);
    const double dt2 = dt * dt;
    double* Z[18];
    for (int a = 0; a < 18; ++a) Z[a] = F + (size_t)a * 2 * G;
    double* Pp[4];
    for (int a = 0; a < 4; ++a) Pp[a] = plus + (size_t)a * 2 * G;
    static const int x0s[4] = {ORCLG_PHI0, ORCLG_AX0, ORCLG_AY0, ORCLG_AZ0};
    static const int xms[4] = {ORCLG_PHIM, ORCLG_AXM, ORCLG_AYM, ORCLG_AZM};
    static const int srcs[4] = {ORCLG_RHO0, ORCLG_JX0, ORCLG_JY0, ORCLG_JZ0};
    for (int32_t j = 0; j < NY; ++j)
        for (int32_t i = 0; i < NX; ++i) {
            const int64_t k = i + (int64_t)NX * j;
            const double num = 2.0 - dt2 * k2[k] * (1.0 - theta);
            const double den = 1.0 + dt2 * k2[k] * theta / 2.0;
            for (int a = 0; a < 4; ++a) {
                const double* x0 = Z[x0s[a]] + 2 * k;
                const double* xm = Z[xms[a]] + 2 * k;
                const double* sr = Z[srcs[a]] + 2 * k;
                Pp[a][2 * k] = (num * x0[0] + dt2 * sr[0]) / den - xm[0];
                Pp[a][2 * k + 1] = (num * x0[1] + dt2 * sr[1]) / den - xm[1];
            }
            const double kxi = kx[i], kyj = ky[j];
            const double* ph = Pp[0] + 2 * k;
            const double *axp = Pp[1] + 2 * k, *ayp = Pp[2] + 2 * k, *azp = Pp[3] + 2 * k;
            const double *ax0 = Z[ORCLG_AX0] + 2 * k, *ay0 = Z[ORCLG_AY0] + 2 * k, *az0 = Z[ORCLG_AZ0] + 2 * k;
            double er, ei;
            /* Ex = -im*kx*phi+ ; Ex -= (Ax+ - Ax0)/dt */
            cmul(0.0 * kxi, -kxi, ph[0], ph[1], &er, &ei);
            Z[ORCLG_EX][2 * k] = er - (axp[0] - ax0[0]) / dt;
            Z[ORCLG_EX][2 * k + 1] = ei - (axp[1] - ax0[1]) / dt;
            cmul(0.0 * kyj, -kyj, ph[0], ph[1], &er, &ei);
            Z[ORCLG_EY][2 * k] = er - (ayp[0] - ay0[0]) / dt;
            Z[ORCLG_EY][2 * k + 1] = ei - (ayp[1] - ay0[1]) / dt;
            /* Ez = -(Az+ - Az0)/dt */
            Z[ORCLG_EZ][2 * k] = -(azp[0] - az0[0]) / dt;
            Z[ORCLG_EZ][2 * k + 1] = -(azp[1] - az0[1]) / dt;
            /* Bx = im*ky*(Az+ + Az0)/2 ; By = -im*kx*(Az+ + Az0)/2 */
            const double szr = azp[0] + az0[0], szi = azp[1] + az0[1];
            cmul(0.0 * kyj, kyj, szr, szi, &er, &ei);
            Z[ORCLG_BX][2 * k] = er / 2.0;
            Z[ORCLG_BX][2 * k + 1] = ei / 2.0;
            cmul(0.0 * kxi, -kxi, szr, szi, &er, &ei);
            Z[ORCLG_BY][2 * k] = er / 2.0;
            Z[ORCLG_BY][2 * k + 1] = ei / 2.0;
            /* Bz = im*kx*(Ay+ + Ay0)/2 ; Bz -= im*ky*(Ax+ + Ax0)/2 */
            double fr, fi;
            cmul(0.0 * kxi, kxi, ayp[0] + ay0[0], ayp[1] + ay0[1], &er, &ei);
            cmul(0.0 * kyj, kyj, axp[0] + ax0[0], axp[1] + ax0[1], &fr, &fi);
            Z[ORCLG_BZ][2 * k] = er / 2.0 - fr / 2.0;
            Z[ORCLG_BZ][2 * k + 1] = ei / 2.0 - fi / 2.0;
        }
    for (int a = ORCLG_EX; a <= ORCLG_BZ; ++a) ifft2_inplace(NX, NY, Z[a]);
    free(kx);
    free(ky);
    free(k2);
}

/* update! (:468-479) on the NX x NY grid: EB = 0.0 + real(F), plus B0 for
 * the magnetic components. */
void orclg_update(int32_t NX, int32_t NY, const double* F, const double* B0, double* EB) {
    const int64_t G = (int64_t)NX * NY;
    for (int c = 0; c < 6; ++c) {
        const double* z = F + (size_t)(ORCLG_EX + c) * 2 * G;
        for (int64_t k = 0; k < G; ++k) {
            const double e = 0.0 + z[2 * k];
            EB[c * G + k] = c < 3 ? e : e + B0[c - 3];
        }
    }
}

static i128 fix_signed(double a, int q) {
    uint64_t lo, hi;
    orc_fxp_from_double(fabs(a), q, &lo, &hi);
    const i128 v = (i128)(((u128)hi << 64) | lo);
    return a < 0.0 ? -v : v;
}

static double unfix_signed(i128 v, int q) {
    if (v < 0) {
        const u128 u = (u128)(-v);
        return -orc_fxp_to_double((uint64_t)u, (uint64_t)(u >> 64), q);
    }
    return orc_fxp_to_double((uint64_t)v, (uint64_t)((u128)v >> 64), q);
}

/* q of an exact sum whose contributions are bounded by sum_s P_s*(|w_s|*M_s) */
int orclg_frac_bits(int32_t ns, const int64_t* P, const double* w, const double* M) {
    double total = 0.0;
    for (int s = 0; s < ns; ++s) total += (double)P[s] * (fabs(w[s]) * (M ? M[s] : 1.0));
    if (!(total > 0.0) || !isfinite(total)) return 64;
    const int q = 110 - ilogb(total);
    return q > 1100 ? 1100 : q;
}

/* deposit (wx*wy)*w[p] (or *w0 when w is NULL) of n particles into acc */
static void deposit_into(i128* acc, int32_t NX, int32_t NY, double NX_Lx, double NY_Ly, int32_t shape, int64_t n,
                         const double* x, const double* y, const double* w, double w0, int q) {
    for (int64_t p = 0; p < n; ++p) {
        int64_t cx[6], cy[6];
        double wx[6], wy[6];
        const int nx = orc23_weights(shape, x[p], NX_Lx, cx, wx);
        const int ny = orc23_weights(shape, y[p], NY_Ly, cy, wy);
        const double ww = w ? w[p] : w0;
        for (int jj = 0; jj < ny; ++jj)
            for (int ii = 0; ii < nx; ++ii) {
                const double c = wx[ii] * wy[jj] * ww;
                acc[wrap_cell(cx[ii], NX) + (int64_t)NX * wrap_cell(cy[jj], NY)] += fix_signed(c, q);
            }
    }
}

/* J0 (slots 9..11, imaginary parts 0) as the exact deposit of
 * (vx, vy, vz)*qw_s of every species at its current position. */
static void deposit_currents(int32_t NX, int32_t NY, double Lx, double Ly, int32_t ns, const int32_t* shape,
                             const int64_t* P, const double* qw, const double* state, double* F) {
    const int64_t G = (int64_t)NX * NY;
    const double NX_Lx = (double)NX / Lx, NY_Ly = (double)NY / Ly;
    double M[3][8];
    const double* st = state;
    for (int s = 0; s < ns; ++s) {
        for (int c = 0; c < 3; ++c) {
            double m = 0.0;
            for (int64_t p = 0; p < P[s]; ++p) {
                const double a = fabs(st[(2 + c) * P[s] + p]);
                m = a > m ? a : m;
            }
            M[c][s] = m;
        }
        st += 5 * P[s];
    }
    i128* acc = (i128*)malloc(sizeof(i128) * (size_t)G);
    for (int c = 0; c < 3; ++c) {
        const int q = orclg_frac_bits(ns, P, qw, M[c]);
        memset(acc, 0, sizeof(i128) * (size_t)G);
        st = state;
        for (int s = 0; s < ns; ++s) {
            const int64_t n = P[s];
            double* w = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
            for (int64_t p = 0; p < n; ++p) w[p] = st[(2 + c) * n + p] * qw[s];
            deposit_into(acc, NX, NY, NX_Lx, NY_Ly, shape[s], n, st, st + n, w, 0.0, q);
            free(w);
            st += 5 * n;
        }
        double* z = F + (size_t)(ORCLG_JX0 + c) * 2 * G;
        for (int64_t k = 0; k < G; ++k) {
            z[2 * k] = unfix_signed(acc[k], q);
            z[2 * k + 1] = 0.0;
        }
    }
    free(acc);
}

/* warmup!(::LorenzGaugeStaggeredField) (:826-859): advect by -dt/2, deposit
 * rho (which the second deposit!'s `rhoJs .= 0` discards, :803), advect by
 * +dt/2, deposit J; the reduction leaves rho0 = 0 and J0 = the currents. */
void orclg_warmup(int32_t NX, int32_t NY, double Lx, double Ly, double dt, int32_t ns, const int32_t* shape,
                  const int64_t* P, const double* charge, const double* weight, double* state, double* F) {
    const int64_t G = (int64_t)NX * NY;
    const double dV = (Lx / (double)NX) * (Ly / (double)NY);
    double qw[8];
    for (int s = 0; s < ns; ++s) qw[s] = charge[s] * weight[s] / dV;
    for (int pass = 0; pass < 2; ++pass) {
        const double h = pass == 0 ? -dt / 2.0 : dt / 2.0;
        double* st = state;
        for (int s = 0; s < ns; ++s) {
            const int64_t n = P[s];
            for (int64_t p = 0; p < n; ++p) {
                st[p] = orc23_unimod(st[p] + st[2 * n + p] * h, Lx);
                st[n + p] = orc23_unimod(st[n + p] + st[3 * n + p] * h, Ly);
            }
            st += 5 * n;
        }
    }
    double* rho = F + (size_t)ORCLG_RHO0 * 2 * G;
    for (int64_t k = 0; k < 2 * G; ++k) rho[k] = 0.0;
    deposit_currents(NX, NY, Lx, Ly, ns, shape, P, qw, state, F);
}

/* the row of diagnose! (:1336-1357): kinetic energy, field energy (mean of
 * abs2 over the six ghosted EBxyz components, /2), particle momentum (3),
 * field momentum (sum of real(E) x real(B) over the cells / NX*NY, 3),
 * characteristic momentum (3). */
void orclg_row(int32_t NX, int32_t NY, int32_t buffer, const double* F, const double* EB, int32_t ns,
               const int64_t* P, const double* mass, const double* weight, const double* state, double* row) {
    const int64_t G = (int64_t)NX * NY;
    const int64_t NXp = NX + 2 * (int64_t)buffer, NYp = NY + 2 * (int64_t)buffer;
    const int64_t nf = 6 * NXp * NYp;
    double* t = (double*)malloc(sizeof(double) * (size_t)(nf > 3 * G ? nf : 3 * G));
    int64_t e = 0;
    for (int64_t jp = 0; jp < NYp; ++jp)
        for (int64_t ip = 0; ip < NXp; ++ip) {
            const int64_t k = wrap_cell(ip - (buffer - 1), NX) + (int64_t)NX * wrap_cell(jp - (buffer - 1), NY);
            for (int c = 0; c < 6; ++c) {
                const double v = EB[c * G + k];
                t[e++] = v * v;
            }
        }
    double pm[3], pc[3];
    orc23_species_moments(ns, P, mass, weight, state, &row[0], pm, pc);
    row[1] = orc_tree_sum(t, nf) / (double)nf / 2.0;
    const double *ex = F + (size_t)ORCLG_EX * 2 * G, *ey = F + (size_t)ORCLG_EY * 2 * G,
                 *ez = F + (size_t)ORCLG_EZ * 2 * G, *bx = F + (size_t)ORCLG_BX * 2 * G,
                 *by = F + (size_t)ORCLG_BY * 2 * G, *bz = F + (size_t)ORCLG_BZ * 2 * G;
    for (int c = 0; c < 3; ++c) {
        for (int64_t k = 0; k < G; ++k) {
            const double Ex = ex[2 * k], Ey = ey[2 * k], Ez = ez[2 * k];
            const double Bx = bx[2 * k], By = by[2 * k], Bz = bz[2 * k];
            t[k] = c == 0 ? Ey * Bz - Ez * By : (c == 1 ? Ez * Bx - Ex * Bz : Ex * By - Ey * Bx);
        }
        row[5 + c] = orc_tree_sum(t, G) / (double)G;
    }
    for (int c = 0; c < 3; ++c) {
        row[2 + c] = pm[c];
        row[8 + c] = pc[c];
    }
    free(t);
}

/* diagnose!(d, field, plasma, t) (:1332-1398) on the state F, EB after a step */
static void lg_diagnose(int32_t NX, int32_t NY, int32_t buffer, int32_t ntskip, int32_t ngskip, double* F,
                        const double* EB, int32_t ns, const int64_t* P, const double* mass, const double* weight,
                        const double* state, int64_t t, double* rows, int64_t* nrow, double* bins, int64_t nbins) {
    const int64_t G = (int64_t)NX * NY;
    const int32_t NXd = NX / ngskip, NYd = NY / ngskip;
    if (t % ntskip == 0 && rows) orclg_row(NX, NY, buffer, F, EB, ns, P, mass, weight, state, rows + 11 * (*nrow)++);
    static const int tr[8] = {ORCLG_AX0, ORCLG_AY0, ORCLG_AZ0, ORCLG_PHI0, ORCLG_RHO0, ORCLG_JX0, ORCLG_JY0, ORCLG_JZ0};
    for (int a = 0; a < 8; ++a) ifft2_inplace(NX, NY, F + (size_t)tr[a] * 2 * G);
    const int64_t b = t / ntskip;
    if (!bins || b >= nbins) return;
    static const int av[14] = {ORCLG_EX,  ORCLG_EY,  ORCLG_EZ,  ORCLG_BX,   ORCLG_BY,   ORCLG_BZ,  ORCLG_AX0,
                               ORCLG_AY0, ORCLG_AZ0, ORCLG_PHI0, ORCLG_RHO0, ORCLG_JX0, ORCLG_JY0, ORCLG_JZ0};
    const double factor = 1.0 / (double)((int64_t)ntskip * ngskip * ngskip);
    for (int a = 0; a < 14; ++a) {
        const double* z = F + (size_t)av[a] * 2 * G;
        double* o = bins + ((size_t)a * nbins + b) * (size_t)NXd * NYd;
        for (int32_t jl = 0; jl < NYd; ++jl)
            for (int32_t il = 0; il < NXd; ++il) {
                double acc2 = o[il + (int64_t)NXd * jl];
                for (int32_t jj = 0; jj < ngskip; ++jj)
                    for (int32_t ii = 0; ii < ngskip; ++ii) {
                        const int64_t k = (int64_t)il * ngskip + ii + (int64_t)NX * ((int64_t)jl * ngskip + jj);
                        acc2 = acc2 + z[2 * k] * factor;
                    }
                o[il + (int64_t)NXd * jl] = acc2;
            }
    }
}

/* ------------------------------------------------------------------------ */
/* Time loop: for t = t0 .. t0+nsteps-1, loop!(plasma, field) (:875-979) and, */
/* if diagnose, diagnose!(d, field, plasma, t) (:1332-1398):                 */
/*   if t % ntskip == 0 the row of bin t/ntskip (11 doubles, orclg_row);      */
/*   ifft! in place of Ax0, Ay0, Az0, phi0, rho0, Jx0, Jy0, Jz0 (:1359-1368); */
/*   every step, bins[f][t/ntskip][il, jl] += real(Z[il*g + ii, jl*g + jj]) *  */
/*   (1/(ntskip*g^2)) for jj outer, ii inner (:1370-1386), g = ngskip, for   */
/*   the fourteen fields Ex, Ey, Ez, Bx, By, Bz, Ax0, Ay0, Az0, phi0, rho0,   */
/*   Jx0, Jy0, Jz0 (after the transforms).                                    */
/* state: species after species, each x, y, vx, vy, vz (P[s] each).  rows     */
/* receives one row per t % ntskip == 0 step of this call; bins (may be NULL) */
/* holds 14 x nbins x (NX/ngskip) x (NY/ngskip).                             */
/* ------------------------------------------------------------------------ */
int orclg_run(int32_t NX, int32_t NY, double Lx, double Ly, double dt, const double* B0, int32_t buffer,
              double theta, int32_t ntskip, int32_t ngskip, int32_t ns, const int32_t* shape, const int64_t* P,
              const double* charge, const double* mass, const double* weight, double* state, double* F,
              int64_t t0, int64_t nsteps, int32_t diagnose, double* rows, double* bins, int64_t nbins) {
    if (ns < 1 || ns > 8 || NX < 2 || NY < 2 || (NX & (NX - 1)) || (NY & (NY - 1)) || ntskip < 1 || ngskip < 1 ||
        NX % ngskip || NY % ngskip)
        return -1;
    const int64_t G = (int64_t)NX * NY;
    const double NX_Lx = (double)NX / Lx, NY_Ly = (double)NY / Ly;
    const double dV = (Lx / (double)NX) * (Ly / (double)NY);
    double qw[8];
    for (int s = 0; s < ns; ++s) qw[s] = charge[s] * weight[s] / dV;
    const int qr = orclg_frac_bits(ns, P, qw, NULL);
    double* plus = (double*)malloc(sizeof(double) * 8 * (size_t)G);
    double* EB = (double*)malloc(sizeof(double) * 6 * (size_t)G);
    i128* acc = (i128*)malloc(sizeof(i128) * (size_t)G);
    int64_t Ptot = 0;
    for (int s = 0; s < ns; ++s) Ptot += P[s];
    double* xh = (double*)malloc(sizeof(double) * (size_t)(Ptot > 0 ? Ptot : 1));
    double* yh = (double*)malloc(sizeof(double) * (size_t)(Ptot > 0 ? Ptot : 1));
    int64_t nrow = 0;
    for (int64_t t = t0; t < t0 + nsteps; ++t) {
        orclg_solve(NX, NY, Lx, Ly, dt, theta, F, plus);
        orclg_update(NX, NY, F, B0, EB);
        /* particles: gather, Boris, half move (the rho position), half move */
        double* st = state;
        int64_t off = 0;
        for (int s = 0; s < ns; ++s) {
            const int64_t n = P[s];
            double *x = st, *y = st + n, *vx = st + 2 * n, *vy = st + 3 * n, *vz = st + 4 * n;
            const double q_m = charge[s] / mass[s];
            for (int64_t p = 0; p < n; ++p) {
                double f[6], v[3] = {vx[p], vy[p], vz[p]};
                orclg_gather(NX, NY, NX_Lx, NY_Ly, shape[s], EB, x[p], y[p], f);
                orclg_boris(dt, q_m, f, f + 3, v);
                vx[p] = v[0];
                vy[p] = v[1];
                vz[p] = v[2];
                x[p] = orc23_unimod(x[p] + vx[p] * dt / 2.0, Lx);
                y[p] = orc23_unimod(y[p] + vy[p] * dt / 2.0, Ly);
                xh[off + p] = x[p];
                yh[off + p] = y[p];
                x[p] = orc23_unimod(x[p] + vx[p] * dt / 2.0, Lx);
                y[p] = orc23_unimod(y[p] + vy[p] * dt / 2.0, Ly);
            }
            st += 5 * n;
            off += n;
        }
        /* reduction: rho0 from the half-step positions, J0 from the new ones */
        memset(acc, 0, sizeof(i128) * (size_t)G);
        off = 0;
        for (int s = 0; s < ns; ++s) {
            deposit_into(acc, NX, NY, NX_Lx, NY_Ly, shape[s], P[s], xh + off, yh + off, NULL, qw[s], qr);
            off += P[s];
        }
        double* rho = F + (size_t)ORCLG_RHO0 * 2 * G;
        for (int64_t k = 0; k < G; ++k) {
            rho[2 * k] = unfix_signed(acc[k], qr);
            rho[2 * k + 1] = 0.0;
        }
        deposit_currents(NX, NY, Lx, Ly, ns, shape, P, qw, state, F);
        /* phi- = phi0, phi0 = phi+, A- = A0, A0 = A+ */
        static const int x0s[4] = {ORCLG_PHI0, ORCLG_AX0, ORCLG_AY0, ORCLG_AZ0};
        static const int xms[4] = {ORCLG_PHIM, ORCLG_AXM, ORCLG_AYM, ORCLG_AZM};
        for (int a = 0; a < 4; ++a) {
            memcpy(F + (size_t)xms[a] * 2 * G, F + (size_t)x0s[a] * 2 * G, sizeof(double) * 2 * (size_t)G);
            memcpy(F + (size_t)x0s[a] * 2 * G, plus + (size_t)a * 2 * G, sizeof(double) * 2 * (size_t)G);
        }
        if (diagnose)
            lg_diagnose(NX, NY, buffer, ntskip, ngskip, F, EB, ns, P, mass, weight, state, t, rows, &nrow, bins,
                        nbins);
    }
    free(plus);
    free(EB);
    free(acc);
    free(xh);
    free(yh);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* LorenzGaugeField (:307-345) and its loop! (:665-745), the centred scheme. */
/* The state slots hold: PHI0 = phi+ (the latest potential), PHIM = phi0     */
/* (the one before), AX0.. = A+, AXM.. = A0, RHO0 = rho+, J0 = J+ (in the    */
/* spectrum, [1,1] = 0, as the step leaves them), EX..BZ as before.  EB (the */
/* 6 x NX x NY interior of EBxyz) is state too: the particle loop runs first */
/* and gathers what the previous step's update! left (zeros at the start).  */
/*   1. particles: gather at x, ElectromagneticBoris, x = unimod(x +          */
/*      (vx0 + vx)/2*dt, Lx), deposit qw, vx*qw, vy*qw, vz*qw at the new x   */
/*      (:672-694); reduction into rho+, J+ (:696-699)                       */
/*   2. fft! of rho+, J+ and their [1,1] = 0 (:700-709)                      */
/*   3. lorenzgauge!(imex, x+, x0, s, k2, dt^2) (:619-627): new = (num*x+ +  */
/*      dt^2*s)/den - x0; x0 = x+; x+ = new (:711-714)                       */
/*   4. E = -im*k*(phi+ + phi0)/2 - (A+ - A0)/dt, B = curl (A+ + A0)/2        */
/*      (:717-725), ifft of the six, update! (:727-735)                      */
/* The reference has no warmup! for this type (warmup!(::AbstractField) is   */
/* the identity, :745) and its diagnose! reads f.rho0, which the type does   */
/* not have (:1366), so only loop! runs.                                      */
/* ------------------------------------------------------------------------ */
static void fft2_inplace(int32_t NX, int32_t NY, double* z) {
    const int64_t G = (int64_t)NX * NY;
    double* re = (double*)malloc(sizeof(double) * (size_t)G);
    double* im = (double*)malloc(sizeof(double) * (size_t)G);
    for (int64_t k = 0; k < G; ++k) {
        re[k] = z[2 * k];
        im[k] = z[2 * k + 1];
    }
    orc2d_fft(NX, NY, re, im, 0);
    for (int64_t k = 0; k < G; ++k) {
        z[2 * k] = re[k];
        z[2 * k + 1] = im[k];
    }
    free(re);
    free(im);
}

void orclg_fft2(int32_t NX, int32_t NY, double* z) { fft2_inplace(NX, NY, z); }

/* steps 3-4 on the state (the sources already in the spectrum) */
void orclg_solve_centred(int32_t NX, int32_t NY, double Lx, double Ly, double dt, double theta, double* F) {
    const int64_t G = (int64_t)NX * NY;
    double* kx = (double*)malloc(sizeof(double) * (size_t)NX);
    double* ky = (double*)malloc(sizeof(double) * (size_t)NY);
    double* k2 = (double*)malloc(sizeof(double) * (size_t)G);
    orclg_wavenumbers(NX, NY, Lx, Ly, kx, ky, k2);
    const double dt2 = dt * dt;
    double* Z[18];
    for (int a = 0; a < 18; ++a) Z[a] = F + (size_t)a * 2 * G;
    static const int xps[4] = {ORCLG_PHI0, ORCLG_AX0, ORCLG_AY0, ORCLG_AZ0};
    static const int x0s[4] = {ORCLG_PHIM, ORCLG_AXM, ORCLG_AYM, ORCLG_AZM};
    static const int srcs[4] = {ORCLG_RHO0, ORCLG_JX0, ORCLG_JY0, ORCLG_JZ0};
    for (int32_t j = 0; j < NY; ++j)
        for (int32_t i = 0; i < NX; ++i) {
            const int64_t k = i + (int64_t)NX * j;
            const double num = 2.0 - dt2 * k2[k] * (1.0 - theta);
            const double den = 1.0 + dt2 * k2[k] * theta / 2.0;
            for (int a = 0; a < 4; ++a) {
                double* xp = Z[xps[a]] + 2 * k;
                double* x0 = Z[x0s[a]] + 2 * k;
                const double* sr = Z[srcs[a]] + 2 * k;
                const double nr = (num * xp[0] + dt2 * sr[0]) / den - x0[0];
                const double ni = (num * xp[1] + dt2 * sr[1]) / den - x0[1];
                x0[0] = xp[0];
                x0[1] = xp[1];
                xp[0] = nr;
                xp[1] = ni;
            }
            const double kxi = kx[i], kyj = ky[j];
            const double *php = Z[ORCLG_PHI0] + 2 * k, *ph0 = Z[ORCLG_PHIM] + 2 * k;
            const double *axp = Z[ORCLG_AX0] + 2 * k, *ayp = Z[ORCLG_AY0] + 2 * k, *azp = Z[ORCLG_AZ0] + 2 * k;
            const double *ax0 = Z[ORCLG_AXM] + 2 * k, *ay0 = Z[ORCLG_AYM] + 2 * k, *az0 = Z[ORCLG_AZM] + 2 * k;
            const double sr = php[0] + ph0[0], si = php[1] + ph0[1];
            double er, ei, fr, fi;
            /* Ex = -im*kx*(phi+ + phi0)/2 ; Ex -= (Ax+ - Ax0)/dt */
            cmul(0.0 * kxi, -kxi, sr, si, &er, &ei);
            Z[ORCLG_EX][2 * k] = er / 2.0 - (axp[0] - ax0[0]) / dt;
            Z[ORCLG_EX][2 * k + 1] = ei / 2.0 - (axp[1] - ax0[1]) / dt;
            cmul(0.0 * kyj, -kyj, sr, si, &er, &ei);
            Z[ORCLG_EY][2 * k] = er / 2.0 - (ayp[0] - ay0[0]) / dt;
            Z[ORCLG_EY][2 * k + 1] = ei / 2.0 - (ayp[1] - ay0[1]) / dt;
            Z[ORCLG_EZ][2 * k] = -(azp[0] - az0[0]) / dt;
            Z[ORCLG_EZ][2 * k + 1] = -(azp[1] - az0[1]) / dt;
            const double szr = azp[0] + az0[0], szi = azp[1] + az0[1];
            cmul(0.0 * kyj, kyj, szr, szi, &er, &ei);
            Z[ORCLG_BX][2 * k] = er / 2.0;
            Z[ORCLG_BX][2 * k + 1] = ei / 2.0;
            cmul(0.0 * kxi, -kxi, szr, szi, &er, &ei);
            Z[ORCLG_BY][2 * k] = er / 2.0;
            Z[ORCLG_BY][2 * k + 1] = ei / 2.0;
            cmul(0.0 * kxi, kxi, ayp[0] + ay0[0], ayp[1] + ay0[1], &er, &ei);
            cmul(0.0 * kyj, kyj, axp[0] + ax0[0], axp[1] + ax0[1], &fr, &fi);
            Z[ORCLG_BZ][2 * k] = er / 2.0 - fr / 2.0;
            Z[ORCLG_BZ][2 * k + 1] = ei / 2.0 - fi / 2.0;
        }
    for (int a = ORCLG_EX; a <= ORCLG_BZ; ++a) ifft2_inplace(NX, NY, Z[a]);
    free(kx);
    free(ky);
    free(k2);
}

int orclg_run_centred(int32_t NX, int32_t NY, double Lx, double Ly, double dt, const double* B0, double theta,
                      int32_t ns, const int32_t* shape, const int64_t* P, const double* charge, const double* mass,
                      const double* weight, double* state, double* F, double* EB, int64_t nsteps) {
    if (ns < 1 || ns > 8 || NX < 2 || NY < 2 || (NX & (NX - 1)) || (NY & (NY - 1))) return -1;
    const int64_t G = (int64_t)NX * NY;
    const double NX_Lx = (double)NX / Lx, NY_Ly = (double)NY / Ly;
    const double dV = (Lx / (double)NX) * (Ly / (double)NY);
    double qw[8];
    for (int s = 0; s < ns; ++s) qw[s] = charge[s] * weight[s] / dV;
    const int qr = orclg_frac_bits(ns, P, qw, NULL);
    i128* acc = (i128*)malloc(sizeof(i128) * (size_t)G);
    for (int64_t t = 0; t < nsteps; ++t) {
        double* st = state;
        for (int s = 0; s < ns; ++s) {
            const int64_t n = P[s];
            double *x = st, *y = st + n, *vx = st + 2 * n, *vy = st + 3 * n, *vz = st + 4 * n;
            const double q_m = charge[s] / mass[s];
            for (int64_t p = 0; p < n; ++p) {
                double f[6], v[3] = {vx[p], vy[p], vz[p]};
                orclg_gather(NX, NY, NX_Lx, NY_Ly, shape[s], EB, x[p], y[p], f);
                const double vx0 = vx[p], vy0 = vy[p];
                orclg_boris(dt, q_m, f, f + 3, v);
                vx[p] = v[0];
                vy[p] = v[1];
                vz[p] = v[2];
                x[p] = orc23_unimod(x[p] + (vx0 + vx[p]) / 2.0 * dt, Lx);
                y[p] = orc23_unimod(y[p] + (vy0 + vy[p]) / 2.0 * dt, Ly);
            }
            st += 5 * n;
        }
        memset(acc, 0, sizeof(i128) * (size_t)G);
        st = state;
        for (int s = 0; s < ns; ++s) {
            deposit_into(acc, NX, NY, NX_Lx, NY_Ly, shape[s], P[s], st, st + P[s], NULL, qw[s], qr);
            st += 5 * P[s];
        }
        double* rho = F + (size_t)ORCLG_RHO0 * 2 * G;
        for (int64_t k = 0; k < G; ++k) {
            rho[2 * k] = unfix_signed(acc[k], qr);
            rho[2 * k + 1] = 0.0;
        }
        deposit_currents(NX, NY, Lx, Ly, ns, shape, P, qw, state, F);
        for (int a = ORCLG_RHO0; a <= ORCLG_JZ0; ++a) {
            double* z = F + (size_t)a * 2 * G;
            fft2_inplace(NX, NY, z);
            z[0] = 0.0;
            z[1] = 0.0;
        }
        orclg_solve_centred(NX, NY, Lx, Ly, dt, theta, F);
        orclg_update(NX, NY, F, B0, EB);
    }
    free(acc);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* LorenzGaugeSemiImplicitField (:403-461): warmup! (:864-871) and loop!     */
/* (:993-1100).  State: the 18 slots above (PHI0/PHIM = phi0/phi-, A0/A-,    */
/* RHO0, J0 the n sources, EX..BZ) and four more, RHOM, JXM, JYM, JZM = rho-, */
/* J-; the buffered deposit grid rhoJs+ (4 x (NX+2b) x (NY+2b), axes         */
/* -(b-1):NX+b, in the reference's element order, component fastest), which  */
/* the next step's first convergence test reads; and EB, the gathered fields */
/* update! left (zeros until the first update!).                             */
/* warmup!: three passes of the four-weight warmup! (:746-775) -- move by h, */
/* deposit qw, vx*qw, vy*qw, vz*qw, move back by h -- with h = -dt (into     */
/* rho-, J-), +dt (rho0, J0) and +dt (rho+, J+ and the grid rhoJs+), then    */
/* advect! by -dt.  rho+, J+ (temporaries the loop overwrites before it      */
/* reads them) and rhoJs-, rhoJs0 (which loop! only copies into each other)  */
/* are not observable and are not kept.                                      */
/* loop!: from the step's start state (x, v), iteration k:                   */
/*   1. gather at (x, y) from EB, xw = unimod(x + vx*dt) with the old v,     */
/*      vw = ElectromagneticBoris(v); the currents vw*qw at xw into the      */
/*      buffered grid (:1018-1045; its charge component stays zero), the    */
/*      reduction into rho+ (zero) and J+ (:1046-1049);                      */
/*   2. fft! of rho+, J+ and [1,1] = 0 (:1050-1059);                         */
/*   3. lorenzgauge! (:638-648) of phi, Ax, Ay, Az with the field theta_f    */
/*      and the source theta_s/2*s- + (1-theta_s)*s0 + theta_s/2*s+          */
/*      (:1060-1065);                                                        */
/*   4. E = -im k (theta_f/2 (phi+ + phi-) + (1-theta_f) phi0) -             */
/*      (A+ - A-)/2dt, B = curl (theta_f/2 (A+ + A-) + (1-theta_f) A0)       */
/*      (:1066-1077), ifft of the six, update! (:1078-1086).                 */
/* The loop stops after iteration k when k > maxiters or                     */
/* isapprox(rhoJs_prev, rhoJs+; rtol, atol = 0) (:1000-1012, orc_isapprox),  */
/* rhoJs_prev being the grid before iteration k (the previous step's, or     */
/* warmup!'s, at k = 1); then x, v = xw, vw and phi- = phi0, phi0 = phi+,    */
/* A- = A0, A0 = A+ (:1089-1100).  The sources rho0, J0, rho-, J- keep their */
/* warmup! values, in real space, except where diagnose! transforms rho0, J0 */
/* in place.  The reference keeps one buffered grid per thread and its test  */
/* compares them all, so its iteration counts depend on nthreads(); this is  */
/* the one-thread grid (julia's default).  Charge and current sums are exact */
/* per buffered cell and per periodic cell (the exact sum over the cell's    */
/* images), each rounded once.  A stencil cell outside the buffer (the       */
/* reference's BoundsError) returns -2.                                      */
/* ------------------------------------------------------------------------ */
static int deposit_buffered(int32_t NX, int32_t NY, int32_t b, double Lx, double Ly, int32_t ns,
                            const int32_t* shape, const int64_t* P, const double* qw, const double* state, int c0,
                            double* grid, double* const* per) {
    const int64_t G = (int64_t)NX * NY;
    const int64_t NXe = NX + 2 * (int64_t)b, NYe = NY + 2 * (int64_t)b, Ge = NXe * NYe;
    const double NX_Lx = (double)NX / Lx, NY_Ly = (double)NY / Ly;
    double M[3][8];
    const double* st = state;
    for (int s = 0; s < ns; ++s) {
        for (int c = 0; c < 3; ++c) {
            double m = 0.0;
            for (int64_t p = 0; p < P[s]; ++p) {
                const double a = fabs(st[(2 + c) * P[s] + p]);
                m = a > m ? a : m;
            }
            M[c][s] = m;
        }
        st += 5 * P[s];
    }
    i128* acc = (i128*)malloc(sizeof(i128) * (size_t)Ge);
    i128* pacc = (i128*)malloc(sizeof(i128) * (size_t)G);
    int rc = 0;
    if (grid)
        for (int64_t e = 0; e < 4 * Ge; ++e) grid[e] = 0.0;
    for (int c = c0; c < 4; ++c) {
        const int q = orclg_frac_bits(ns, P, qw, c == 0 ? NULL : M[c - 1]);
        memset(acc, 0, sizeof(i128) * (size_t)Ge);
        memset(pacc, 0, sizeof(i128) * (size_t)G);
        st = state;
        for (int s = 0; s < ns; ++s) {
            const int64_t n = P[s];
            for (int64_t p = 0; p < n; ++p) {
                const double w = c == 0 ? qw[s] : st[(1 + c) * n + p] * qw[s];
                int64_t cx[6], cy[6];
                double wx[6], wy[6];
                const int nx = orc23_weights(shape[s], st[p], NX_Lx, cx, wx);
                const int ny = orc23_weights(shape[s], st[n + p], NY_Ly, cy, wy);
                for (int jj = 0; jj < ny; ++jj)
                    for (int ii = 0; ii < nx; ++ii) {
                        const double term = wx[ii] * wy[jj] * w;
                        const int64_t u = cx[ii] - 1 + b, v = cy[jj] - 1 + b;
                        if (u < 0 || u >= NXe || v < 0 || v >= NYe) {
                            rc = -2;
                            continue;
                        }
                        const i128 f = fix_signed(term, q);
                        acc[u + NXe * v] += f;
                        pacc[wrap_cell(cx[ii], NX) + (int64_t)NX * wrap_cell(cy[jj], NY)] += f;
                    }
            }
            st += 5 * n;
        }
        if (grid)
            for (int64_t e = 0; e < Ge; ++e) grid[c + 4 * e] = unfix_signed(acc[e], q);
        if (per && per[c])
            for (int64_t k = 0; k < G; ++k) {
                per[c][2 * k] = unfix_signed(pacc[k], q);
                per[c][2 * k + 1] = 0.0;
            }
    }
    free(acc);
    free(pacc);
    return rc;
}

/* the deposits alone (tests): c0 = 0 all four weights, 1 the currents */
int orclg_si_deposit(int32_t NX, int32_t NY, int32_t buffer, double Lx, double Ly, int32_t ns, const int32_t* shape,
                     const int64_t* P, const double* charge, const double* weight, const double* state, int32_t c0,
                     double* grid, double* per) {
    const double dV = (Lx / (double)NX) * (Ly / (double)NY);
    double qw[8];
    for (int s = 0; s < ns; ++s) qw[s] = charge[s] * weight[s] / dV;
    const int64_t G = (int64_t)NX * NY;
    double* pp[4] = {per, per + 2 * G, per + 4 * G, per + 6 * G};
    return deposit_buffered(NX, NY, buffer, Lx, Ly, ns, shape, P, qw, state, c0, grid, per ? pp : NULL);
}

static void advect_all(int32_t ns, const int64_t* P, double Lx, double Ly, double h, double* state) {
    double* st = state;
    for (int s = 0; s < ns; ++s) {
        const int64_t n = P[s];
        for (int64_t p = 0; p < n; ++p) {
            st[p] = orc23_unimod(st[p] + st[2 * n + p] * h, Lx);
            st[n + p] = orc23_unimod(st[n + p] + st[3 * n + p] * h, Ly);
        }
        st += 5 * n;
    }
}

/* F: 22 complex slots; grid: 4 x (NX+2b) x (NY+2b) */
int orclg_si_warmup(int32_t NX, int32_t NY, double Lx, double Ly, double dt, int32_t buffer, int32_t ns,
                    const int32_t* shape, const int64_t* P, const double* charge, const double* weight, double* state,
                    double* F, double* grid) {
    if (ns < 1 || ns > 8 || NX < 2 || NY < 2 || (NX & (NX - 1)) || (NY & (NY - 1)) || buffer < 0) return -1;
    const int64_t G = (int64_t)NX * NY;
    const double dV = (Lx / (double)NX) * (Ly / (double)NY);
    double qw[8];
    for (int s = 0; s < ns; ++s) qw[s] = charge[s] * weight[s] / dV;
    int rc = 0;
    for (int pass = 0; pass < 3; ++pass) {
        const double h = pass == 0 ? -dt : dt;
        const int first = pass == 0 ? ORCLG_RHOM : ORCLG_RHO0;
        double* pp[4] = {F + (size_t)first * 2 * G, F + (size_t)(first + 1) * 2 * G, F + (size_t)(first + 2) * 2 * G,
                         F + (size_t)(first + 3) * 2 * G};
        advect_all(ns, P, Lx, Ly, h, state);
        const int r = deposit_buffered(NX, NY, buffer, Lx, Ly, ns, shape, P, qw, state, 0, pass == 2 ? grid : NULL,
                                       pass == 2 ? NULL : pp);
        if (r) rc = r;
        advect_all(ns, P, Lx, Ly, -h, state);
    }
    advect_all(ns, P, Lx, Ly, -dt, state);
    return rc;
}

/* steps 2-4 of an iteration: src (4 complex: rho+, Jx+, Jy+, Jz+ in real
 * space) is transformed in place and its [1,1] cleared; plus gets phi+, A+;
 * the six E, B slots of F get their real-space values. */
void orclg_si_solve(int32_t NX, int32_t NY, double Lx, double Ly, double dt, double theta_f, double theta_s,
                    double* F, double* src, double* plus) {
    const int64_t G = (int64_t)NX * NY;
    for (int a = 0; a < 4; ++a) {
        double* z = src + (size_t)a * 2 * G;
        fft2_inplace(NX, NY, z);
        z[0] = 0.0;
        z[1] = 0.0;
    }
    double* kx = (double*)malloc(sizeof(double) * (size_t)NX);
    double* ky = (double*)malloc(sizeof(double) * (size_t)NY);
    double* k2 = (double*)malloc(sizeof(double) * (size_t)G);
    orclg_wavenumbers(NX, NY, Lx, Ly, kx, ky, k2);
    const double dt2 = dt * dt, dt_2 = 2.0 * dt;
    const double hs = theta_s / 2.0, ws = 1.0 - theta_s, hf = theta_f / 2.0, wf = 1.0 - theta_f;
    static const int x0s[4] = {ORCLG_PHI0, ORCLG_AX0, ORCLG_AY0, ORCLG_AZ0};
    static const int xms[4] = {ORCLG_PHIM, ORCLG_AXM, ORCLG_AYM, ORCLG_AZM};
    static const int s0s[4] = {ORCLG_RHO0, ORCLG_JX0, ORCLG_JY0, ORCLG_JZ0};
    static const int sms[4] = {ORCLG_RHOM, ORCLG_JXM, ORCLG_JYM, ORCLG_JZM};
    for (int32_t j = 0; j < NY; ++j)
        for (int32_t i = 0; i < NX; ++i) {
            const int64_t k = i + (int64_t)NX * j;
            const double num = 2.0 - dt2 * k2[k] * (1.0 - theta_f);
            const double den = 1.0 + dt2 * k2[k] * theta_f / 2.0;
            double pl[4][2], c0[4][2];
            for (int a = 0; a < 4; ++a) {
                const double* sm = F + (size_t)sms[a] * 2 * G + 2 * k;
                const double* s0 = F + (size_t)s0s[a] * 2 * G + 2 * k;
                const double* sp = src + (size_t)a * 2 * G + 2 * k;
                const double* x0 = F + (size_t)x0s[a] * 2 * G + 2 * k;
                const double* xm = F + (size_t)xms[a] * 2 * G + 2 * k;
                for (int r = 0; r < 2; ++r) {
                    const double sv = hs * sm[r] + ws * s0[r] + hs * sp[r];
                    pl[a][r] = (num * x0[r] + dt2 * sv) / den - xm[r];
                    /* theta_f/2*(x+ + x-) + (1-theta_f)*x0 */
                    c0[a][r] = hf * (pl[a][r] + xm[r]) + wf * x0[r];
                    plus[(size_t)a * 2 * G + 2 * k + r] = pl[a][r];
                }
            }
            const double kxi = kx[i], kyj = ky[j];
            const double* axm = F + (size_t)ORCLG_AXM * 2 * G + 2 * k;
            const double* aym = F + (size_t)ORCLG_AYM * 2 * G + 2 * k;
            const double* azm = F + (size_t)ORCLG_AZM * 2 * G + 2 * k;
            double er, ei, fr, fi;
            double* Z;
            cmul(0.0 * kxi, -kxi, c0[0][0], c0[0][1], &er, &ei);
            Z = F + (size_t)ORCLG_EX * 2 * G + 2 * k;
            Z[0] = er - (pl[1][0] - axm[0]) / dt_2;
            Z[1] = ei - (pl[1][1] - axm[1]) / dt_2;
            cmul(0.0 * kyj, -kyj, c0[0][0], c0[0][1], &er, &ei);
            Z = F + (size_t)ORCLG_EY * 2 * G + 2 * k;
            Z[0] = er - (pl[2][0] - aym[0]) / dt_2;
            Z[1] = ei - (pl[2][1] - aym[1]) / dt_2;
            Z = F + (size_t)ORCLG_EZ * 2 * G + 2 * k;
            Z[0] = -(pl[3][0] - azm[0]) / dt_2;
            Z[1] = -(pl[3][1] - azm[1]) / dt_2;
            cmul(0.0 * kyj, kyj, c0[3][0], c0[3][1], &er, &ei);
            Z = F + (size_t)ORCLG_BX * 2 * G + 2 * k;
            Z[0] = er;
            Z[1] = ei;
            cmul(0.0 * kxi, -kxi, c0[3][0], c0[3][1], &er, &ei);
            Z = F + (size_t)ORCLG_BY * 2 * G + 2 * k;
            Z[0] = er;
            Z[1] = ei;
            cmul(0.0 * kxi, kxi, c0[2][0], c0[2][1], &er, &ei);
            cmul(0.0 * kyj, kyj, c0[1][0], c0[1][1], &fr, &fi);
            Z = F + (size_t)ORCLG_BZ * 2 * G + 2 * k;
            Z[0] = er - fr;
            Z[1] = ei - fi;
        }
    for (int a = ORCLG_EX; a <= ORCLG_BZ; ++a) ifft2_inplace(NX, NY, F + (size_t)a * 2 * G);
    free(kx);
    free(ky);
    free(k2);
}

/* nsteps of loop! and (diagnose) diagnose!; iters[t - t0] gets the step's
 * iteration count (may be NULL). */
int orclg_si_run(int32_t NX, int32_t NY, double Lx, double Ly, double dt, const double* B0, int32_t buffer,
                 double theta_f, double theta_s, double rtol, int64_t maxiters, int32_t ntskip, int32_t ngskip,
                 int32_t ns, const int32_t* shape, const int64_t* P, const double* charge, const double* mass,
                 const double* weight, double* state, double* F, double* grid, double* EB, int64_t t0,
                 int64_t nsteps, int32_t diagnose, double* rows, double* bins, int64_t nbins, int64_t* iters) {
    if (ns < 1 || ns > 8 || NX < 2 || NY < 2 || (NX & (NX - 1)) || (NY & (NY - 1)) || ntskip < 1 || ngskip < 1 ||
        NX % ngskip || NY % ngskip || buffer < 0 || maxiters < 0)
        return -1;
    const int64_t G = (int64_t)NX * NY;
    const int64_t Ge = (NX + 2 * (int64_t)buffer) * (NY + 2 * (int64_t)buffer);
    const double NX_Lx = (double)NX / Lx, NY_Ly = (double)NY / Ly;
    const double dV = (Lx / (double)NX) * (Ly / (double)NY);
    double qw[8];
    for (int s = 0; s < ns; ++s) qw[s] = charge[s] * weight[s] / dV;
    int64_t Ptot = 0;
    for (int s = 0; s < ns; ++s) Ptot += P[s];
    double* work = (double*)malloc(sizeof(double) * 5 * (size_t)(Ptot > 0 ? Ptot : 1));
    double* src = (double*)malloc(sizeof(double) * 8 * (size_t)G);
    double* plus = (double*)malloc(sizeof(double) * 8 * (size_t)G);
    double* prev = (double*)malloc(sizeof(double) * 4 * (size_t)Ge);
    double* pp[4] = {src, src + 2 * G, src + 4 * G, src + 6 * G};
    int rc = 0;
    int64_t nrow = 0;
    for (int64_t t = t0; t < t0 + nsteps && rc == 0; ++t) {
        int64_t k = 0;
        int conv = 0;
        while (!(k > 0 && (k > maxiters || conv))) {
            ++k;
            memcpy(prev, grid, sizeof(double) * 4 * (size_t)Ge);
            const double* st = state;
            double* wk = work;
            for (int s = 0; s < ns; ++s) {
                const int64_t n = P[s];
                const double q_m = charge[s] / mass[s];
                for (int64_t p = 0; p < n; ++p) {
                    double f[6], v[3] = {st[2 * n + p], st[3 * n + p], st[4 * n + p]};
                    orclg_gather(NX, NY, NX_Lx, NY_Ly, shape[s], EB, st[p], st[n + p], f);
                    wk[p] = orc23_unimod(st[p] + v[0] * dt, Lx);
                    wk[n + p] = orc23_unimod(st[n + p] + v[1] * dt, Ly);
                    orclg_boris(dt, q_m, f, f + 3, v);
                    wk[2 * n + p] = v[0];
                    wk[3 * n + p] = v[1];
                    wk[4 * n + p] = v[2];
                }
                st += 5 * n;
                wk += 5 * n;
            }
            for (int64_t e = 0; e < 2 * G; ++e) src[e] = 0.0;
            const int r = deposit_buffered(NX, NY, buffer, Lx, Ly, ns, shape, P, qw, work, 1, grid, pp);
            if (r) {
                rc = r;
                break;
            }
            conv = orc_isapprox(prev, grid, 4 * Ge, rtol, 0.0);
            orclg_si_solve(NX, NY, Lx, Ly, dt, theta_f, theta_s, F, src, plus);
            orclg_update(NX, NY, F, B0, EB);
        }
        if (rc) break;
        if (iters) iters[t - t0] = k;
        memcpy(state, work, sizeof(double) * 5 * (size_t)Ptot);
        static const int x0s[4] = {ORCLG_PHI0, ORCLG_AX0, ORCLG_AY0, ORCLG_AZ0};
        static const int xms[4] = {ORCLG_PHIM, ORCLG_AXM, ORCLG_AYM, ORCLG_AZM};
        for (int a = 0; a < 4; ++a) {
            memcpy(F + (size_t)xms[a] * 2 * G, F + (size_t)x0s[a] * 2 * G, sizeof(double) * 2 * (size_t)G);
            memcpy(F + (size_t)x0s[a] * 2 * G, plus + (size_t)a * 2 * G, sizeof(double) * 2 * (size_t)G);
        }
        if (diagnose)
            lg_diagnose(NX, NY, buffer, ntskip, ngskip, F, EB, ns, P, mass, weight, state, t, rows, &nrow, bins,
                        nbins);
    }
    free(work);
    free(src);
    free(plus);
    free(prev);
    return rc;
}
