/*
 * em1d.c -- CPU FP64 restatement of the 1D electromagnetic fixed-point PIC
 * step of src/GaussianEMFixedPoint.jl: Gaussian +-6 stencil centred on the
 * wrapped cell, exact signed charge and current sums, the spectral
 * Gauss/Ampere/Faraday updates of Ex, Ey, Ez, By, Bz with c = 1/eps, the
 * implicit-midpoint velocity update, up to three fixed-point sweeps tested on
 * Ex, the wrap, the per-step sort by x and the twelve moment fields; and the
 * explicit step of src/GaussianEMExplicit.jl on the same operators (Gauss,
 * leapfrog Faraday then Ampere, the split-kick velocity update, the current
 * at (X + x)/2 with X left unsorted).
 *
 * TEST INFRASTRUCTURE ONLY (see picoracle.h): the oracle the GPU EM path is
 * checked against.
 *
 * Pinned by (tests/test_oracle_em1d.py):
 *   orc_em_stencil    hand cases of irn = Int(mod1(round(c*N), N)), including
 *                     the wrapped-centre cases (c*N < 1/2 and c > 1 get no
 *                     weight), partition of unity elsewhere, scipy's erf.
 *   orc_em_deposit    exact rational sums; total = w * sum(u).
 *   orc_em_fields     single-mode closed forms of each spectral update (Gauss,
 *                     Ampere with the current, Faraday), against numpy.fft.
 *   orc_em_push       closed forms: v' = u + (a + (u + a dt/2) x b) dt.
 *   orc_em_run        the vacuum light wave's amplification matrix, derived per
 *                     mode; the script's Fz .*= Ez (Ez restarts from the
 *                     current term each step); one sweep when Ex does not move;
 *                     the moment fields equal deposits of the final state; the
 *                     sort equals numpy's stable argsort; two calls equal one.
 *   orc_emx_*         single-mode closed forms of Gauss and both curls (numpy
 *                     FFT), the split kick's closed form in complex notation
 *                     (E kick, B turn by (1 - i b dt), E kick), the vacuum
 *                     leapfrog's per-mode matrix, the unsorted-X pairing of
 *                     the current against a numpy restatement of the step,
 *                     and two calls equal one.
 */
#include "picoracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef __int128 i128;
typedef unsigned __int128 u128;

static const double kTwoPiEM = 6.283185307179586; /* Julia 2pi = 2*Float64(pi) */

/* irn(c) = Int(mod1(round(c*N), N)) (:8): round ties to even, then mod1 on the
 * Float64 result, 1..N. */
int64_t orc_em_irn(double c, int32_t N) {
    const double r = rint(c * (double)N);
    if (!(r > -4.0e18 && r < 4.0e18)) return N;
    int64_t m = (int64_t)r % N;
    if (m < 0) m += N;
    return m == 0 ? N : m;
}

static double f_edge(double g, double c, double Nd) { return orc_erf((g - c) * Nd) / 2.0; }

/* s(c) (:9-10): cells mod1(i, N) for i in irn(c)-6 .. irn(c)+6 (written
 * 0-based) with ff(i, c) = f((i+0.5)/N, c) - f((i-0.5)/N, c),
 * f(g, c) = erf((g - c)*N)/2.  The centre is the wrapped cell while c is not
 * wrapped, so a particle with round(c*N) = 0 or N+1 (c*N < 1/2, or c past 1
 * during a sweep) is evaluated 63 cells away and gets zero weight. */
void orc_em_stencil(int32_t N, double c, int64_t* cells, double* ff) {
    const double Nd = (double)N;
    const int64_t i0 = orc_em_irn(c, N);
    for (int k = 0; k < 13; ++k) {
        const int64_t i = i0 - 6 + k;
        ff[k] = f_edge(((double)i + 0.5) / Nd, c, Nd) - f_edge(((double)i - 0.5) / Nd, c, Nd);
        int64_t m = i % N;
        if (m <= 0) m += N;
        cells[k] = m - 1;
    }
}

/* Fraction bits of an exact signed deposit of P contributions bounded by
 * |w|*M: P*|w|*M*2^q < 2^111. */
int orc_em_frac_bits(int64_t P, double w, double M) {
    const double total = (double)P * (fabs(w) * M);
    if (!(total > 0.0) || !isfinite(total)) return 64;
    const int q = 110 - ilogb(total);
    return q > 1100 ? 1100 : q;
}

static i128 fix_s(double a, int q) {
    uint64_t lo, hi;
    orc_fxp_from_double(fabs(a), q, &lo, &hi);
    const i128 v = (i128)(((u128)hi << 64) | lo);
    return a < 0.0 ? -v : v;
}

static double unfix_s(i128 v, int q) {
    if (v < 0) {
        const u128 u = (u128)(-v);
        return -orc_fxp_to_double((uint64_t)u, (uint64_t)(u >> 64), q);
    }
    return orc_fxp_to_double((uint64_t)v, (uint64_t)((u128)v >> 64), q);
}

/* d(y, u) (:11): r = 0; r[cell] += (u[j]*ff)*w over every particle's stencil
 * at y[j] (u == NULL: u = 1).  Per-cell sums are exact: contributions rounded
 * to 2^-q, q = orc_em_frac_bits(P, w, max|u|), summed in 128 bits, rounded
 * once. */
void orc_em_deposit(int32_t N, int64_t P, const double* y, const double* u, double w, double* rho) {
    double M = 1.0;
    if (u) {
        M = 0.0;
        for (int64_t j = 0; j < P; ++j) M = fabs(u[j]) > M ? fabs(u[j]) : M;
    }
    const int q = orc_em_frac_bits(P, w, M);
    i128* acc = (i128*)calloc((size_t)N, sizeof(i128));
    for (int64_t j = 0; j < P; ++j) {
        int64_t cells[13];
        double ff[13];
        orc_em_stencil(N, y[j], cells, ff);
        for (int k = 0; k < 13; ++k) {
            const double a = (u ? u[j] * ff[k] : ff[k]) * w;
            acc[cells[k]] += fix_s(a, q);
        }
    }
    for (int32_t k = 0; k < N; ++k) rho[k] = unfix_s(acc[k], q);
    free(acc);
}

/* ik = 2pi*im*vcat(0, 1:N/2, -N/2+1:-1) (:8) as Julia builds it:
 * Complex(0.0, 2pi) * m = (0.0*m, 2pi*m). */
static void ik_of(int32_t N, int32_t j, double* re, double* im) {
    const double m = (double)(j <= N / 2 ? j : j - N);
    *re = 0.0 * m;
    *im = kTwoPiEM * m;
}

static void fft_real(int32_t N, const double* a, double* re, double* im) {
    for (int32_t j = 0; j < N; ++j) {
        re[j] = a[j];
        im[j] = 0.0;
    }
    orc_fft(N, re, im, 0);
}

/* rif(x) = real.(ifft(x)): the unnormalised inverse, then 1/N. */
static void rif(int32_t N, double* re, double* im, double* out) {
    orc_fft(N, re, im, 1);
    const double s = 1.0 / (double)N;
    for (int32_t j = 0; j < N; ++j) out[j] = re[j] * s;
}

/* Gauss (:20): Ex = real(ifft((xi = fft(rho) ./ ik; xi[1] *= false; xi))).
 * fft(rho)[1]/ik[1] = (0,0)/(0,0) is NaN, and `*= false` makes it
 * copysign(0, NaN): x86's default NaN is negative, so xi[1] = (-0.0, -0.0).
 * The other ratios are Julia's robust division by (0.0*m, 2pi*m), which
 * reduces to t = 1/(2pi*m), (Im*t, -(Re*t)).  re, im: N scratch each. */
static void gauss(int32_t N, const double* rho, double* re, double* im, double* Ex) {
    fft_real(N, rho, re, im);
    for (int32_t j = 0; j < N; ++j) {
        if (j == 0) {
            re[0] = -0.0;
            im[0] = -0.0;
            continue;
        }
        double kr, ki;
        ik_of(N, j, &kr, &ki);
        const double t = 1.0 / ki;
        const double xr = im[j] * t, xi = -(re[j] * t);
        re[j] = xr;
        im[j] = xi;
    }
    rif(N, re, im, Ex);
}

/* One sweep's field updates (:20-27), given rho, Jy, Jz at the sweep's
 * midpoints and the saved Fx, Fy, Fz, Cy, Cz:
 *   Ex = real(ifft((xi = fft(rho) ./ ik; xi[1] *= false; xi)))
 *   by = (By + Cy)/2, bz = (Bz + Cz)/2
 *   Ey = Fy + dt*rif(-ik .* fft(bz)/eps^2 .- fft(Jy))
 *   Ez = Fz + dt*rif(+ik .* fft(by)/eps^2 .- fft(Jz))
 *   ex, ey, ez = (E + F)/2
 *   By = Cy - dt*rif(-ik .* fft(ez)),  Bz = Cz - dt*rif(ik .* fft(ey))
 *   by, bz = (B + C)/2
 * Complex products are Julia's (ac - bd, ad + bc), complex/real divides both
 * parts; Gauss is gauss() above.  Outputs: E[3N] (Ex, Ey, Ez),
 * B[2N] (By, Bz), e3[3N] (ex, ey, ez), b2[2N] (by, bz). */
void orc_em_fields(int32_t N, double eps, double dt, const double* rho, const double* Jy, const double* Jz,
                   const double* Fy, const double* Fz, const double* Cy, const double* Cz, double* E, double* B,
                   double* e3, double* b2) {
    const size_t n = (size_t)N;
    double* re = (double*)malloc(sizeof(double) * n);
    double* im = (double*)malloc(sizeof(double) * n);
    double* jr = (double*)malloc(sizeof(double) * n);
    double* ji = (double*)malloc(sizeof(double) * n);
    double* tmp = (double*)malloc(sizeof(double) * n);
    double *Ex = E, *Ey = E + N, *Ez = E + 2 * N, *By = B, *Bz = B + N;
    double *ex = e3, *ey = e3 + N, *ez = e3 + 2 * N, *by = b2, *bz = b2 + N;
    const double e2 = eps * eps;
    gauss(N, rho, re, im, Ex);
    /* the averaged B of the current iterate */
    for (int32_t j = 0; j < N; ++j) {
        by[j] = (By[j] + Cy[j]) / 2.0;
        bz[j] = (Bz[j] + Cz[j]) / 2.0;
    }
    /* Ampere, y and z */
    for (int comp = 0; comp < 2; ++comp) {
        fft_real(N, comp == 0 ? bz : by, re, im);
        fft_real(N, comp == 0 ? Jy : Jz, jr, ji);
        for (int32_t j = 0; j < N; ++j) {
            double kr, ki;
            ik_of(N, j, &kr, &ki);
            if (comp == 0) {
                kr = -kr;
                ki = -ki;
            }
            const double pr = kr * re[j] - ki * im[j];
            const double pi = kr * im[j] + ki * re[j];
            re[j] = pr / e2 - jr[j];
            im[j] = pi / e2 - ji[j];
        }
        rif(N, re, im, tmp);
        double* out = comp == 0 ? Ey : Ez;
        const double* F = comp == 0 ? Fy : Fz;
        for (int32_t j = 0; j < N; ++j) out[j] = F[j] + dt * tmp[j];
    }
    for (int32_t j = 0; j < N; ++j) ey[j] = (Ey[j] + Fy[j]) / 2.0;
    for (int32_t j = 0; j < N; ++j) ez[j] = (Ez[j] + Fz[j]) / 2.0;
    /* Faraday, y and z (ex needs Fx, which the caller averages) */
    for (int comp = 0; comp < 2; ++comp) {
        fft_real(N, comp == 0 ? ez : ey, re, im);
        for (int32_t j = 0; j < N; ++j) {
            double kr, ki;
            ik_of(N, j, &kr, &ki);
            if (comp == 0) {
                kr = -kr;
                ki = -ki;
            }
            const double pr = kr * re[j] - ki * im[j];
            const double pi = kr * im[j] + ki * re[j];
            re[j] = pr;
            im[j] = pi;
        }
        rif(N, re, im, tmp);
        double* out = comp == 0 ? By : Bz;
        const double* C = comp == 0 ? Cy : Cz;
        for (int32_t j = 0; j < N; ++j) out[j] = C[j] - dt * tmp[j];
    }
    for (int32_t j = 0; j < N; ++j) {
        by[j] = (By[j] + Cy[j]) / 2.0;
        bz[j] = (Bz[j] + Cz[j]) / 2.0;
    }
    (void)ex;
    free(re); free(im); free(jr); free(ji); free(tmp);
}

/* The velocity update of :28-29 for one particle: theta = s((x + X)/2),
 * a = sum(k -> [ex, ey, ez][k]*ff_k), b = sum(k -> [Bx, by, bz][k]*ff_k)
 * (left folds from the first term; Bx is the script's zero array),
 * v = u + (a + cross(u + a*dt/2, b))*dt with u = V. */
void orc_em_push(int32_t N, double c, const double* ex, const double* ey, const double* ez, const double* by,
                 const double* bz, double dt, const double* u, double* v) {
    int64_t cells[13];
    double ff[13];
    orc_em_stencil(N, c, cells, ff);
    double a1 = 0, a2 = 0, a3 = 0, b1 = 0, b2 = 0, b3 = 0;
    for (int k = 0; k < 13; ++k) {
        const int64_t m = cells[k];
        const double t1 = ex[m] * ff[k], t2 = ey[m] * ff[k], t3 = ez[m] * ff[k];
        const double s1 = 0.0 * ff[k], s2 = by[m] * ff[k], s3 = bz[m] * ff[k];
        if (k == 0) {
            a1 = t1; a2 = t2; a3 = t3;
            b1 = s1; b2 = s2; b3 = s3;
        } else {
            a1 = a1 + t1; a2 = a2 + t2; a3 = a3 + t3;
            b1 = b1 + s1; b2 = b2 + s2; b3 = b3 + s3;
        }
    }
    const double w1 = u[0] + a1 * dt / 2.0, w2 = u[1] + a2 * dt / 2.0, w3 = u[2] + a3 * dt / 2.0;
    const double c1 = w2 * b3 - w3 * b2, c2 = w3 * b1 - w1 * b3, c3 = w1 * b2 - w2 * b1;
    v[0] = u[0] + (a1 + c1) * dt;
    v[1] = u[1] + (a2 + c2) * dt;
    v[2] = u[2] + (a3 + c3) * dt;
}

typedef struct {
    uint64_t k;
    int64_t i;
} EmKey;

static int cmp_em(const void* pa, const void* pb) {
    const EmKey* a = (const EmKey*)pa;
    const EmKey* b = (const EmKey*)pb;
    if (a->k != b->k) return a->k < b->k ? -1 : 1;
    return a->i < b->i ? -1 : (a->i > b->i);
}

/* p = sortperm(x); x = x[p]; v = v[p, :] (:41): stable, isless order. */
void orc_em_sort(int64_t P, double* x, double* v1, double* v2, double* v3, int64_t* perm) {
    EmKey* k = (EmKey*)malloc(sizeof(EmKey) * (size_t)(P > 0 ? P : 1));
    for (int64_t j = 0; j < P; ++j) {
        k[j].k = orc23_order_key(x[j]);
        k[j].i = j;
    }
    qsort(k, (size_t)P, sizeof(EmKey), cmp_em);
    double* t = (double*)malloc(sizeof(double) * (size_t)(P > 0 ? P : 1));
    double* cols[4] = {x, v1, v2, v3};
    for (int c = 0; c < 4; ++c) {
        for (int64_t j = 0; j < P; ++j) t[j] = cols[c][k[j].i];
        memcpy(cols[c], t, sizeof(double) * (size_t)P);
    }
    if (perm)
        for (int64_t j = 0; j < P; ++j) perm[j] = k[j].i;
    free(t);
    free(k);
}

/* nsteps of the loop of :13-50 from the state (x, v = (v1, v2, v3), E = (Ex,
 * Ey, Ez), B = (By, Bz), Fz).  fields (may be NULL) receives, for each step t
 * of the call, fields[:, t, 1:12] (N x nsteps x 12, Julia's layout): Ex, Ey,
 * Ez, By, Bz, d(x), d(x, v1), d(x, v2), d(x, v3), d(x, v1.^2), d(x, v2.^2),
 * d(x, v3.^2) at the wrapped, sorted end-of-step state.  sweeps (may be NULL)
 * receives the sweeps of each step.  sort != 0 sorts the particles by x after
 * every step as the script does. */
int orc_em_run(int32_t N, int64_t P, double eps, double dt, double w, double rtol, int32_t max_sweeps, int32_t sort,
               double* x, double* v1, double* v2, double* v3, double* E, double* B, double* Fz, int64_t nsteps,
               double* fields, int32_t* sweeps) {
    if (N < 8 || (N & (N - 1)) || P < 1 || max_sweeps < 1) return -1;
    const size_t n = (size_t)N, np = (size_t)P;
    double *X = malloc(sizeof(double) * np), *V1 = malloc(sizeof(double) * np), *V2 = malloc(sizeof(double) * np),
           *V3 = malloc(sizeof(double) * np), *c = malloc(sizeof(double) * np), *uy = malloc(sizeof(double) * np),
           *uz = malloc(sizeof(double) * np), *sq = malloc(sizeof(double) * np);
    double *rho = malloc(sizeof(double) * n), *Jy = malloc(sizeof(double) * n), *Jz = malloc(sizeof(double) * n),
           *Fx = malloc(sizeof(double) * n), *Fy = malloc(sizeof(double) * n), *Cy = malloc(sizeof(double) * n),
           *Cz = malloc(sizeof(double) * n), *e3 = malloc(sizeof(double) * 3 * n),
           *b2 = malloc(sizeof(double) * 2 * n);
    double *Ex = E, *Ey = E + N, *Ez = E + 2 * N, *By = B, *Bz = B + N;
    for (int64_t t = 0; t < nsteps; ++t) {
        memcpy(X, x, sizeof(double) * np);
        memcpy(V1, v1, sizeof(double) * np);
        memcpy(V2, v2, sizeof(double) * np);
        memcpy(V3, v3, sizeof(double) * np);
        /* Fx .*= NaN (all NaN: the first test never passes); Fy .= Ey;
         * Fz .*= Ez (the script multiplies where it means to copy, so Fz stays
         * a zero of either sign); Cy .= By; Cz .= Bz */
        for (int32_t j = 0; j < N; ++j) {
            Fx[j] = NAN;
            Fy[j] = Ey[j];
            Fz[j] = Fz[j] * Ez[j];
            Cy[j] = By[j];
            Cz[j] = Bz[j];
        }
        int s = 0;
        for (; s < max_sweeps; ++s) {
            if (orc_isapprox(Fx, Ex, N, rtol, 0.0)) break;
            memcpy(Fx, Ex, sizeof(double) * n);
            for (int64_t j = 0; j < P; ++j) {
                x[j] = X[j] + (V1[j] + v1[j]) / 2.0 * dt;
                c[j] = (X[j] + x[j]) / 2.0;
                uy[j] = (V2[j] + v2[j]) / 2.0;
                uz[j] = (V3[j] + v3[j]) / 2.0;
            }
            orc_em_deposit(N, P, c, NULL, w, rho);
            orc_em_deposit(N, P, c, uy, w, Jy);
            orc_em_deposit(N, P, c, uz, w, Jz);
            orc_em_fields(N, eps, dt, rho, Jy, Jz, Fy, Fz, Cy, Cz, E, B, e3, b2);
            for (int32_t j = 0; j < N; ++j) e3[j] = (Ex[j] + Fx[j]) / 2.0;
            for (int64_t j = 0; j < P; ++j) {
                const double u[3] = {V1[j], V2[j], V3[j]};
                double vv[3];
                orc_em_push(N, (x[j] + X[j]) / 2.0, e3, e3 + N, e3 + 2 * N, b2, b2 + N, dt, u, vv);
                v1[j] = vv[0];
                v2[j] = vv[1];
                v3[j] = vv[2];
            }
        }
        if (sweeps) sweeps[t] = s;
        for (int64_t j = 0; j < P; ++j) x[j] = orc_jl_mod1f(x[j]);
        if (sort) orc_em_sort(P, x, v1, v2, v3, NULL);
        if (fields) {
            double* o = fields + (size_t)t * n;
            const size_t stride = (size_t)nsteps * n;
            memcpy(o, Ex, sizeof(double) * n);
            memcpy(o + stride, Ey, sizeof(double) * n);
            memcpy(o + 2 * stride, Ez, sizeof(double) * n);
            memcpy(o + 3 * stride, By, sizeof(double) * n);
            memcpy(o + 4 * stride, Bz, sizeof(double) * n);
            orc_em_deposit(N, P, x, NULL, w, o + 5 * stride);
            const double* vs[3] = {v1, v2, v3};
            for (int k = 0; k < 3; ++k) orc_em_deposit(N, P, x, vs[k], w, o + (6 + k) * stride);
            for (int k = 0; k < 3; ++k) {
                for (int64_t j = 0; j < P; ++j) sq[j] = vs[k][j] * vs[k][j];
                orc_em_deposit(N, P, x, sq, w, o + (9 + k) * stride);
            }
        }
    }
    free(X); free(V1); free(V2); free(V3); free(c); free(uy); free(uz); free(sq);
    free(rho); free(Jy); free(Jz); free(Fx); free(Fy); free(Cy); free(Cz); free(e3); free(b2);
    return 0;
}


/* ------------------------------------------------------------------------ */
/* The explicit script, src/GaussianEMExplicit.jl:12-30.  One step:          */
/*   X .= x; Ex .= Gauss(d(X))                                          :13  */
/*   x .= X .+ v[:,1]*dt; x .= mod.(x, 1); p = sortperm(x); x = x[p];        */
/*   v = v[p, :]                                                        :14  */
/*   Cy .= By; Cz .= Bz; By .-= dt.*rif(-ik.*fft(Ez));                        */
/*   Bz .-= dt.*rif(ik.*fft(Ey)); by, bz = (B .+ C)./2                :15-16 */
/*   v[j] += a*dt/2; v[j] += cross(v[j], b)*dt; v[j] += a*dt/2 at s(x[j]) :17-22 */
/*   Ey .+= dt.*rif(-ik.*fft(Bz)/eps^2 .- fft(d((X.+x)/2, v[:,2])))           */
/*   Ez .+= dt.*rif(+ik.*fft(By)/eps^2 .- fft(d((X.+x)/2, v[:,3])))    :23-24 */
/*   fields[:, t, 1:12] as in the fixed point                         :25-29 */
/* X is not permuted with x: (X .+ x)/2 pairs the pre-sort position of one   */
/* particle with the sorted position of another, and the script does so.     */
/* ------------------------------------------------------------------------ */

/* out = rif(sign*ik .* fft(a)) (sign < 0: -ik), or, with J, rif(sign*ik .*
 * fft(a)/e2 .- fft(J)) */
static void curl_term(int32_t N, int sign, const double* a, double e2, const double* J, double* out) {
    const size_t n = (size_t)N;
    double* re = (double*)malloc(sizeof(double) * n);
    double* im = (double*)malloc(sizeof(double) * n);
    double* jr = (double*)malloc(sizeof(double) * n);
    double* ji = (double*)malloc(sizeof(double) * n);
    fft_real(N, a, re, im);
    if (J) fft_real(N, J, jr, ji);
    for (int32_t j = 0; j < N; ++j) {
        double kr, ki;
        ik_of(N, j, &kr, &ki);
        if (sign < 0) {
            kr = -kr;
            ki = -ki;
        }
        double pr = kr * re[j] - ki * im[j];
        double pi = kr * im[j] + ki * re[j];
        if (J) {
            pr = pr / e2 - jr[j];
            pi = pi / e2 - ji[j];
        }
        re[j] = pr;
        im[j] = pi;
    }
    rif(N, re, im, out);
    free(re); free(im); free(jr); free(ji);
}

/* :13 and :15-16: Ex from rho; By, Bz from the old Ey, Ez; b2 = (by, bz). */
void orc_emx_gauss_faraday(int32_t N, double dt, const double* rho, double* E, double* B, double* b2) {
    const size_t n = (size_t)N;
    double* re = (double*)malloc(sizeof(double) * n);
    double* im = (double*)malloc(sizeof(double) * n);
    double* ty = (double*)malloc(sizeof(double) * n);
    double* tz = (double*)malloc(sizeof(double) * n);
    double *Ex = E, *Ey = E + N, *Ez = E + 2 * N, *By = B, *Bz = B + N;
    gauss(N, rho, re, im, Ex);
    curl_term(N, -1, Ez, 0.0, NULL, ty);
    curl_term(N, +1, Ey, 0.0, NULL, tz);
    for (int32_t j = 0; j < N; ++j) {
        const double cy = By[j], cz = Bz[j];
        By[j] = cy - dt * ty[j];
        Bz[j] = cz - dt * tz[j];
        b2[j] = (By[j] + cy) / 2.0;
        b2[N + j] = (Bz[j] + cz) / 2.0;
    }
    free(re); free(im); free(ty); free(tz);
}

/* :23-24: Ey, Ez from the new By, Bz and the currents Jy, Jz. */
void orc_emx_ampere(int32_t N, double eps, double dt, const double* Jy, const double* Jz, double* E, const double* B) {
    const size_t n = (size_t)N;
    double* ty = (double*)malloc(sizeof(double) * n);
    double* tz = (double*)malloc(sizeof(double) * n);
    const double e2 = eps * eps;
    curl_term(N, -1, B + N, e2, Jy, ty);
    curl_term(N, +1, B, e2, Jz, tz);
    for (int32_t j = 0; j < N; ++j) {
        E[N + j] = E[N + j] + dt * ty[j];
        E[2 * N + j] = E[2 * N + j] + dt * tz[j];
    }
    free(ty); free(tz);
}

/* :17-22 for one particle at x: a and b are left folds over the stencil from
 * its first term (Bx is the script's zero array); then
 * v += a*dt/2; v += cross(v, b)*dt; v += a*dt/2. */
void orc_emx_push(int32_t N, double x, const double* E, const double* b2, double dt, double* v) {
    int64_t cells[13];
    double ff[13];
    orc_em_stencil(N, x, cells, ff);
    const double *Ex = E, *Ey = E + N, *Ez = E + 2 * N, *by = b2, *bz = b2 + N;
    double a1 = 0, a2 = 0, a3 = 0, b1 = 0, b2y = 0, b3 = 0;
    for (int k = 0; k < 13; ++k) {
        const int64_t m = cells[k];
        const double t1 = Ex[m] * ff[k], t2 = Ey[m] * ff[k], t3 = Ez[m] * ff[k];
        const double s1 = 0.0 * ff[k], s2 = by[m] * ff[k], s3 = bz[m] * ff[k];
        if (k == 0) {
            a1 = t1; a2 = t2; a3 = t3;
            b1 = s1; b2y = s2; b3 = s3;
        } else {
            a1 = a1 + t1; a2 = a2 + t2; a3 = a3 + t3;
            b1 = b1 + s1; b2y = b2y + s2; b3 = b3 + s3;
        }
    }
    const double u1 = v[0] + a1 * dt / 2.0, u2 = v[1] + a2 * dt / 2.0, u3 = v[2] + a3 * dt / 2.0;
    const double c1 = u2 * b3 - u3 * b2y, c2 = u3 * b1 - u1 * b3, c3 = u1 * b2y - u2 * b1;
    const double w1 = u1 + c1 * dt, w2 = u2 + c2 * dt, w3 = u3 + c3 * dt;
    v[0] = w1 + a1 * dt / 2.0;
    v[1] = w2 + a2 * dt / 2.0;
    v[2] = w3 + a3 * dt / 2.0;
}

int orc_emx_run(int32_t N, int64_t P, double eps, double dt, double w, int32_t sort, double* x, double* v1,
                double* v2, double* v3, double* E, double* B, int64_t nsteps, double* fields) {
    if (N < 8 || (N & (N - 1)) || P < 1) return -1;
    const size_t n = (size_t)N, np = (size_t)P;
    double* X = (double*)malloc(sizeof(double) * np);
    double* c = (double*)malloc(sizeof(double) * np);
    double* sq = (double*)malloc(sizeof(double) * np);
    double* rho = (double*)malloc(sizeof(double) * n);
    double* Jy = (double*)malloc(sizeof(double) * n);
    double* Jz = (double*)malloc(sizeof(double) * n);
    double* b2 = (double*)malloc(sizeof(double) * 2 * n);
    for (int64_t t = 0; t < nsteps; ++t) {
        memcpy(X, x, sizeof(double) * np);
        orc_em_deposit(N, P, X, NULL, w, rho);
        for (int64_t j = 0; j < P; ++j) x[j] = orc_jl_mod1f(X[j] + v1[j] * dt);
        if (sort) orc_em_sort(P, x, v1, v2, v3, NULL);
        orc_emx_gauss_faraday(N, dt, rho, E, B, b2);
        for (int64_t j = 0; j < P; ++j) {
            double vv[3] = {v1[j], v2[j], v3[j]};
            orc_emx_push(N, x[j], E, b2, dt, vv);
            v1[j] = vv[0];
            v2[j] = vv[1];
            v3[j] = vv[2];
        }
        for (int64_t j = 0; j < P; ++j) c[j] = (X[j] + x[j]) / 2.0;
        orc_em_deposit(N, P, c, v2, w, Jy);
        orc_em_deposit(N, P, c, v3, w, Jz);
        orc_emx_ampere(N, eps, dt, Jy, Jz, E, B);
        if (fields) {
            double* o = fields + (size_t)t * n;
            const size_t stride = (size_t)nsteps * n;
            memcpy(o, E, sizeof(double) * n);
            memcpy(o + stride, E + N, sizeof(double) * n);
            memcpy(o + 2 * stride, E + 2 * N, sizeof(double) * n);
            memcpy(o + 3 * stride, B, sizeof(double) * n);
            memcpy(o + 4 * stride, B + N, sizeof(double) * n);
            orc_em_deposit(N, P, x, NULL, w, o + 5 * stride);
            const double* vs[3] = {v1, v2, v3};
            for (int k = 0; k < 3; ++k) orc_em_deposit(N, P, x, vs[k], w, o + (6 + k) * stride);
            for (int k = 0; k < 3; ++k) {
                for (int64_t j = 0; j < P; ++j) sq[j] = vs[k][j] * vs[k][j];
                orc_em_deposit(N, P, x, sq, w, o + (9 + k) * stride);
            }
        }
    }
    free(X); free(c); free(sq); free(rho); free(Jy); free(Jz); free(b2);
    return 0;
}
