/*
 * pic2d3v.c -- CPU FP64 restatement of the electrostatic path of the
 * structured 2D3V module src/PIC2D3V.jl: species of any charge and mass, an
 * Lx x Ly box, a uniform magnetic field of any direction, the NGP, area and
 * B-spline (orders 0..5) shapes, loop!(plasma, ::ElectrostaticField),
 * diagnose!(::ElectrostaticDiagnostics) and sort!(::Species, dx, dy).
 *
 * TEST INFRASTRUCTURE ONLY (see picoracle.h): the oracle the GPU PIC2D3V path
 * is checked against.
 *
 * Grids are column-major NX x NY (x fastest).  Cells are Julia indices
 * (1-based, unwrapped, as depositindicesfractions returns them); the
 * reference reads and writes them in arrays with `buffer` ghost layers
 * (:284-292) that the reduction and update fold periodically (:11-27), which
 * is the same as wrapping the index: cell c lives at (c - 1) mod N.
 *
 * Pinned by (tests/test_oracle_pic2d3v.py):
 *   orc23_bspline      exact cardinal B-splines (Cox-de Boor in rationals) at
 *                      dyadic offsets, partition of unity, the reference's
 *                      polynomial coefficients term by term for orders 0, 1.
 *   orc23_weights      hand cases of the index rule of every shape (the
 *                      r > 1/2 switch of even orders), linear reproduction.
 *   orc23_gather       constant and linear fields (linear reproduction of the
 *                      B-splines of order >= 1), brute force on hand grids.
 *   orc23_boris        both variants: E = 0 closed forms (rotation about B,
 *                      the reference's q_m^2 scaling and its energy change),
 *                      B = 0 plain kicks.
 *   orc23_deposit      exact rational sums of the contributions; charge total;
 *                      first moment (B-splines of order >= 1).
 *   orc23_fieldsolve   single-mode closed forms in an Lx x Ly box; equals the
 *                      2D solve of src/Electrostatic2D3V.jl at Lx = Ly = 1.
 *   orc23_sort         lexicographic order of (ceil(x/dx), ceil(y/dy), vx)
 *                      with ties by index, against numpy's stable lexsort.
 *   orc23_diagnostics  closed forms on hand states.
 *   orc23_run          momentum conservation (one species, area shape),
 *                      plasma frequency, accumulate vs replace field updates.
 */
#include "picoracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef __int128 i128;
typedef unsigned __int128 u128;

static const double kTwoPi23 = 6.283185307179586; /* Julia 2pi = 2*Float64(pi) */

/* unimod(x, n) = x > n ? x - n : x > 0 ? x : x + n   (src/PIC2D3V.jl:9) */
double orc23_unimod(double x, double n) {
    return x > n ? x - n : (x > 0.0 ? x : x + n);
}

/* bspline(BSplineWeighting{n}(), x) (:1124-1162).  The reference evaluates
 * the polynomials under @fastmath, whose reassociation is unpinned; this
 * restatement fixes the order as written, left to right, with integer powers
 * as LLVM's powi expands them (x^2 = x*x, x^3 = x*(x*x), x^4 = (x*x)*(x*x),
 * x^5 = x*((x*x)*(x*x))) and every coefficient a Float64 division as Julia
 * folds it (9/8, 5/24, ...).  Returns n + 1 weights. */
static double p2(double t) { return t * t; }
static double p3(double t) { return t * (t * t); }
static double p4(double t) { double s = t * t; return s * s; }
static double p5(double t) { double s = t * t; return t * (s * s); }

int orc23_bspline(int32_t n, double x, double* w) {
    switch (n) {
    case 0:
        w[0] = 1.0;
        return 1;
    case 1:
        w[0] = x;
        w[1] = 1.0 - x;
        return 2;
    case 2: {
        const double a = x - 1.5, b = x - 0.5, c = x + 0.5;
        w[0] = 9.0 / 8.0 + 3.0 / 2.0 * a + 1.0 / 2.0 * p2(a);
        w[1] = 3.0 / 4.0 - p2(b);
        w[2] = 9.0 / 8.0 - 3.0 / 2.0 * c + 1.0 / 2.0 * p2(c);
        return 3;
    }
    case 3: {
        const double a = x - 2.0, b = x - 1.0, c = x, d = x + 1.0;
        w[0] = 4.0 / 3.0 + 2.0 * a + p2(a) + 1.0 / 6.0 * p3(a);
        w[1] = 2.0 / 3.0 - p2(b) - 1.0 / 2.0 * p3(b);
        w[2] = 2.0 / 3.0 - p2(c) + 1.0 / 2.0 * p3(c);
        w[3] = 4.0 / 3.0 - 2.0 * d + p2(d) - 1.0 / 6.0 * p3(d);
        return 4;
    }
    case 4: {
        const double a = x - 2.5, b = x - 1.5, c = x - 0.5, d = x + 0.5, e = x + 1.5;
        w[0] = 625.0 / 384.0 + 125.0 / 48.0 * a + 25.0 / 16.0 * p2(a) + 5.0 / 12.0 * p3(a) + 1.0 / 24.0 * p4(a);
        w[1] = 55.0 / 96.0 - 5.0 / 24.0 * b - 5.0 / 4.0 * p2(b) - 5.0 / 6.0 * p3(b) - 1.0 / 6.0 * p4(b);
        w[2] = 115.0 / 192.0 - 5.0 / 8.0 * p2(c) + 1.0 / 4.0 * p4(c);
        w[3] = 55.0 / 96.0 + 5.0 / 24.0 * d - 5.0 / 4.0 * p2(d) + 5.0 / 6.0 * p3(d) - 1.0 / 6.0 * p4(d);
        w[4] = 625.0 / 384.0 - 125.0 / 48.0 * e + 25.0 / 16.0 * p2(e) - 5.0 / 12.0 * p3(e) + 1.0 / 24.0 * p4(e);
        return 5;
    }
    case 5: {
        const double a = x - 3.0, b = x - 2.0, c = x - 1.0, d0 = x + 0.0, d1 = x - 0.0, e = x + 1.0, f = x + 2.0;
        w[0] = 243.0 / 120.0 + 81.0 / 24.0 * a + 9.0 / 4.0 * p2(a) + 3.0 / 4.0 * p3(a) + 1.0 / 8.0 * p4(a) +
               1.0 / 120.0 * p5(a);
        w[1] = 17.0 / 40.0 - 5.0 / 8.0 * b - 7.0 / 4.0 * p2(b) - 5.0 / 4.0 * p3(b) - 3.0 / 8.0 * p4(b) -
               1.0 / 24.0 * p5(b);
        w[2] = 22.0 / 40.0 - 1.0 / 2.0 * p2(c) + 1.0 / 4.0 * p4(c) + 1.0 / 12.0 * p5(c);
        w[3] = 22.0 / 40.0 - 1.0 / 2.0 * p2(d0) + 1.0 / 4.0 * p4(d1) - 1.0 / 12.0 * p5(d1);
        w[4] = 17.0 / 40.0 + 5.0 / 8.0 * e - 7.0 / 4.0 * p2(e) + 5.0 / 4.0 * p3(e) - 3.0 / 8.0 * p4(e) +
               1.0 / 24.0 * p5(e);
        w[5] = 243.0 / 120.0 - 81.0 / 24.0 * f + 9.0 / 4.0 * p2(f) - 3.0 / 4.0 * p3(f) + 1.0 / 8.0 * p4(f) -
               1.0 / 120.0 * p5(f);
        return 6;
    }
    }
    return 0;
}

/* depositindicesfractions(s, z, NZ, NZ_Lz) (:1106-1114) and the shape's
 * gridinteractiontuple (:1116-1121, 1165-1187): zN = z*NZ_Lz, i = ceil(zN),
 * r = i - zN;
 *   NGP:        ((i, 1),)
 *   area:       ((i, 1 - r), (i + 1, r))
 *   B-spline n: odd n: j = i, x = 1 - r; even n: q = r > 1/2, j = i + q,
 *               x = (q + 1/2) - r; cells j - fld(n,2) .. j + cld(n,2) with
 *               weights bspline(n, x).
 * Writes Julia (1-based, unwrapped) cell indices; returns the count.  The
 * reference asserts 0 < r <= 1 (:1112), which fails for a particle exactly on
 * a cell edge (r = 0); the weights there are the limits of the formulas. */
int orc23_weights(int32_t shape, double z, double NZ_Lz, int64_t* cell, double* w) {
    const double zN = z * NZ_Lz;
    const double ci = ceil(zN);
    const int64_t i = (ci > -4.0e18 && ci < 4.0e18) ? (int64_t)ci : 0;
    const double r = (double)i - zN;
    if (shape == ORC23_NGP) {
        cell[0] = i;
        w[0] = 1.0;
        return 1;
    }
    if (shape == ORC23_AREA) {
        cell[0] = i;
        cell[1] = i + 1;
        w[0] = 1.0 - r;
        w[1] = r;
        return 2;
    }
    const int n = shape - ORC23_BSPLINE0;
    int64_t j;
    double x;
    if (n & 1) {
        j = i;
        x = 1.0 - r;
    } else {
        const int q = r > 0.5;
        j = i + q;
        x = ((double)q + 0.5) - r;
    }
    const int m = orc23_bspline(n, x, w);
    const int64_t lo = j - n / 2; /* fld(n, 2) for n >= 0 */
    for (int k = 0; k < m; ++k) cell[k] = lo + k;
    return m;
}

static inline int64_t wrap_cell(int64_t c, int32_t N) { return (c - 1) & (int64_t)(N - 1); }

/* (f::ElectrostaticField)(shape, x, y) (:1217-1228): j outer, i inner,
 * wxy = wx*wy, E = muladd(Exy[., i, j], wxy, E) taken as a fused
 * multiply-add.  Ex, Ey are the NX x NY field the ghost layers copy. */
void orc23_gather(int32_t NX, int32_t NY, double NX_Lx, double NY_Ly, int32_t shape, const double* Ex,
                  const double* Ey, double x, double y, double* ex, double* ey) {
    int64_t cx[6], cy[6];
    double wx[6], wy[6];
    const int nx = orc23_weights(shape, x, NX_Lx, cx, wx);
    const int ny = orc23_weights(shape, y, NY_Ly, cy, wy);
    double a = 0.0, b = 0.0;
    for (int jj = 0; jj < ny; ++jj)
        for (int ii = 0; ii < nx; ++ii) {
            const double wxy = wx[ii] * wy[jj];
            const int64_t k = wrap_cell(cx[ii], NX) + (int64_t)NX * wrap_cell(cy[jj], NY);
            a = fma(Ex[k], wxy, a);
            b = fma(Ey[k], wxy, b);
        }
    *ex = a;
    *ey = b;
}

/* Boris push with E = (Ex, Ey, 0) and the uniform B.
 * variant 0: ElectrostaticBoris (:39-54), as the electrostatic loop uses it:
 *   t = (B*dt)/2, t2 = dot(t, t), e = (E*(dt/2))*q_m, m = v + e,
 *   v+ = m + ((cross(m + cross(m, t), t)*q_m^2)*2)/(1 + q_m^2*t2), v = v+ + e.
 *   The rotation scales with q_m^2, not q_m: a negative charge turns the same
 *   way as a positive one, and |v| is kept only for q_m^2 = 1.
 * variant 1: ElectromagneticBoris (:56-69) with Ez = 0, the textbook form:
 *   theta = (dt/2)*q_m, t = B*theta, s = 2/(1 + dot(t, t)), e = E*theta,
 *   m = v + e, v+ = m + cross(m + cross(m, t), t)*s, v = v+ + e.
 * Cross products and dot products as StaticArrays writes them. */
void orc23_boris(int32_t variant, const double* B, double dt, double q_m, double Ex, double Ey, double* vx,
                 double* vy, double* vz) {
    double t1, t2, t3, e1, e2, e3, sc = 0.0, den = 0.0, qm2 = 0.0;
    const double dt_2 = dt / 2.0;
    if (variant == 0) {
        t1 = B[0] * dt / 2.0;
        t2 = B[1] * dt / 2.0;
        t3 = B[2] * dt / 2.0;
        const double tt = t1 * t1 + t2 * t2 + t3 * t3;
        e1 = Ex * dt_2 * q_m;
        e2 = Ey * dt_2 * q_m;
        e3 = 0.0 * dt_2 * q_m;
        qm2 = q_m * q_m;
        den = 1.0 + qm2 * tt;
    } else {
        const double th = dt_2 * q_m;
        t1 = B[0] * th;
        t2 = B[1] * th;
        t3 = B[2] * th;
        sc = 2.0 / (1.0 + (t1 * t1 + t2 * t2 + t3 * t3));
        e1 = Ex * th;
        e2 = Ey * th;
        e3 = 0.0 * th;
    }
    const double m1 = *vx + e1, m2 = *vy + e2, m3 = *vz + e3;
    const double c1 = m2 * t3 - m3 * t2, c2 = m3 * t1 - m1 * t3, c3 = m1 * t2 - m2 * t1;
    const double u1 = m1 + c1, u2 = m2 + c2, u3 = m3 + c3;
    const double d1 = u2 * t3 - u3 * t2, d2 = u3 * t1 - u1 * t3, d3 = u1 * t2 - u2 * t1;
    double p1, pp2, pp3;
    if (variant == 0) {
        p1 = m1 + d1 * qm2 * 2.0 / den;
        pp2 = m2 + d2 * qm2 * 2.0 / den;
        pp3 = m3 + d3 * qm2 * 2.0 / den;
    } else {
        p1 = m1 + d1 * sc;
        pp2 = m2 + d2 * sc;
        pp3 = m3 + d3 * sc;
    }
    *vx = p1 + e1;
    *vy = pp2 + e2;
    *vz = pp3 + e3;
}

/* ------------------------------------------------------------------------ */
/* Exact per-cell charge: each contribution (wx*wy)*w (:1249-1256) is rounded */
/* to the grid 2^-q (ties to even, sign kept), summed exactly and rounded     */
/* once.  q puts the total |charge| of every species below 2^111, so signed   */
/* sums fit the GPU's three 38-bit limbs (two's complement, 114 bits).        */
/* ------------------------------------------------------------------------ */
int orc23_frac_bits(int32_t ns, const int64_t* P, const double* w) {
    double total = 0.0;
    for (int s = 0; s < ns; ++s) total += (double)P[s] * fabs(w[s]);
    if (!(total > 0.0) || !isfinite(total)) return 64;
    const int q = 110 - ilogb(total);
    return q > 1100 ? 1100 : q;
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

static void deposit_into(i128* acc, int32_t NX, int32_t NY, double NX_Lx, double NY_Ly, int32_t shape, int64_t P,
                         const double* x, const double* y, double w, int q) {
    for (int64_t p = 0; p < P; ++p) {
        int64_t cx[6], cy[6];
        double wx[6], wy[6];
        const int nx = orc23_weights(shape, x[p], NX_Lx, cx, wx);
        const int ny = orc23_weights(shape, y[p], NY_Ly, cy, wy);
        for (int jj = 0; jj < ny; ++jj)
            for (int ii = 0; ii < nx; ++ii) {
                const double c = wx[ii] * wy[jj] * w;
                acc[wrap_cell(cx[ii], NX) + (int64_t)NX * wrap_cell(cy[jj], NY)] += fix_signed(c, q);
            }
    }
}

/* deposit!(rho, shape, x, y, NX_Lx, NY_Ly, w) for P particles, folded onto
 * the NX x NY grid as reduction! does (:481-486), q = orc23_frac_bits of this
 * one set. */
void orc23_deposit(int32_t NX, int32_t NY, double NX_Lx, double NY_Ly, int32_t shape, int64_t P, const double* x,
                   const double* y, double w, double* rho) {
    const int64_t G = (int64_t)NX * NY;
    const int q = orc23_frac_bits(1, &P, &w);
    i128* acc = (i128*)calloc((size_t)G, sizeof(i128));
    deposit_into(acc, NX, NY, NX_Lx, NY_Ly, shape, P, x, y, w, q);
    for (int64_t k = 0; k < G; ++k) rho[k] = unfix_signed(acc[k], q);
    free(acc);
}

/* The electrostatic solve of loop! (:559-576) in an Lx x Ly box:
 * kx = (2pi/Lx)*[0:NX/2-1, -NX/2:-1], ky likewise, k2 = kx^2 + ky^2,
 * im_k^-2 = -im ./ k2 with [1,1] = 0 (FFTHelper, :251-268), phi = fft(rho),
 * phi[1,1] = 0, tmp = phi*im_k^-2, Ex = tmp*kx[i], Ey = tmp*ky[j], inverse
 * transforms scaled by 1/(NX*NY).  phi_out (may be NULL) receives
 * real(ifft(phi)) as diagnose! stores in phis (:1326).  Returns 1 if a tmp
 * was not finite (the @assert of :570), else 0. */
int orc23_fieldsolve(int32_t NX, int32_t NY, double Lx, double Ly, const double* rho, double* Ex, double* Ey,
                     double* phi_out) {
    return orc2d_fieldsolve_k(NX, NY, kTwoPi23 / Lx, kTwoPi23 / Ly, rho, Ex, Ey, phi_out);
}

/* ------------------------------------------------------------------------ */
/* sort!(s, dx, dy) (:210-215): the stable permutation ordering particles by */
/* (ceil(x/dx), ceil(y/dy), vx) under isless (-0.0 < 0.0, NaN last), ties by */
/* index; then every column of xyv is permuted.                              */
/* ------------------------------------------------------------------------ */
typedef struct {
    int64_t cx, cy;
    uint64_t v;
    int64_t idx;
} SortKey;

uint64_t orc23_order_key(double v) {
    if (v != v) return UINT64_MAX;
    uint64_t b;
    memcpy(&b, &v, 8);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

int64_t orc23_ceil_index(double a) {
    const double c = ceil(a);
    return (c > -4.0e18 && c < 4.0e18) ? (int64_t)c : 0;
}

static int cmp_key(const void* pa, const void* pb) {
    const SortKey* a = (const SortKey*)pa;
    const SortKey* b = (const SortKey*)pb;
    if (a->cx != b->cx) return a->cx < b->cx ? -1 : 1;
    if (a->cy != b->cy) return a->cy < b->cy ? -1 : 1;
    if (a->v != b->v) return a->v < b->v ? -1 : 1;
    return a->idx < b->idx ? -1 : (a->idx > b->idx);
}

void orc23_sort(int64_t P, double* x, double* y, double* vx, double* vy, double* vz, double dx, double dy,
                int64_t* perm) {
    SortKey* k = (SortKey*)malloc(sizeof(SortKey) * (size_t)(P > 0 ? P : 1));
    for (int64_t p = 0; p < P; ++p) {
        k[p].cx = orc23_ceil_index(x[p] / dx);
        k[p].cy = orc23_ceil_index(y[p] / dy);
        k[p].v = orc23_order_key(vx[p]);
        k[p].idx = p;
    }
    qsort(k, (size_t)P, sizeof(SortKey), cmp_key);
    double* t = (double*)malloc(sizeof(double) * (size_t)(P > 0 ? P : 1));
    double* cols[5] = {x, y, vx, vy, vz};
    for (int c = 0; c < 5; ++c) {
        for (int64_t p = 0; p < P; ++p) t[p] = cols[c][k[p].idx];
        memcpy(cols[c], t, sizeof(double) * (size_t)P);
    }
    if (perm)
        for (int64_t p = 0; p < P; ++p) perm[p] = k[p].idx;
    free(t);
    free(k);
}

/* ------------------------------------------------------------------------ */
/* Diagnostics of ElectrostaticDiagnostics (:1301-1330), one row of 8:       */
/*   [0] kinetic energy  sum_s ((sum(abs2, v_s)*mass)/2)*weight               */
/*   [1] field energy    mean(abs2, Exy)/2 over the ghosted 2 x (NX+2b) x     */
/*                       (NY+2b) array (ghost cells copy the periodic image)  */
/*   [2..4] momentum     sum_s (sum_p v)*(mass*weight)                        */
/*   [5..7] characteristic momentum, the same with |v|                        */
/* Every sum over particles or cells uses the canonical tree (orc_tree_sum):  */
/* the kinetic sum over the 3P velocity entries in xyv's column order; the    */
/* reference's momentum loop (:174-183) is a sequential fold, which the tree  */
/* replaces (the two agree to (P-1)*eps*sum|v|).  Species are summed left to  */
/* right as sum(generator) does.                                              */
/* ------------------------------------------------------------------------ */
void orc23_diagnostics(int32_t NX, int32_t NY, int32_t buffer, const double* Eacc_x, const double* Eacc_y,
                       int32_t ns, const int64_t* P, const double* mass, const double* weight,
                       const double* state, double* row) {
    const int64_t NXp = NX + 2 * (int64_t)buffer, NYp = NY + 2 * (int64_t)buffer;
    const int64_t nf = 2 * NXp * NYp;
    double* t = (double*)malloc(sizeof(double) * (size_t)nf);
    int64_t e = 0;
    for (int64_t jp = 0; jp < NYp; ++jp)
        for (int64_t ip = 0; ip < NXp; ++ip) {
            const int64_t k = wrap_cell(ip - (buffer - 1), NX) + (int64_t)NX * wrap_cell(jp - (buffer - 1), NY);
            t[e++] = Eacc_x[k] * Eacc_x[k];
            t[e++] = Eacc_y[k] * Eacc_y[k];
        }
    row[1] = orc_tree_sum(t, nf) / (double)nf / 2.0;
    free(t);
    double pm[3], pc[3];
    orc23_species_moments(ns, P, mass, weight, state, &row[0], pm, pc);
    for (int c = 0; c < 3; ++c) {
        row[2 + c] = pm[c];
        row[5 + c] = pc[c];
    }
}

/* kineticenergy, momentum and characteristicmomentum summed over the
 * species (:174-200, :1304-1309): state holds each species as x, y, vx, vy,
 * vz (P[s] each), species after species. */
void orc23_species_moments(int32_t ns, const int64_t* P, const double* mass, const double* weight,
                           const double* state, double* ke_out, double* pm_out, double* pc_out) {
    double ke = 0.0, pm[3] = {0, 0, 0}, pc[3] = {0, 0, 0};
    const double* st = state;
    for (int s = 0; s < ns; ++s) {
        const int64_t n = P[s];
        const double* v[3] = {st + 2 * n, st + 3 * n, st + 4 * n};
        double* a = (double*)malloc(sizeof(double) * (size_t)(3 * n > 0 ? 3 * n : 1));
        for (int64_t p = 0; p < n; ++p)
            for (int c = 0; c < 3; ++c) a[3 * p + c] = v[c][p] * v[c][p];
        const double kes = orc_tree_sum(a, 3 * n) * mass[s] / 2.0 * weight[s];
        double ms[3], cs[3];
        for (int c = 0; c < 3; ++c) {
            ms[c] = orc_tree_sum(v[c], n) * (mass[s] * weight[s]);
            for (int64_t p = 0; p < n; ++p) a[p] = fabs(v[c][p]);
            cs[c] = orc_tree_sum(a, n) * (mass[s] * weight[s]);
        }
        free(a);
        if (s == 0) {
            ke = kes;
            for (int c = 0; c < 3; ++c) { pm[c] = ms[c]; pc[c] = cs[c]; }
        } else {
            ke = ke + kes;
            for (int c = 0; c < 3; ++c) { pm[c] = pm[c] + ms[c]; pc[c] = pc[c] + cs[c]; }
        }
        st += 5 * n;
    }
    *ke_out = ke;
    for (int c = 0; c < 3; ++c) {
        pm_out[c] = pm[c];
        pc_out[c] = pc[c];
    }
}

/* ------------------------------------------------------------------------ */
/* Time loop: for t = t0 .. t0+nsteps-1, loop!(plasma, field) (:530-581)     */
/* then diagnose!(d, field, plasma, t) (:1310-1330).                         */
/*   loop!: for every species (charge, mass, weight, shape): gather at x,    */
/*     v = boris(v, E, q_m = charge/mass), x = unimod(x + ((vx0 + vx)/2)*dt,  */
/*     Lx) (y likewise), deposit (charge*weight)/dV at the new position;      */
/*     then the exact per-cell charge, the solve, and update! (:294-297):     */
/*     accumulate != 0 adds the new field into Exy as the reference does     */
/*     (Exy is never cleared, so the gathered field is the running sum of    */
/*     every step's field); accumulate == 0 replaces it (0.0 + E).            */
/*   diagnose!: if t % ntskip == 0 the row of bin t/ntskip (8 doubles); every */
/*     step Exs[:, :, b] += Ex[1:ngskip:end, 1:ngskip:end]/ntskip, Eys and    */
/*     phis likewise, b = t/ntskip.                                           */
/* state holds each species as x, y, vx, vy, vz (P[s] each), species after    */
/* species.  rows receives the rows of this call (row r of the call is bin    */
/* ceil(t0/ntskip) + r); Exs/Eys/phis (may be NULL) hold bins 0 .. nbins-1 of */
/* (NX/ngskip) x (NY/ngskip) each.  Returns 1 if the solve met a non-finite   */
/* value (the @assert of :570), else 0.                                       */
/* ------------------------------------------------------------------------ */
int orc23_run(int32_t NX, int32_t NY, double Lx, double Ly, double dt, const double* B, int32_t buffer,
              int32_t ntskip, int32_t ngskip, int32_t accumulate, int32_t boris, int32_t ns, const int32_t* shape,
              const int64_t* P, const double* charge, const double* mass, const double* weight, double* state,
              double* Eacc_x, double* Eacc_y, int64_t t0, int64_t nsteps, double* rows, double* Exs, double* Eys,
              double* phis, int64_t nbins, int32_t nthreads) {
    const int64_t G = (int64_t)NX * NY;
    const double NX_Lx = (double)NX / Lx, NY_Ly = (double)NY / Ly;
    const double dV = (Lx / (double)NX) * (Ly / (double)NY);
    double qw[8];
    for (int s = 0; s < ns; ++s) qw[s] = charge[s] * weight[s] / dV;
    const int q = orc23_frac_bits(ns, P, qw);
    const int32_t NXd = NX / ngskip, NYd = NY / ngskip;
    i128* acc = (i128*)malloc(sizeof(i128) * (size_t)G);
    double* rho = (double*)malloc(sizeof(double) * (size_t)G);
    double* ex = (double*)malloc(sizeof(double) * (size_t)G);
    double* ey = (double*)malloc(sizeof(double) * (size_t)G);
    double* ph = (double*)malloc(sizeof(double) * (size_t)G);
    int bad = 0;
    int64_t nrow = 0;
    (void)nthreads;
    for (int64_t t = t0; t < t0 + nsteps; ++t) {
        memset(acc, 0, sizeof(i128) * (size_t)G);
        double* st = state;
        for (int s = 0; s < ns; ++s) {
            const int64_t n = P[s];
            double *x = st, *y = st + n, *vx = st + 2 * n, *vy = st + 3 * n, *vz = st + 4 * n;
            const double q_m = charge[s] / mass[s];
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
            for (int64_t p = 0; p < n; ++p) {
                double fx, fy;
                orc23_gather(NX, NY, NX_Lx, NY_Ly, shape[s], Eacc_x, Eacc_y, x[p], y[p], &fx, &fy);
                const double vx0 = vx[p], vy0 = vy[p];
                orc23_boris(boris, B, dt, q_m, fx, fy, &vx[p], &vy[p], &vz[p]);
                x[p] = orc23_unimod(x[p] + (vx0 + vx[p]) / 2.0 * dt, Lx);
                y[p] = orc23_unimod(y[p] + (vy0 + vy[p]) / 2.0 * dt, Ly);
            }
            deposit_into(acc, NX, NY, NX_Lx, NY_Ly, shape[s], n, x, y, qw[s], q);
            st += 5 * n;
        }
        for (int64_t k = 0; k < G; ++k) rho[k] = unfix_signed(acc[k], q);
        bad |= orc23_fieldsolve(NX, NY, Lx, Ly, rho, ex, ey, ph);
        for (int64_t k = 0; k < G; ++k) {
            Eacc_x[k] = accumulate ? Eacc_x[k] + ex[k] : 0.0 + ex[k];
            Eacc_y[k] = accumulate ? Eacc_y[k] + ey[k] : 0.0 + ey[k];
        }
        const int64_t b = t / ntskip;
        if (t % ntskip == 0 && rows) orc23_diagnostics(NX, NY, buffer, Eacc_x, Eacc_y, ns, P, mass, weight, state,
                                                        rows + 8 * nrow++);
        if (b < nbins) {
            for (int32_t j = 0; j < NYd; ++j)
                for (int32_t i = 0; i < NXd; ++i) {
                    const int64_t k = (int64_t)i * ngskip + (int64_t)NX * ((int64_t)j * ngskip);
                    const int64_t o = b * (int64_t)NXd * NYd + i + (int64_t)NXd * j;
                    if (Exs) Exs[o] = Exs[o] + ex[k] / (double)ntskip;
                    if (Eys) Eys[o] = Eys[o] + ey[k] / (double)ntskip;
                    if (phis) phis[o] = phis[o] + ph[k] / (double)ntskip;
                }
        }
    }
    free(acc);
    free(rho);
    free(ex);
    free(ey);
    free(ph);
    return bad;
}
