/*
 * pic2d.c -- CPU FP64 restatement of the electrostatic 2D3V step of
 * src/Electrostatic2D3V.jl (area weighting, Boris rotation about x, 2D
 * spectral Poisson solve).
 *
 * TEST INFRASTRUCTURE ONLY (see picoracle.h): the oracle the GPU 2D path is
 * checked against.
 *
 * Arrays follow Julia's column-major layout: grid value [i, j] (1-based) is
 * element (i-1) + NX*(j-1); x runs fastest.
 *
 * Pinned by (tests/test_oracle_2d.py):
 *   orc2d_area_weights    hand cases, partition of unity, the i/i+1 pairing of
 *                         the reference and the textbook i-1/i variant.
 *   orc2d_gather          constant and linear fields (exact), brute force.
 *   orc2d_boris           E = 0: rotation about x by 2 atan(B0 dt/2) with |v|
 *                         preserved; B0 = 0: plain kick; closed form.
 *   orc2d_deposit         exact rational sums equal the fixed-point result bit
 *                         for bit; total charge P*w.
 *   orc2d_fft             numpy.fft.fft2 / ifft2, O(N^4) long-double DFT.
 *   orc2d_fieldsolve      single-mode closed forms E = -grad phi, zero mean,
 *                         Gauss's law spectrally.
 *   orc2d_diagnostics     closed forms on hand states.
 *   orc2d_run             momentum along x conserved; cold single-beam plasma
 *                         frequency; upper-hybrid frequency sqrt(n0 + B0^2) with
 *                         k perpendicular to B.
 */
#include "picoracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef unsigned __int128 u128;

static const double kTwoPi2d = 6.283185307179586; /* Julia 2pi = 2*Float64(pi) */

/* unimod(x, n) = 0 < x <= n ? x : x > n ? x - n : x + n
 * (src/Electrostatic2D3V.jl:83) */
double orc2d_unimod(double x, double n) {
    return (0.0 < x && x <= n) ? x : (x > n ? x - n : x + n);
}

int64_t orc2d_unimodi(int64_t x, int64_t n) {
    return (0 < x && x <= n) ? x : (x > n ? x - n : x + n);
}

/* g(z, NZ) (src/Electrostatic2D3V.jl:84-92): zN = z*NZ, i = unimod(ceil(zN),
 * NZ), r = i - zN; ((i, 1-r), (unimod(i+1, NZ), r)).  Cells are returned
 * 0-based.  textbook != 0 pairs i with i-1 instead (cloud in cell about the
 * nearest upper node), the variant offered beside the reference's pairing. */
void orc2d_area_weights(double z, int32_t NZ, int32_t textbook, int32_t* c0, int32_t* c1,
                        double* w0, double* w1) {
    double zN = z * (double)NZ;
    int64_t i = orc2d_unimodi((int64_t)ceil(zN), NZ);
    double r = (double)i - zN;
    int64_t i2 = textbook ? orc2d_unimodi(i - 1, NZ) : orc2d_unimodi(i + 1, NZ);
    *c0 = (int32_t)(i - 1);
    *c1 = (int32_t)(i2 - 1);
    *w0 = 1.0 - r;
    *w1 = r;
}

/* eval(F1, F2, xi, yi) (src/Electrostatic2D3V.jl:94-103): j outer, i inner,
 * wxy = wx*wy, F += real(F[i,j])*wxy under @muladd (fused multiply-add). */
void orc2d_gather(int32_t NX, int32_t NY, const double* Ex, const double* Ey, double x, double y,
                  int32_t textbook, double* ex, double* ey) {
    int32_t ci[2], cj[2];
    double wi[2], wj[2];
    orc2d_area_weights(x, NX, textbook, &ci[0], &ci[1], &wi[0], &wi[1]);
    orc2d_area_weights(y, NY, textbook, &cj[0], &cj[1], &wj[0], &wj[1]);
    double f1 = 0.0, f2 = 0.0;
    for (int b = 0; b < 2; ++b)
        for (int a = 0; a < 2; ++a) {
            double wxy = wi[a] * wj[b];
            int64_t k = (int64_t)ci[a] + (int64_t)NX * cj[b];
            f1 = fma(Ex[k], wxy, f1);
            f2 = fma(Ey[k], wxy, f2);
        }
    *ex = f1;
    *ey = f2;
}

/* boris(vx, vy, vz, Ex, Ey, dt) (src/Electrostatic2D3V.jl:32-41) with
 * tvec = (B0*dt/2, 0, 0) and tscale = 2/(1 + dot(tvec, tvec)).  Every product
 * of the two cross products is written out, zeros included, as StaticArrays
 * evaluates them. */
void orc2d_boris(double* vx, double* vy, double* vz, double Ex, double Ey, double dt, double B0) {
    const double t1 = B0 * dt / 2.0;
    const double tscale = 2.0 / (1.0 + (t1 * t1 + 0.0 * 0.0 + 0.0 * 0.0));
    const double dt_2 = dt / 2.0;
    const double e1 = Ex * dt_2, e2 = Ey * dt_2, e3 = 0.0;
    const double m1 = *vx + e1, m2 = *vy + e2, m3 = *vz + e3;
    const double c1 = m2 * 0.0 - m3 * 0.0, c2 = m3 * t1 - m1 * 0.0, c3 = m1 * 0.0 - m2 * t1;
    const double u1 = m1 + c1, u2 = m2 + c2, u3 = m3 + c3;
    const double d1 = u2 * 0.0 - u3 * 0.0, d2 = u3 * t1 - u1 * 0.0, d3 = u1 * 0.0 - u2 * t1;
    const double p1 = m1 + d1 * tscale, p2 = m2 + d2 * tscale, p3 = m3 + d3 * tscale;
    *vx = p1 + e1;
    *vy = p2 + e2;
    *vz = p3 + e3;
}

/* The particle loop of src/Electrostatic2D3V.jl:128-136 without the deposit:
 * gather at (x, y), Boris, x = unimod(x + vx*dt, 1), y = unimod(y + vy*dt, 1). */
void orc2d_push(int32_t NX, int32_t NY, int64_t P, double* x, double* y, double* vx, double* vy,
                double* vz, const double* Ex, const double* Ey, double dt, double B0, int32_t textbook,
                int32_t nthreads) {
#pragma omp parallel for num_threads(nthreads > 0 ? nthreads : 1) schedule(static)
    for (int64_t p = 0; p < P; ++p) {
        double ex, ey;
        orc2d_gather(NX, NY, Ex, Ey, x[p], y[p], textbook, &ex, &ey);
        orc2d_boris(&vx[p], &vy[p], &vz[p], ex, ey, dt, B0);
        x[p] = orc2d_unimod(x[p] + vx[p] * dt, 1.0);
        y[p] = orc2d_unimod(y[p] + vy[p] * dt, 1.0);
    }
}

/* depositcharge!(F, x, y, w) (src/Electrostatic2D3V.jl:105-109): F[i,j] +=
 * wx*wy*w for j outer, i inner.  Each product (wx*wy)*w is summed exactly
 * per cell (fixed point on the 2^-q grid of orc_fxp_frac_bits, rounded once),
 * so the result does not depend on the per-thread grids of the reference
 * (:72,139-141) or on their number. */
void orc2d_deposit(int32_t NX, int32_t NY, int64_t P, const double* x, const double* y, double w,
                   int32_t textbook, int32_t nthreads, double* rho) {
    const int64_t G = (int64_t)NX * NY;
    const int q = orc_fxp_frac_bits(P, w);
    int nt = nthreads > 0 ? nthreads : 1;
    u128* acc = (u128*)calloc((size_t)(G * nt), sizeof(u128));
#pragma omp parallel num_threads(nt)
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        u128* a = acc + (int64_t)tid * G;
#pragma omp for schedule(static)
        for (int64_t p = 0; p < P; ++p) {
            int32_t ci[2], cj[2];
            double wi[2], wj[2];
            orc2d_area_weights(x[p], NX, textbook, &ci[0], &ci[1], &wi[0], &wi[1]);
            orc2d_area_weights(y[p], NY, textbook, &cj[0], &cj[1], &wj[0], &wj[1]);
            for (int b = 0; b < 2; ++b)
                for (int c = 0; c < 2; ++c) {
                    uint64_t lo, hi;
                    orc_fxp_from_double(wi[c] * wj[b] * w, q, &lo, &hi);
                    a[(int64_t)ci[c] + (int64_t)NX * cj[b]] += ((u128)hi << 64) | lo;
                }
        }
    }
    for (int64_t k = 0; k < G; ++k) {
        u128 s = 0;
        for (int t = 0; t < nt; ++t) s += acc[(int64_t)t * G + k];
        rho[k] = orc_fxp_to_double((uint64_t)s, (uint64_t)(s >> 64), q);
    }
    free(acc);
}

/* 2D transform, unnormalised in both directions: forward = 1D transforms
 * along x for every j, then along y for every i; inverse = along y, then
 * along x.  Each 1D transform is orc_fft's radix-2 decimation in time. */
void orc2d_fft(int32_t NX, int32_t NY, double* re, double* im, int32_t inverse) {
    double* cr = (double*)malloc(sizeof(double) * (size_t)NY);
    double* ci = (double*)malloc(sizeof(double) * (size_t)NY);
    for (int pass = 0; pass < 2; ++pass) {
        int along_x = inverse ? (pass == 1) : (pass == 0);
        if (along_x) {
            for (int32_t j = 0; j < NY; ++j)
                orc_fft(NX, re + (int64_t)NX * j, im + (int64_t)NX * j, inverse);
        } else {
            for (int32_t i = 0; i < NX; ++i) {
                for (int32_t j = 0; j < NY; ++j) {
                    cr[j] = re[i + (int64_t)NX * j];
                    ci[j] = im[i + (int64_t)NX * j];
                }
                orc_fft(NY, cr, ci, inverse);
                for (int32_t j = 0; j < NY; ++j) {
                    re[i + (int64_t)NX * j] = cr[j];
                    im[i + (int64_t)NX * j] = ci[j];
                }
            }
        }
    }
    free(cr);
    free(ci);
}

/* Field solve (src/Electrostatic2D3V.jl:70-81,142-157):
 *   phi = fft(rho); phi[1,1] = 0
 *   minvkk = -im ./ (kx.^2 .+ (ky').^2), minvkk[1,1] = 0
 *     (Complex{Int}(0,-1) / s = (0.0/s, -1.0/s))
 *   tmp = phi*minvkk (Julia complex product (ac - bd, ad + bc))
 *   Ex = tmp*kx[i], Ey = tmp*ky[j]; E = real(ifft(.)) (scale 1/(NX*NY))
 * kx = 2pi*[0:NX/2-1, -NX/2:-1].  phi_out (may be NULL) receives
 * real(ifft(phi)) as the script stores in `phis` (:173): rho without its
 * mean. */
void orc2d_fieldsolve(int32_t NX, int32_t NY, const double* rho, double* Ex, double* Ey,
                      double* phi_out) {
    (void)orc2d_fieldsolve_k(NX, NY, kTwoPi2d, kTwoPi2d, rho, Ex, Ey, phi_out);
}

/* The same solve with wavenumber scales kxs, kys (kx = kxs*m): 2pi/Lx and
 * 2pi/Ly in an Lx x Ly box (src/PIC2D3V.jl:252-253).  Returns 1 if a tmp was
 * not finite (the @assert of src/Electrostatic2D3V.jl:148), else 0. */
int orc2d_fieldsolve_k(int32_t NX, int32_t NY, double kxs, double kys, const double* rho, double* Ex,
                       double* Ey, double* phi_out) {
    const int64_t G = (int64_t)NX * NY;
    int bad = 0;
    double* pr = (double*)malloc(sizeof(double) * (size_t)G);
    double* pi = (double*)malloc(sizeof(double) * (size_t)G);
    double* xr = (double*)malloc(sizeof(double) * (size_t)G);
    double* xi = (double*)malloc(sizeof(double) * (size_t)G);
    double* yr = (double*)malloc(sizeof(double) * (size_t)G);
    double* yi = (double*)malloc(sizeof(double) * (size_t)G);
    for (int64_t k = 0; k < G; ++k) { pr[k] = rho[k]; pi[k] = 0.0; }
    orc2d_fft(NX, NY, pr, pi, 0);
    pr[0] = 0.0;
    pi[0] = 0.0;
    for (int32_t j = 0; j < NY; ++j) {
        const double ky = kys * (double)(j < NY / 2 ? j : j - NY);
        for (int32_t i = 0; i < NX; ++i) {
            const double kx = kxs * (double)(i < NX / 2 ? i : i - NX);
            const int64_t k = i + (int64_t)NX * j;
            double c = 0.0, d = 0.0;
            if (k != 0) {
                const double s = kx * kx + ky * ky;
                c = 0.0 / s;
                d = -1.0 / s;
            }
            const double a = pr[k], b = pi[k];
            const double tr = a * c - b * d;
            const double ti = a * d + b * c;
            bad |= !(isfinite(tr) && isfinite(ti));
            xr[k] = tr * kx;
            xi[k] = ti * kx;
            yr[k] = tr * ky;
            yi[k] = ti * ky;
        }
    }
    orc2d_fft(NX, NY, xr, xi, 1);
    orc2d_fft(NX, NY, yr, yi, 1);
    const double scale = 1.0 / (double)G;
    for (int64_t k = 0; k < G; ++k) {
        Ex[k] = xr[k] * scale;
        Ey[k] = yr[k] * scale;
    }
    if (phi_out) {
        orc2d_fft(NX, NY, pr, pi, 1);
        for (int64_t k = 0; k < G; ++k) phi_out[k] = pr[k] * scale;
    }
    free(pr); free(pi); free(xr); free(xi); free(yr); free(yi);
    return bad;
}

/* Diagnostics (src/Electrostatic2D3V.jl:166-170):
 *   K1 = mean(Ex.^2 .+ Ey.^2), K2 = sum((vx.^2 + vy.^2).*w), K3 = K1 + K2,
 *   K4 = sum(vx)/P, K5 = sum(vy)/P
 * with every sum over the canonical tree (orc_tree_sum). */
void orc2d_diagnostics(int32_t NX, int32_t NY, int64_t P, const double* Ex, const double* Ey,
                       const double* vx, const double* vy, double w, double* K5) {
    const int64_t G = (int64_t)NX * NY;
    const int64_t n = P > G ? P : G;
    double* t = (double*)malloc(sizeof(double) * (size_t)n);
    for (int64_t k = 0; k < G; ++k) t[k] = Ex[k] * Ex[k] + Ey[k] * Ey[k];
    K5[0] = orc_tree_sum(t, G) / (double)G;
    for (int64_t p = 0; p < P; ++p) t[p] = (vx[p] * vx[p] + vy[p] * vy[p]) * w;
    K5[1] = orc_tree_sum(t, P);
    K5[2] = K5[0] + K5[1];
    K5[3] = orc_tree_sum(vx, P) / (double)P;
    K5[4] = orc_tree_sum(vy, P) / (double)P;
    free(t);
}

/* The time loop of src/Electrostatic2D3V.jl:120-176: push every particle with
 * the current field, deposit at the new positions, solve; after step t with
 * t % ns == 0 write diagnostics row t/ns - 1 of K (5 columns).  Ex, Ey carry
 * the field between calls (zeros at the start of the script). */
int orc2d_run(int32_t NX, int32_t NY, int64_t P, double dt, double B0, double w, int32_t textbook,
              int32_t ns, double* x, double* y, double* vx, double* vy, double* vz, double* Ex,
              double* Ey, int64_t nsteps, double* K, int32_t nthreads) {
    if (NX < 4 || NY < 4 || (NX & (NX - 1)) || (NY & (NY - 1)) || P < 1 || ns < 1) return -1;
    const int64_t G = (int64_t)NX * NY;
    double* rho = (double*)malloc(sizeof(double) * (size_t)G);
    for (int64_t t = 1; t <= nsteps; ++t) {
        orc2d_push(NX, NY, P, x, y, vx, vy, vz, Ex, Ey, dt, B0, textbook, nthreads);
        orc2d_deposit(NX, NY, P, x, y, w, textbook, nthreads, rho);
        orc2d_fieldsolve(NX, NY, rho, Ex, Ey, NULL);
        if (K && t % ns == 0) orc2d_diagnostics(NX, NY, P, Ex, Ey, vx, vy, w, K + 5 * (t / ns - 1));
    }
    free(rho);
    return 0;
}
