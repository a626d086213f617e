/*
 * pic1d2v.c -- CPU FP64 restatement of the magnetised 1D2V scripts of
 * jwscook/ParticleInCellCodeGolf.jl: an explicit Boris push with the Gaussian
 * stencil, one species (src/NGP1D2V.jl) or two (src/NGP1D2V2S.jl), with the
 * time-binned field E(x, t) they accumulate for omega-k spectra.
 *
 * TEST INFRASTRUCTURE ONLY, like the rest of oracle/.
 *
 * Per step t (src/NGP1D2V2S.jl:31-58; src/NGP1D2V.jl:39-62 is the one-species
 * case with q = q/m = 1):
 *   E = real(ifft(fft(rho()) ./ ik)), rho() = sum over species of
 *       r[k[1]] += q*k[2]*w over d.(x) (the stencil centred on x)
 *   for each particle: Ej = sum(k->E[k[1]]*k[2], d(x[j]));
 *       (vx, vy) = boris(vx, vy, Ej, B0, dt, q_m); x += vx*dt
 *   x .= mod.(x, 1); Es[:, cld(t, bin)] .+= E
 *   t % bin == 0: D[t/bin, 1:2] = (sum(abs2, E)/N, sum(kinetic)*n0/P) ./ 2,
 *       D[3:5] = (D1 + D2, sum(vx/P ...), sum(vy/P ...)), D[1:3] .*= 2/n0
 * The scripts divide D and Es by bin after the loop; that normalisation is
 * left to the caller.  Charge sums are exact (signed contributions summed in
 * 128-bit integers and rounded once) and every other sum uses the canonical
 * tree, as in picoracle.c.
 *
 * Pinned by tests/test_oracle_1d2v.py: the Boris rotation against a closed
 * form (E = 0: rotation by 2 atan(B dt/2) with |v| kept; E != 0: the textbook
 * vector form in long double), a transcription of the loop over the pinned
 * primitives, momentum and energy bookkeeping, and the cyclotron frequency of
 * a cold magnetised plasma (the upper-hybrid frequency sqrt(n0 + B0^2)).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "picoracle.h"

typedef unsigned __int128 u128;

/* boris(vx, vy, E, B, dt, q_m) (src/NGP1D2V2S.jl:5-11): with dt2q_m = dt/2*q_m,
 * v- = [vx + E*dt2q_m, vy, 0.0], t = [0, 0, B*dt2q_m],
 * v+ = v- + 2*cross(v- + cross(v-, t), t)/(1 + dot(t, t)),
 * return (v+[1] + E*dt2q_m, v+[2]).  Every product of the cross products is
 * made, zeros included, as the SVector arithmetic does.  For q_m = 1 this is
 * src/NGP1D2V.jl:5-10 exactly (E*dt/2 == E*(dt/2) and B*dt/2 == B*(dt/2)). */
void orc_boris12(double* vx, double* vy, double E, double B, double dt, double q_m) {
    const double h = (dt / 2.0) * q_m;
    const double m1 = *vx + E * h, m2 = *vy, m3 = 0.0;
    const double t1 = 0.0, t2 = 0.0, t3 = B * h;
    const double c1 = m2 * t3 - m3 * t2, c2 = m3 * t1 - m1 * t3, c3 = m1 * t2 - m2 * t1;
    const double a1 = m1 + c1, a2 = m2 + c2, a3 = m3 + c3;
    const double d1 = a2 * t3 - a3 * t2, d2 = a3 * t1 - a1 * t3;
    const double s = 1.0 + ((t1 * t1 + t2 * t2) + t3 * t3);
    *vx = (m1 + (2.0 * d1) / s) + E * h;
    *vy = m2 + (2.0 * d2) / s;
}

/* rho() of all species: exact signed per-cell sums rounded once. */
void orc_deposit_1d2v(int32_t N, int32_t R, int32_t ns, int64_t P, const double* q, double w,
                      const double* x, double* rho) {
    double qmax = 0.0;
    for (int s = 0; s < ns; ++s) qmax = fabs(q[s]) > qmax ? fabs(q[s]) : qmax;
    const int qb = orc_fxp_frac_bits((int64_t)ns * P, w * qmax);
    u128* pos = (u128*)calloc((size_t)N, sizeof(u128));
    u128* neg = (u128*)calloc((size_t)N, sizeof(u128));
    int64_t cells[2 * 16 + 1];
    double wt[2 * 16 + 1];
    for (int s = 0; s < ns; ++s)
        for (int64_t j = 0; j < P; ++j) {
            orc_gauss_stencil(N, R, x[s * P + j], cells, wt);
            for (int32_t k = 0; k < 2 * R + 1; ++k) {
                const double a = (q[s] * wt[k]) * w;
                uint64_t lo, hi;
                if (a >= 0.0) {
                    orc_fxp_from_double(a, qb, &lo, &hi);
                    pos[cells[k]] += ((u128)hi << 64) | lo;
                } else {
                    orc_fxp_from_double(-a, qb, &lo, &hi);
                    neg[cells[k]] += ((u128)hi << 64) | lo;
                }
            }
        }
    for (int32_t i = 0; i < N; ++i) {
        if (pos[i] >= neg[i]) {
            const u128 d = pos[i] - neg[i];
            rho[i] = orc_fxp_to_double((uint64_t)d, (uint64_t)(d >> 64), qb);
        } else {
            const u128 d = neg[i] - pos[i];
            rho[i] = -orc_fxp_to_double((uint64_t)d, (uint64_t)(d >> 64), qb);
        }
    }
    free(pos);
    free(neg);
}

int orc_run_1d2v(int32_t N, int32_t R, int32_t ns, int64_t P, const double* q, const double* q_m,
                 const double* mass, double w, double n0, double B0, double dt, int32_t bin,
                 double* x, double* vx, double* vy, int64_t t0, int64_t nsteps, double* Es, int64_t nbins,
                 double* D, int32_t nthreads) {
    if (N < 4 || (N & (N - 1)) != 0 || P < 1 || (ns != 1 && ns != 2) || bin < 1) return -1;
    if (R != 6 && R != 7) return -1;
    const int nt = nthreads > 1 ? nthreads : 1;
    (void)nt;
    const size_t Nz = (size_t)N, Pz = (size_t)P;
    double* rho = (double*)malloc(sizeof(double) * Nz);
    double* E = (double*)malloc(sizeof(double) * Nz);
    double* tmp = (double*)malloc(sizeof(double) * (Pz > Nz ? Pz : Nz));
    for (int64_t it = 1; it <= nsteps; ++it) {
        const int64_t t = t0 + it;
        orc_deposit_1d2v(N, R, ns, P, q, w, x, rho);
        orc_fieldsolve(N, rho, E);
        for (int s = 0; s < ns; ++s) {
            double* xs = x + s * P;
            double* vxs = vx + s * P;
            double* vys = vy + s * P;
#pragma omp parallel for num_threads(nt) schedule(static)
            for (int64_t j = 0; j < P; ++j) {
                const double Ej = orc_gather_at(N, R, E, xs[j]);
                orc_boris12(&vxs[j], &vys[j], Ej, B0, dt, q_m[s]);
                xs[j] += vxs[j] * dt;
            }
        }
        orc_wrap((int64_t)ns * P, x);
        const int64_t b = (t + bin - 1) / bin - 1; /* cld(t, bin), 0-based */
        if (Es && b < nbins)
            for (int32_t i = 0; i < N; ++i) Es[b * N + i] += E[i];
        if (D && t % bin == 0 && b < nbins) {
            for (int32_t i = 0; i < N; ++i) tmp[i] = E[i] * E[i];
            const double se = orc_tree_sum(tmp, N);
            for (int64_t j = 0; j < P; ++j) {
                double k = mass[0] * ((vy[j] * vy[j]) + (vx[j] * vx[j]));
                if (ns == 2) k = k + mass[1] * ((vy[P + j] * vy[P + j]) + (vx[P + j] * vx[P + j]));
                tmp[j] = k;
            }
            const double sk = orc_tree_sum(tmp, P);
            double* row = D + 5 * b;
            row[0] = row[0] + (se / (double)N) / 2.0;
            row[1] = row[1] + ((sk * n0) / (double)P) / 2.0;
            for (int c = 0; c < 2; ++c) {
                const double* v = c == 0 ? vx : vy;
                for (int64_t j = 0; j < P; ++j) {
                    double m = (mass[0] * v[j]) / (double)P;
                    if (ns == 2) m = m + (mass[1] * v[P + j]) / (double)P;
                    tmp[j] = m;
                }
                const double sm = orc_tree_sum(tmp, P);
                row[3 + c] = row[3 + c] + sm;
            }
            row[2] = row[2] + (row[0] + row[1]);
            for (int c = 0; c < 3; ++c) row[c] = row[c] * (2.0 / n0);
        }
    }
    free(rho);
    free(E);
    free(tmp);
    return 0;
}
