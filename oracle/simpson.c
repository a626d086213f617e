/*
 * simpson.c -- CPU FP64 restatement of the Simpson-1/3 fixed-point step of
 * jwscook/ParticleInCellCodeGolf.jl (src/GaussianFixedPointQuietSimpson13.jl,
 * src/AreaFixedPointQuietSimpson13.jl).
 *
 * TEST INFRASTRUCTURE ONLY, like the rest of oracle/: loaded by the tests and
 * never by the product.
 *
 * Per step (GaussianFixedPointQuietSimpson13.jl:8-18):
 *   X .= x; V .= v; F .*= NaN
 *   E[1,:] = solve(rho(X, X))
 *   for _ in 0:9
 *     isapprox(F, E, rtol=l, atol=0) && break; F .= E      (3 x N matrices)
 *     x .= X .+ (v .+ V)/2*dt
 *     v[j] = V[j] + sum(k->E[1,k[1]]*k[2], d((X[j]+X[j])/2))*dt/6
 *     E[2,:] = solve(rho(X, x))
 *     v[j] = v[j] + sum(k->E[2,k[1]]*k[2], d((X[j]+x[j])/2))*4dt/6
 *     E[3,:] = solve(rho(x, x))
 *     v[j] = v[j] + sum(k->E[3,k[1]]*k[2], d((x[j]+x[j])/2))*dt/6
 *   end; x .= mod.(x, 1)
 *   D from E[end,:] and v
 * Julia binds `4dt` tighter than `*` and `/`, so the middle kick is
 * (sum*(4*dt))/6; the outer ones are (sum*dt)/6.  The deposits use the exact
 * per-cell sums of picoracle.c and isapprox runs over the 3N entries in
 * column-major order with the canonical tree (the order Julia's norm visits a
 * matrix in; its pairwise grouping is not pinned).
 *
 * Pinned by tests/test_oracle_simpson.py: a transcription of the loop above in
 * Python over the separately pinned primitives (bitwise, serial and threaded);
 * the analytic two-stream growth rate of the quiet start (within 1%); momentum
 * conservation to 1e-14; the area weights by hand cases and exact linear
 * interpolation.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "picoracle.h"

#ifdef _OPENMP
#include <omp.h>
#endif

static void solve_row(int32_t N, const double* rho, double* E3, int32_t r, double* tmp) {
    orc_fieldsolve(N, rho, tmp);
    for (int32_t i = 0; i < N; ++i) E3[r + 3 * i] = tmp[i];
}

static void row_of(int32_t N, const double* E3, int32_t r, double* out) {
    for (int32_t i = 0; i < N; ++i) out[i] = E3[r + 3 * i];
}

int orc_run_simpson(int32_t N, int32_t shape, int64_t P, double W, double w, double dt,
                    double rtol, int32_t max_sweeps, double* x, double* v, double* E3,
                    int64_t nsteps, double* diag, int32_t* sweeps, int32_t nthreads) {
    if (N < 4 || (N & (N - 1)) != 0 || P < 1) return -1;
    if (shape != ORC_SHAPE_AREA && shape != 6 && shape != 7) return -1;
    const size_t Pz = (size_t)P, Nz = (size_t)N;
    double* X = (double*)malloc(sizeof(double) * Pz);
    double* V = (double*)malloc(sizeof(double) * Pz);
    double* F = (double*)malloc(sizeof(double) * 3 * Nz);
    double* rho = (double*)malloc(sizeof(double) * Nz);
    double* tmp = (double*)malloc(sizeof(double) * Nz);
    double* E1 = (double*)malloc(sizeof(double) * Nz);
    double* E2 = (double*)malloc(sizeof(double) * Nz);
    double* Ek = (double*)malloc(sizeof(double) * Nz);
    const int nt = nthreads > 1 ? nthreads : 1;
    (void)nt;
    for (int64_t t = 0; t < nsteps; ++t) {
        memcpy(X, x, sizeof(double) * Pz);
        memcpy(V, v, sizeof(double) * Pz);
        for (size_t e = 0; e < 3 * Nz; ++e) F[e] = NAN; /* F .*= NaN */
        orc_deposit_par(N, shape, P, X, X, w, rho, nthreads);
        solve_row(N, rho, E3, 0, tmp);
        row_of(N, E3, 0, E1);
        int32_t it;
        for (it = 0; it < max_sweeps; ++it) {
            if (orc_isapprox(F, E3, 3 * (int64_t)N, rtol, 0.0)) break;
            memcpy(F, E3, sizeof(double) * 3 * Nz);
            orc_midpoint_position(P, X, V, v, x, dt);
#pragma omp parallel for num_threads(nt) schedule(static)
            for (int64_t j = 0; j < P; ++j)
                v[j] = V[j] + (orc_gather_at(N, shape, E1, (X[j] + X[j]) / 2.0) * dt) / 6.0;
            orc_deposit_par(N, shape, P, X, x, w, rho, nthreads);
            solve_row(N, rho, E3, 1, tmp);
            row_of(N, E3, 1, E2);
#pragma omp parallel for num_threads(nt) schedule(static)
            for (int64_t j = 0; j < P; ++j)
                v[j] = v[j] + (orc_gather_at(N, shape, E2, (X[j] + x[j]) / 2.0) * (4.0 * dt)) / 6.0;
            orc_deposit_par(N, shape, P, x, x, w, rho, nthreads);
            solve_row(N, rho, E3, 2, tmp);
            row_of(N, E3, 2, Ek);
#pragma omp parallel for num_threads(nt) schedule(static)
            for (int64_t j = 0; j < P; ++j)
                v[j] = v[j] + (orc_gather_at(N, shape, Ek, (x[j] + x[j]) / 2.0) * dt) / 6.0;
        }
        orc_wrap(P, x);
        if (diag) {
            row_of(N, E3, 2, Ek);
            orc_diagnostics(N, P, Ek, v, W, diag + 4 * t);
        }
        if (sweeps) sweeps[t] = it;
    }
    free(X); free(V); free(F); free(rho); free(tmp); free(E1); free(E2); free(Ek);
    return 0;
}
