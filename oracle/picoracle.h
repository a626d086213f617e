/*
 * picoracle.h -- CPU FP64 restatement of the electrostatic PIC step of
 * jwscook/ParticleInCellCodeGolf.jl.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the oracle: it is loaded by the
 * tests under tests/, by __graft_entry__.smoke() and by the cpu_baseline leg of
 * bench.py, always as the checker and never as the thing measured or shipped.
 * The product (libpicgolf.so) never links, loads or calls it.
 *
 * Parity status of each function (how it is pinned) is listed in picoracle.c.
 */
#ifndef PICORACLE_H
#define PICORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Julia scalar semantics (src/NGPFourier.jl:2-3, src/GaussianFixedPointQuiet.jl:6,10) */
double  orc_jl_mod1f(double x);                 /* Julia mod(x, 1) on Float64     */
int64_t orc_jl_mod1i(int64_t i, int64_t n);     /* Julia mod1(i, n) on Int        */
double  orc_jl_round(double x);                 /* Julia round(x): ties to even   */

/* OpenLibm 0.8.1 (FreeBSD msun) erf / exp restatements */
double orc_exp(double x);
double orc_erf(double x);

/* canonical reductions shared with the GPU path */
double orc_tree_sum(const double* x, int64_t n);
double orc_norm2(const double* x, int64_t n);
int    orc_isapprox(const double* F, const double* E, int64_t n, double rtol, double atol);

/* fixed-point exact per-cell accumulation */
int    orc_fxp_frac_bits(int64_t P, double w);
void   orc_fxp_from_double(double a, int q, uint64_t* lo, uint64_t* hi);
double orc_fxp_to_double(uint64_t lo, uint64_t hi, int q);

/* spectral solve (A5) */
void orc_twiddles(int32_t N, double* re, double* im);
void orc_fft(int32_t N, double* re, double* im, int32_t inverse);
void orc_dft_bruteforce(int32_t N, const double* re, const double* im,
                        double* out_re, double* out_im, int32_t inverse);
void orc_fieldsolve(int32_t N, const double* rho, double* E);

/* stencils (A2, A7) */
int64_t orc_ngp_cell(double x, int32_t N);      /* 0-based cell of f(x)          */
void    orc_gauss_stencil(int32_t N, int32_t R, double c, int64_t* cells, double* wt);

/* linear area weights of src/AreaFixedPointQuietSimpson13.jl:5 (shape code
 * ORC_SHAPE_AREA wherever a shape is taken) */
#define ORC_SHAPE_AREA 1
void    orc_area_stencil(int32_t N, double c, int64_t* cells, double* wt);
/* sum(k -> E[k[1]]*k[2], d(c)) for shape 1 / 6 / 7 (left fold) */
double  orc_gather_at(int32_t N, int32_t shape, const double* E, double c);

/* deposits (A3, A8): mode 0 = exact per-cell sum (the semantics the GPU
 * reproduces bitwise), mode 1 = the scripts' literal sequential order.
 * shape: 0 = NGP, 1 = area, 6 / 7 = Gaussian.  orc_deposit_par is mode 0 on
 * nthreads OpenMP threads (bitwise equal). */
void orc_deposit(int32_t N, int32_t shape, int64_t P, const double* x,
                 const double* y, double w, int32_t mode, double* rho);
void orc_deposit_par(int32_t N, int32_t shape, int64_t P, const double* x,
                     const double* y, double w, double* rho, int32_t nthreads);

/* particle updates (A4, A6, A10, A11) */
void orc_drift_half(int64_t P, double* x, const double* v, double dt);
void orc_kick(int32_t N, int32_t shape, int64_t P, double* v, const double* x,
              const double* E, double dt);
void orc_midpoint_position(int64_t P, const double* X, const double* V,
                           const double* v, double* x, double dt);
void orc_push_midpoint(int32_t N, int32_t R, int64_t P, const double* X,
                       const double* V, const double* x, double* v,
                       const double* E, double dt);
void orc_wrap(int64_t P, double* x);

/* diagnostics (A12) */
void orc_diagnostics(int32_t N, int64_t P, const double* E, const double* v,
                     double W, double* out4);

/* initial conditions (A13) */
void orc_quiet_start(int64_t P, double* x, double* v);

/* whole runs.  shape: 0 = NGP, 6 / 7 = Gaussian half-width.  scheme: 0 =
 * leapfrog (src/NGPFourier.jl, src/Gaussian.jl), 1 = implicit midpoint
 * (src/GaussianFixedPoint.jl, src/GaussianFixedPointQuiet.jl).
 * E (length N) carries the field between steps (the scripts start from zeros).
 * diag: nsteps x 4 (may be NULL); sweeps: nsteps (may be NULL).
 * deposit_mode as for orc_deposit; nthreads > 1 uses OpenMP (bitwise equal to
 * the serial run because the per-cell sums are exact). */
int orc_run(int32_t N, int32_t shape, int32_t scheme, int64_t P, double W,
            double w, double dt, double rtol, int32_t max_sweeps,
            double* x, double* v, double* E, int64_t nsteps,
            double* diag, int32_t* sweeps, int32_t deposit_mode,
            int32_t nthreads);

/* Simpson-1/3 fixed point (src/GaussianFixedPointQuietSimpson13.jl:8-18,
 * src/AreaFixedPointQuietSimpson13.jl:7-17), simpson.c.  shape: 1 = area,
 * 6 / 7 = Gaussian.  E3 is the script's 3 x N field matrix in column-major
 * order (E[r + 3*i] is row r+1, cell i+1), carried between steps. */
int orc_run_simpson(int32_t N, int32_t shape, int64_t P, double W, double w, double dt,
                    double rtol, int32_t max_sweeps, double* x, double* v, double* E3,
                    int64_t nsteps, double* diag, int32_t* sweeps, int32_t nthreads);

/* magnetised 1D2V, one or two species (src/NGP1D2V.jl, src/NGP1D2V2S.jl),
 * pic1d2v.c.  Species s particle j is at s*P + j.  q, q_m, mass: per species
 * charge, charge-to-mass ratio and kinetic weight in the diagnostics
 * (1D2V: 1, 1, 1; 1D2V2S: (-1, 1), (-1, 1/M), (1, M)).  t0: steps already run;
 * Es (N x nbins, N fastest) accumulates E into bin ceil(t/bin); D (nbins x 5)
 * receives the script's row at every t with t % bin == 0 (both before the
 * scripts' final division by bin). */
void orc_boris12(double* vx, double* vy, double E, double B, double dt, double q_m);
void orc_deposit_1d2v(int32_t N, int32_t R, int32_t nspecies, int64_t P, const double* q, double w,
                      const double* x, double* rho);
int  orc_run_1d2v(int32_t N, int32_t R, int32_t nspecies, int64_t P, const double* q, const double* q_m,
                  const double* mass, double w, double n0, double B0, double dt, int32_t bin,
                  double* x, double* vx, double* vy, int64_t t0, int64_t nsteps, double* Es,
                  int64_t nbins, double* D, int32_t nthreads);

/* electrostatic 2D3V (src/Electrostatic2D3V.jl), pic2d.c.  Grids are
 * column-major NX x NY (x fastest); cells are returned 0-based.  textbook = 0
 * is the reference's i / i+1 area pairing, 1 the i-1 / i variant. */
double  orc2d_unimod(double x, double n);
int64_t orc2d_unimodi(int64_t x, int64_t n);
void orc2d_area_weights(double z, int32_t NZ, int32_t textbook, int32_t* c0, int32_t* c1,
                        double* w0, double* w1);
void orc2d_gather(int32_t NX, int32_t NY, const double* Ex, const double* Ey, double x, double y,
                  int32_t textbook, double* ex, double* ey);
void orc2d_boris(double* vx, double* vy, double* vz, double Ex, double Ey, double dt, double B0);
void orc2d_push(int32_t NX, int32_t NY, int64_t P, double* x, double* y, double* vx, double* vy,
                double* vz, const double* Ex, const double* Ey, double dt, double B0,
                int32_t textbook, int32_t nthreads);
void orc2d_deposit(int32_t NX, int32_t NY, int64_t P, const double* x, const double* y, double w,
                   int32_t textbook, int32_t nthreads, double* rho);
void orc2d_fft(int32_t NX, int32_t NY, double* re, double* im, int32_t inverse);
void orc2d_fieldsolve(int32_t NX, int32_t NY, const double* rho, double* Ex, double* Ey,
                      double* phi_out);
void orc2d_diagnostics(int32_t NX, int32_t NY, int64_t P, const double* Ex, const double* Ey,
                       const double* vx, const double* vy, double w, double* K5);
int  orc2d_run(int32_t NX, int32_t NY, int64_t P, double dt, double B0, double w,
               int32_t textbook, int32_t ns, double* x, double* y, double* vx, double* vy,
               double* vz, double* Ex, double* Ey, int64_t nsteps, double* K, int32_t nthreads);
int  orc2d_fieldsolve_k(int32_t NX, int32_t NY, double kxs, double kys, const double* rho, double* Ex,
                        double* Ey, double* phi_out);

/* structured 2D3V module, electrostatic path (src/PIC2D3V.jl), in pic2d3v.c */
#define ORC23_NGP 0
#define ORC23_AREA 1
#define ORC23_BSPLINE0 2 /* B-spline of order n: ORC23_BSPLINE0 + n, n = 0..5 */
double   orc23_unimod(double x, double n);
int      orc23_bspline(int32_t n, double x, double* w);
int      orc23_weights(int32_t shape, double z, double NZ_Lz, int64_t* cell, double* w);
void     orc23_gather(int32_t NX, int32_t NY, double NX_Lx, double NY_Ly, int32_t shape, const double* Ex,
                      const double* Ey, double x, double y, double* ex, double* ey);
void     orc23_boris(int32_t variant, const double* B, double dt, double q_m, double Ex, double Ey, double* vx,
                     double* vy, double* vz);
int      orc23_frac_bits(int32_t ns, const int64_t* P, const double* w);
void     orc23_deposit(int32_t NX, int32_t NY, double NX_Lx, double NY_Ly, int32_t shape, int64_t P,
                       const double* x, const double* y, double w, double* rho);
int      orc23_fieldsolve(int32_t NX, int32_t NY, double Lx, double Ly, const double* rho, double* Ex, double* Ey,
                          double* phi_out);
uint64_t orc23_order_key(double v);
int64_t  orc23_ceil_index(double a);
void     orc23_sort(int64_t P, double* x, double* y, double* vx, double* vy, double* vz, double dx, double dy,
                    int64_t* perm);
void     orc23_species_moments(int32_t ns, const int64_t* P, const double* mass, const double* weight,
                               const double* state, double* ke_out, double* pm_out, double* pc_out);
void     orc23_diagnostics(int32_t NX, int32_t NY, int32_t buffer, const double* Eacc_x, const double* Eacc_y,
                           int32_t ns, const int64_t* P, const double* mass, const double* weight,
                           const double* state, double* row);
int      orc23_run(int32_t NX, int32_t NY, double Lx, double Ly, double dt, const double* B, int32_t buffer,
                   int32_t ntskip, int32_t ngskip, int32_t accumulate, int32_t boris, int32_t ns,
                   const int32_t* shape, const int64_t* P, const double* charge, const double* mass,
                   const double* weight, double* state, double* Eacc_x, double* Eacc_y, int64_t t0,
                   int64_t nsteps, double* rows, double* Exs, double* Eys, double* phis, int64_t nbins,
                   int32_t nthreads);

/* Lorenz-gauge electromagnetic 2D3V (src/PIC2D3V.jl LorenzGaugeStaggeredField), in
 * lorenz2d.c.  Field state: 18 complex NX x NY arrays (interleaved re, im). */
#define ORCLG_PHI0 0
#define ORCLG_PHIM 1
#define ORCLG_AX0 2
#define ORCLG_AY0 3
#define ORCLG_AZ0 4
#define ORCLG_AXM 5
#define ORCLG_AYM 6
#define ORCLG_AZM 7
#define ORCLG_RHO0 8
#define ORCLG_JX0 9
#define ORCLG_JY0 10
#define ORCLG_JZ0 11
#define ORCLG_EX 12
#define ORCLG_EY 13
#define ORCLG_EZ 14
#define ORCLG_BX 15
#define ORCLG_BY 16
#define ORCLG_BZ 17
#define ORCLG_RHOM 18 /* LorenzGaugeSemiImplicitField: rho-, Jx-, Jy-, Jz- */
#define ORCLG_JXM 19
#define ORCLG_JYM 20
#define ORCLG_JZM 21
void orclg_wavenumbers(int32_t NX, int32_t NY, double Lx, double Ly, double* kx, double* ky, double* k2);
void orclg_boris(double dt, double q_m, const double* E, const double* B, double* v);
void orclg_gather(int32_t NX, int32_t NY, double NX_Lx, double NY_Ly, int32_t shape, const double* EB, double x,
                  double y, double* out);
void orclg_ifft2(int32_t NX, int32_t NY, double* z);
void orclg_solve(int32_t NX, int32_t NY, double Lx, double Ly, double dt, double theta, double* F, double* plus);
void orclg_update(int32_t NX, int32_t NY, const double* F, const double* B0, double* EB);
int  orclg_frac_bits(int32_t ns, const int64_t* P, const double* w, const double* M);
void orclg_warmup(int32_t NX, int32_t NY, double Lx, double Ly, double dt, int32_t ns, const int32_t* shape,
                  const int64_t* P, const double* charge, const double* weight, double* state, double* F);
void orclg_row(int32_t NX, int32_t NY, int32_t buffer, const double* F, const double* EB, int32_t ns,
               const int64_t* P, const double* mass, const double* weight, const double* state, double* row);
void orclg_fft2(int32_t NX, int32_t NY, double* z);
void orclg_solve_centred(int32_t NX, int32_t NY, double Lx, double Ly, double dt, double theta, double* F);
int  orclg_run_centred(int32_t NX, int32_t NY, double Lx, double Ly, double dt, const double* B0, double theta,
                       int32_t ns, const int32_t* shape, const int64_t* P, const double* charge, const double* mass,
                       const double* weight, double* state, double* F, double* EB, int64_t nsteps);
int  orclg_run(int32_t NX, int32_t NY, double Lx, double Ly, double dt, const double* B0, int32_t buffer,
               double theta, int32_t ntskip, int32_t ngskip, int32_t ns, const int32_t* shape, const int64_t* P,
               const double* charge, const double* mass, const double* weight, double* state, double* F,
               int64_t t0, int64_t nsteps, int32_t diagnose, double* rows, double* bins, int64_t nbins);
int  orclg_si_deposit(int32_t NX, int32_t NY, int32_t buffer, double Lx, double Ly, int32_t ns, const int32_t* shape,
                      const int64_t* P, const double* charge, const double* weight, const double* state, int32_t c0,
                      double* grid, double* per);
int  orclg_si_warmup(int32_t NX, int32_t NY, double Lx, double Ly, double dt, int32_t buffer, int32_t ns,
                     const int32_t* shape, const int64_t* P, const double* charge, const double* weight,
                     double* state, double* F, double* grid);
void orclg_si_solve(int32_t NX, int32_t NY, double Lx, double Ly, double dt, double theta_f, double theta_s,
                    double* F, double* src, double* plus);
int  orclg_si_run(int32_t NX, int32_t NY, double Lx, double Ly, double dt, const double* B0, int32_t buffer,
                  double theta_f, double theta_s, double rtol, int64_t maxiters, int32_t ntskip, int32_t ngskip,
                  int32_t ns, const int32_t* shape, const int64_t* P, const double* charge, const double* mass,
                  const double* weight, double* state, double* F, double* grid, double* EB, int64_t t0,
                  int64_t nsteps, int32_t diagnose, double* rows, double* bins, int64_t nbins, int64_t* iters);

/* 1D electromagnetic fixed point (src/GaussianEMFixedPoint.jl), in em1d.c */
int64_t orc_em_irn(double c, int32_t N);
void    orc_em_stencil(int32_t N, double c, int64_t* cells, double* ff);
int     orc_em_frac_bits(int64_t P, double w, double M);
void    orc_em_deposit(int32_t N, int64_t P, const double* y, const double* u, double w, double* rho);
void    orc_em_fields(int32_t N, double eps, double dt, const double* rho, const double* Jy, const double* Jz,
                      const double* Fy, const double* Fz, const double* Cy, const double* Cz, double* E, double* B,
                      double* e3, double* b2);
void    orc_em_push(int32_t N, double c, const double* ex, const double* ey, const double* ez, const double* by,
                    const double* bz, double dt, const double* u, double* v);
void    orc_em_sort(int64_t P, double* x, double* v1, double* v2, double* v3, int64_t* perm);
int     orc_em_run(int32_t N, int64_t P, double eps, double dt, double w, double rtol, int32_t max_sweeps,
                   int32_t sort, double* x, double* v1, double* v2, double* v3, double* E, double* B, double* Fz,
                   int64_t nsteps, double* fields, int32_t* sweeps);
/* 1D electromagnetic explicit step (src/GaussianEMExplicit.jl), in em1d.c */
void    orc_emx_gauss_faraday(int32_t N, double dt, const double* rho, double* E, double* B, double* b2);
void    orc_emx_ampere(int32_t N, double eps, double dt, const double* Jy, const double* Jz, double* E,
                       const double* B);
void    orc_emx_push(int32_t N, double x, const double* E, const double* b2, double dt, double* v);
int     orc_emx_run(int32_t N, int64_t P, double eps, double dt, double w, int32_t sort, double* x, double* v1,
                    double* v2, double* v3, double* E, double* B, int64_t nsteps, double* fields);


#ifdef __cplusplus
}
#endif
#endif
