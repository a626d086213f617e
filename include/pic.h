/*
 * pic.h -- C ABI of libpicgolf, the MI355X (gfx950) implementation of the
 * electrostatic particle-in-cell time step of jwscook/ParticleInCellCodeGolf.jl.
 *
 * The reference has no FFI of its own: its operators are script-local closures
 * and inline expressions.  Each entry point below replaces one of them and
 * cites the reference expression it stands in for.  A Julia drop-in module
 * binds these with `ccall`; Python binds them with ctypes.
 *
 * Conventions
 *  - Every function returns 0 (PIC_OK) or a negative PIC_E* code; the message
 *    of the last failure on a context is available from pic_last_error().
 *  - Arrays are plain host arrays of Float64 in the scripts' order; Julia index
 *    k (1-based) is C index k-1.  A cell index k corresponds to position k/N,
 *    Julia index N to position 0.
 *  - The library copies host arrays in and out and keeps no host pointer after
 *    a call returns.  Calls are synchronous.
 *  - A context is bound to one HIP device and one stream and is not
 *    thread-safe; distinct contexts may be used from distinct threads.
 */
#ifndef PIC_H
#define PIC_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PIC_OK 0
#define PIC_EINVAL (-1)     /* bad argument: size, shape, null pointer, ...      */
#define PIC_EHIP (-2)       /* HIP runtime error (message copied)                */
#define PIC_ENOMEM (-3)     /* device allocation failed                          */
#define PIC_ENONFINITE (-4) /* NaN/Inf detected in the charge or the field       */
#define PIC_ESTATE (-5)     /* call out of order (e.g. run before upload)        */

/* particle shapes: NGP, or the erf-integrated Gaussian with stencil half-width
 * 6 (src/GaussianFixedPoint.jl:5, src/Gaussian.jl:6) or 7
 * (src/GaussianFixedPointQuiet.jl:6) */
#define PIC_SHAPE_NGP 0
#define PIC_SHAPE_AREA 1   /* linear area weights, src/AreaFixedPointQuietSimpson13.jl:5 (Simpson only) */
#define PIC_SHAPE_GAUSS6 6
#define PIC_SHAPE_GAUSS7 7

/* time schemes: explicit leapfrog with half drifts (src/NGPFourier.jl:4-6,
 * src/Gaussian.jl:8-10) or the implicit midpoint Picard loop
 * (src/GaussianFixedPoint.jl:7-9, src/GaussianFixedPointQuiet.jl:8-10) */
#define PIC_SCHEME_LEAPFROG 0
#define PIC_SCHEME_MIDPOINT 1
#define PIC_SCHEME_SIMPSON 2  /* Simpson-1/3 fixed point, src/GaussianFixedPointQuietSimpson13.jl:8-17 (N <= 1024) */

/* precision of pic_run: FP64 throughout (bitwise equal to the CPU
 * restatement), or MIXED (midpoint and Simpson schemes): FP32 particle push
 * and shape weights with exact integer charge accumulation; FFT, convergence
 * test and diagnostics FP64 */
#define PIC_PREC_FP64 0
#define PIC_PREC_MIXED 1

typedef struct pic_config {
    int32_t N;          /* cells: power of two, 8 <= N <= 4096                  */
    int32_t shape;      /* PIC_SHAPE_*                                          */
    int32_t scheme;     /* PIC_SCHEME_*                                         */
    int32_t precision;  /* PIC_PREC_*                                           */
    int64_t P;          /* particles                                            */
    double W;           /* plasma frequency squared (diagnostics scale)         */
    double w;           /* deposit weight per particle, W/P*N in the scripts    */
                        /* (W/NP/dx for src/Gaussian.jl:7)                      */
    double dt;          /* time step                                            */
    double rtol;        /* Picard tolerance: 4eps (Quiet) or 1e-8               */
    int32_t max_sweeps; /* Picard iterations per step, 10 in the scripts        */
    int32_t device;     /* HIP device ordinal                                   */
} pic_config;

typedef struct pic_ctx pic_ctx;

/* per-kernel timing of pic_run_profiled: HIP events on the context stream */
typedef struct pic_kernel_stat {
    char name[32];
    int64_t launches;   /* launches that did work (early exits not counted)     */
    double total_ms;
} pic_kernel_stat;

/* ---- context --------------------------------------------------------- */
int pic_create(pic_ctx** out, const pic_config* cfg);
/* A batch of independent instances (uneven N, P, W, ...) advanced together by
 * one launch per kernel per sweep; all instances share shape, scheme,
 * precision and max_sweeps.  Instance i is addressed by its index below. */
int pic_create_batch(pic_ctx** out, const pic_config* cfgs, int32_t n_instances);
int pic_destroy(pic_ctx* ctx);
const char* pic_last_error(const pic_ctx* ctx);
const char* pic_strerror(int code);
int pic_version(void);

/* ---- device-resident runs (the scripts' @gif for t in 1:T loops) ------ */
/* Upload particle state (x in [0,1), v) and the field E carried between
 * steps (NULL = zeros, as E=zeros(N) in the scripts) for instance `inst`. */
int pic_upload(pic_ctx* ctx, int32_t inst, const double* x, const double* v, const double* E);
int pic_download(pic_ctx* ctx, int32_t inst, double* x, double* v, double* E);
/* Advance every instance nsteps steps on the device (hipGraph replay).
 * diag (nsteps x 4 per instance, row-major, instance-major; may be NULL): the
 * D[t,1:4] trace of src/GaussianFixedPointQuiet.jl:11,13.  sweeps (nsteps per
 * instance; may be NULL): Picard sweeps taken each step (1 for leapfrog). */
int pic_run(pic_ctx* ctx, int64_t nsteps, double* diag, int32_t* sweeps);
/* As pic_run, but returns at once: the steps run on a thread of the library
 * while the caller goes on.  diag and sweeps must stay valid until pic_sync.
 * Until then every other call on the context returns PIC_ESTATE (pic_destroy
 * waits for the run first).  The reference has no counterpart: its step loop
 * is the script's own (src/GaussianFixedPointQuiet.jl:7). */
int pic_run_async(pic_ctx* ctx, int64_t nsteps, double* diag, int32_t* sweeps);
/* Waits for the context's pic_run_async run and returns its status (PIC_OK
 * when none is in flight); pic_last_error then describes a failure. */
int pic_sync(pic_ctx* ctx);
/* As pic_run, but launched eagerly with HIP events around every kernel.
 * Fills up to *n_stats entries of stats and sets *n_stats to the count. */
int pic_run_profiled(pic_ctx* ctx, int64_t nsteps, double* diag, int32_t* sweeps,
                     pic_kernel_stat* stats, int32_t* n_stats);

/* ---- run path ---------------------------------------------------------- */
/* How pic_run advances the steps.  The reference has no counterpart: its
 * step loop is the script's own `for t in 1:T` (src/NGPFourier.jl:4,
 * src/GaussianFixedPointQuiet.jl:7).  Both paths give the same bits.
 *  GRAPH       captured step graphs: one launch per kernel per sweep slot
 *  PERSISTENT  one launch for the whole call: the step loop, the Picard
 *              sweeps and their test stay on the device.  Single FP64
 *              instances only: NGP leapfrog with N <= 256 and P <= 8192, or
 *              the Gaussian shapes (leapfrog or midpoint) with N <= 256 and
 *              P <= 16384
 *  AUTO        PERSISTENT where it applies, else GRAPH (the default)
 * workgroups: the Gaussian persistent run's particle workgroups (1..128,
 * each 32 particles per slot, at most 4 slots), 0 for the default. */
#define PIC_RUN_AUTO 0
#define PIC_RUN_GRAPH 1
#define PIC_RUN_PERSISTENT 2
int pic_set_run_path(pic_ctx* ctx, int32_t path, int32_t workgroups);
/* The path the next pic_run takes (GRAPH or PERSISTENT) and its workgroups
 * (0 for GRAPH). */
int pic_get_run_path(const pic_ctx* ctx, int32_t* path, int32_t* workgroups);

/* ---- particle tiling of the push launches ------------------------------ */
/* The reference has no counterpart: its particle loops are serial
 * (src/GaussianFixedPointQuiet.jl:10).  A push launch runs one workgroup of
 * `block` threads per tile of at most tile_particles particles of one
 * instance.  The tiling changes how fast a run is, never its results: state,
 * field, diagnostics and sweep counts are the same bit for bit under any
 * tiling. */
typedef struct pic_tuning {
    int64_t tile_particles;    /* particles per tile chosen                        */
    int32_t block;             /* threads per push workgroup                       */
    int32_t n_tiles;           /* tiles over all instances                         */
    int32_t candidates;        /* tilings timed                                    */
    double ms_per_step;        /* trial time per step of the chosen tiling         */
    double entry_ms_per_step;  /* trial time per step of the tiling on entry       */
} pic_tuning;
/* tile_particles >= 1 (rounded up to a multiple of 64, at most 2^20), block a
 * power of two in [64, 1024] (halved while it exceeds twice the tile). */
int pic_set_tiling(pic_ctx* ctx, int64_t tile_particles, int32_t block);
int pic_get_tiling(const pic_ctx* ctx, int64_t* tile_particles, int32_t* block, int32_t* n_tiles);
/* Autotuner: times trial_steps steps (1..64, after one warm-up step) of the
 * context's current state under each candidate tiling -- 1, 2, 3, 4, 6, 8, 12
 * or 16 tiles per compute unit, blocks of 256, 512 or 1024 threads -- and
 * keeps the fastest (the entry tiling unless a candidate beats it).  The state
 * is restored after every trial, so the call leaves the simulation where it
 * was.  out may be NULL. */
int pic_autotune(pic_ctx* ctx, int32_t trial_steps, pic_tuning* out);

/* ---- operators (FP64, bitwise equal to the CPU restatement) ---------- */
/* rho = r(x, y): r .*= 0; for l in d.((x.+y)./2); for k in l; r[k[1]] += k[2]*w
 * (src/GaussianFixedPointQuiet.jl:7; y = NULL means y = x); NGP:
 * n[f(j)] += w (src/NGPFourier.jl:5); PIC_SHAPE_AREA: the linear weights of
 * src/AreaFixedPointQuietSimpson13.jl:5-6.  rho has the context's N cells. */
int pic_deposit(pic_ctx* ctx, const double* x, const double* y, int64_t P,
                double w, int32_t shape, double* rho);
/* E = real(ifft((xi = fft(rho) ./ ik; xi[1] *= 0; xi)))
 * (src/GaussianFixedPointQuiet.jl:9, src/NGPFourier.jl:5) */
int pic_fieldsolve(pic_ctx* ctx, const double* rho, double* E);
/* u(): x .= mod.(x .+ v/2*dt, 1)   (src/NGPFourier.jl:2) */
int pic_drift_half(pic_ctx* ctx, double* x, const double* v, int64_t P, double dt);
/* u(); v += E[f.(x)]*dt (src/NGPFourier.jl:6) for NGP, or
 * u(); v[j] += sum(k->E[k[1]]*k[2], d(x[j]))*dt (src/Gaussian.jl:9-10) */
int pic_push_leapfrog(pic_ctx* ctx, double* x, double* v, const double* E,
                      int64_t P, double dt, int32_t shape);
/* x .= X .+ (v .+ V)/2*dt   (src/GaussianFixedPointQuiet.jl:9) */
int pic_midpoint_position(pic_ctx* ctx, const double* X, const double* V,
                          const double* v, double* x, int64_t P, double dt);
/* v[j] = V[j] + sum(k->E[k[1]]*k[2], d((x[j]+X[j])/2))*dt
 * (src/GaussianFixedPointQuiet.jl:10) */
int pic_push_midpoint(pic_ctx* ctx, const double* X, const double* V,
                      const double* x, double* v, const double* E, int64_t P,
                      double dt, int32_t shape);
/* v_out[j] = v_in[j] + sum(k->E[k[1]]*k[2], d((a[j]+b[j])/2))*coef/6: a stage
 * kick of the Simpson-1/3 sweep (src/GaussianFixedPointQuietSimpson13.jl:12,14,16,
 * src/AreaFixedPointQuietSimpson13.jl:11,13,15) with (a, b, coef) = (X, X, dt),
 * (X, x, 4dt) and (x, x, dt).  shape: PIC_SHAPE_AREA, GAUSS6 or GAUSS7. */
int pic_simpson_kick(pic_ctx* ctx, const double* a, const double* b, const double* v_in, double* v_out,
                     const double* E, int64_t P, double coef, int32_t shape);
/* x .= mod.(x, 1)   (src/GaussianFixedPointQuiet.jl:10) */
int pic_wrap(pic_ctx* ctx, double* x, int64_t P);
/* isapprox(F, E, rtol=rtol, atol=atol)   (src/GaussianFixedPointQuiet.jl:8) */
int pic_isapprox(pic_ctx* ctx, const double* F, const double* E, int32_t n,
                 double rtol, double atol, int32_t* result);
/* D[t,1:4] of src/GaussianFixedPointQuiet.jl:11,13 for one state */
int pic_diagnostics(pic_ctx* ctx, const double* E, const double* v, int64_t P,
                    double W, double* out4);

/* ---- allocation-free deposit ------------------------------------------ */
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory -- code under HIP graph capture, tight serving loops -- the deposit
 * of pic_deposit (src/GaussianFixedPointQuiet.jl:7, src/NGPFourier.jl:5) on
 * DEVICE arrays x, y (y = NULL means y = x) and rho (N cells), enqueued on
 * `stream` (a hipStream_t; NULL = the null stream) and returning without
 * waiting.  No context: N is an argument (a power of two, 8 <= N <= 4096).
 * Its scratch, an exact per-cell accumulator of 3 x 4096 u64 limbs (96 KiB),
 * is a static device array each GPU holds for the life of the process; the
 * call allocates, frees and synchronises nothing and clears the accumulator
 * with a kernel (safe to capture).  Calls that overlap on one GPU share that
 * accumulator and race: serialise them.  rho is bit for bit pic_deposit's.
 * Errors: PIC_EINVAL (null array, P < 1, bad N, shape or w), PIC_EHIP; the
 * message of the calling thread's last failure is pic_noalloc_last_error(). */
int pic_deposit_noalloc(const double* x, const double* y, int64_t P, double w, int32_t N,
                        int32_t shape, double* rho, void* stream);
const char* pic_noalloc_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
