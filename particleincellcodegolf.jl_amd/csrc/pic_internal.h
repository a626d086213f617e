// pic_internal.h -- descriptors shared by the host runtime and the kernels,
// and the host-side launchers the kernels file exports.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

namespace pic {

constexpr int kMaxN = 4096;       // largest grid (one LDS-resident transform)
constexpr int kTreeLanes = 256;   // canonical reduction: lanes per chunk
constexpr int kTreeChunk = 4096;  // canonical reduction: elements per chunk
constexpr int kRedDoubles = 6 * kTreeLanes + 8;  // LDS scratch of the norm trees
constexpr int kV2Bits = 50;  // mixed kinetic sums: grid of v^2 is 2^-50
constexpr int kV1Bits = 40;  // and of v, 2^-40
constexpr int kSolveThreads = 512;  // threads of the one-block-per-instance solve
constexpr int kPassStages = 3;      // radix-2 FFT stages applied per LDS pass
constexpr int kXiOffset = 8192;     // doubles before the 1/(2 pi k) factors in the twiddle table
constexpr double kTwoPi = 6.283185307179586;  // Julia 2pi = 2*Float64(pi)

// Twiddle table (double2 pairs per radix-2 stage, from W[m] = exp(-2 pi i m/4096))
// followed by the solve's 1/(2 pi k) factors at double kXiOffset.
std::vector<double> pass_twiddle_table();

// One simulation instance (a batch holds many).  Device-resident, read-only
// during a run.
struct InstDesc {
    int32_t N, log2N, nmask, R;  // cells, log2(N), N-1, stencil half-width (0: NGP)
    int64_t P;                   // particles
    int64_t poff;                // first particle in the packed particle arrays
    int32_t goff;                // first cell in the packed grid arrays
    int32_t tile0, ntiles;       // tiles of this instance in the tile list
    int32_t chunk0, nchunks;     // diagnostics chunks (4096 particles each)
    int32_t q;                   // fixed-point fraction bits for the FP64 deposit
    int32_t max_sweeps;
    int64_t red0;                // first cell of this instance in the charge accumulator
    double w, dt, W, rtol;
    double wq;                   // mixed deposit: w * 2^-30 (charge per weight quantum)
    // mixed: velocities are held in steps of the 2^-32 fixed-point position,
    // v_s = v * vstep with vstep = dt * 2^32, so every position update is a
    // power-of-two scaling; the gather reads E * escale2 (escale2 = dt^2 * 4,
    // applied in FP64 once per cell) against weights in 2^-30 units
    double vstep, inv_vstep, escale2;
};

// Device-mutable per-instance loop state.
struct InstState {
    int32_t e_cur;       // E buffer (0/1) holding the latest field
    int32_t last_sweep;  // -1 while sweeping; k once solve k ended the Picard loop
    int32_t step;        // steps completed since upload
    int32_t err;         // bit 0: non-finite charge/field
    int32_t run_base;    // step at the start of the current pic_run
    int32_t pad[3];
};

struct TileDesc {
    int32_t inst;
    int32_t local;       // tile index within the instance
    int64_t start;       // first particle (instance-relative)
    int64_t count;
};

struct ChunkDesc {       // one diagnostics chunk
    int32_t inst, pad;
    int64_t start, count;
};

struct RunBuffers {
    InstDesc* insts;
    InstState* st;
    TileDesc* tiles;
    ChunkDesc* chunks;
    int32_t ninst, ntiles, nchunks;
    int32_t maxN, R, scheme, precision, max_sweeps;
    int32_t block_push, block_solve, lds_push;
    // FP64 particle state, ping-pong by step parity, and the Picard iterate
    double* X[2];
    double* V[2];
    double* vit[2];
    // mixed particle state: 32-bit fixed-point positions (2^-32 of the box),
    // FP32 velocity as a compensated hi + lo pair, FP32 Picard iterate
    uint32_t* Xq[2];
    float* Vh[2];
    float* Vl[2];
    float* vf[2];
    void* racc;        // per-cell charge accumulator: 3 u64 limbs (FP64) or u64 quanta (mixed)
    double* rho;       // packed charge, sum of N over instances
    double* E[2];      // packed field, ping-pong (InstState::e_cur)
    float* E32;        // FP32 copy of the latest field for the mixed gather
    const double* tw;  // twiddles of the 4096-point transform, (re, im) pairs
    double* dpart;     // diagnostics partials, 2 per chunk
    double* diag;      // [ninst][diag_cap][4]
    int32_t* sweeps;   // [ninst][diag_cap]
    int64_t diag_cap;
    // Speculative sweep count.  A step graph may hold fewer sweep slots than
    // max_sweeps; a solve in the last slot that finds an instance unconverged
    // sets *halt = 1, the push of that slot still runs, diagnostics turn it
    // into 2 and every later kernel returns at once.  The host then finishes
    // the step with the remaining sweeps and resumes.
    int32_t* halt;
    // midpoint: the final sweep also deposits the next step's first charge,
    // so steps after the first carry no first-deposit launch
    int32_t prime_next;
    int64_t racc_cells;  // u64 words of the charge accumulator
    // Simpson-1/3 (FP64): E[0..1] hold E[3,:] (ping-pong by e_cur), Es1 is
    // E[1,:], Es2[0..1] E[2,:]; racc collects rho(X, X), racc2 rho(X, x) and
    // racc3 rho(x, x); Vg[p] = V + (gather(E[1,:], X)*dt)/6 of the step.
    double* Es1;
    double* Es2[2];
    void* racc2;
    void* racc3;
    double* Vg;
    // mixed Simpson: FP32 copies of E[1,:] and E[2,:] for the gathers (E32
    // holds E[3,:]'s), and G1s[p] = gather(E[1,:], X)/6 in steps of the grid;
    // racc2, racc3 hold u64 quanta
    float* E32s1;
    float* E32s2;
    float* G1s;
};

// kernels_1d.hip
void launch_mid_push0(const RunBuffers& b, hipStream_t s);
// last_slot: this is the graph's last sweep slot (it runs while *halt == 1)
void launch_mid_push(const RunBuffers& b, int k, bool last_slot, hipStream_t s);
// stall: an instance still unconverged after this solve halts the graph
void launch_solve(const RunBuffers& b, int k, bool stall, hipStream_t s);
// Simpson-1/3: the step's prologue (rho(X,X), E[1,:], Vg and the charges of
// sweep 0), then per sweep slot a solve (both fields and the test) and a push
int simpson_push_lds_bytes(int N);
void launch_simp_first(const RunBuffers& b, hipStream_t s);
void launch_simp_solve(const RunBuffers& b, int k, bool stall, hipStream_t s);
void launch_simp_push(const RunBuffers& b, int k, bool last_slot, hipStream_t s);
// kernels_mixed_aux.hip: the mixed forms (called by the launchers above)
int simpson32_push_lds_bytes(int N);
void launch_simp32_first(const RunBuffers& b, hipStream_t s);
void launch_simp32_start(const RunBuffers& b, hipStream_t s);
void launch_simp32_push(const RunBuffers& b, int k, bool last_slot, hipStream_t s);
void launch_lf_first(const RunBuffers& b, hipStream_t s);
void launch_lf_second(const RunBuffers& b, hipStream_t s);
void launch_diag(const RunBuffers& b, hipStream_t s);
int push_lds_bytes(int N, int precision);
int configure_kernels();      // raise the dynamic LDS limit; 0 on success
bool mixed_available();

// kernels_mixed.hip
void launch_mid_push32(const RunBuffers& b, int mode, int k, bool last_slot, hipStream_t s);
int push32_lds_bytes(int N);
int configure_kernels_mixed();

// kernels_small.hip: the whole run of a small single FP64 instance in one
// launch (the step loop, the Picard sweeps and their test on the device)
constexpr int kSmallMaxN = 256;       // the solve runs in one wave, 4 grid cells per lane
constexpr int kSmallNgpSlots = 16;    // NGP: particles per thread of the one 512-thread workgroup (P <= 8192)
constexpr int kSmallMaxSlots = 4;     // Gaussian: particles per 16-lane segment
constexpr int kSmallMaxGroups = 128;  // Gaussian: particle workgroups (32 segments each; P <= 16384)
struct SmallRun {
    InstDesc I;
    InstState* st;
    double *X0, *X1, *V0, *V1, *E0, *E1;
    const double2* tw;
    double* diag;                // nsteps x 4
    int32_t* sweeps;             // nsteps
    int64_t nsteps;
    unsigned long long* gacc;    // Gaussian: three device charge accumulators of 3N limbs
    uint32_t* bar;               // Gaussian: barrier counter, generation, failure flag
    int32_t K;                   // Gaussian: particle workgroups (one more computes diagnostics)
    uint32_t spin_limit;         // Gaussian: polls before a barrier wait gives up
};
// whether the instance can run persistently; *K gets the default workgroups
bool small_eligible(const InstDesc& I, int scheme, int precision, int64_t* K);
int small_workgroups_ok(const InstDesc& I, int64_t K);
hipError_t launch_small_run(const SmallRun& a, int scheme, hipStream_t s);
int configure_kernels_small();

// operators (single grid of N cells, contiguous particle arrays)
struct OpBuffers {
    int32_t N, log2N;
    const double* tw;
    void* part;          // >= ntiles * N u128
    double* dpart;       // >= 2 * ceil(P/4096)
    int32_t* flag;
};
void launch_op_deposit(const OpBuffers& o, const double* x, const double* y, int64_t P,
                       double w, int R, int q, double* rho, hipStream_t s);
void launch_op_fieldsolve(const OpBuffers& o, const double* rho, double* E, hipStream_t s);
void launch_op_drift_half(double* x, const double* v, int64_t P, double dt, hipStream_t s);
void launch_op_push_leapfrog(const OpBuffers& o, double* x, double* v, const double* E,
                             int64_t P, double dt, int R, hipStream_t s);
void launch_op_midpos(const double* X, const double* V, const double* v, double* x,
                      int64_t P, double dt, hipStream_t s);
void launch_op_push_midpoint(const OpBuffers& o, const double* X, const double* V,
                             const double* x, double* v, const double* E, int64_t P,
                             double dt, int R, hipStream_t s);
void launch_op_stage_kick(const OpBuffers& o, const double* a, const double* b, const double* vin, double* vout,
                          const double* E, int64_t P, double coef, int shape, hipStream_t s);
void launch_op_wrap(double* x, int64_t P, hipStream_t s);
void launch_op_isapprox(const double* F, const double* E, int32_t n, double rtol,
                        double atol, int32_t* out, hipStream_t s);
void launch_op_diagnostics(const OpBuffers& o, const double* E, const double* v, int64_t P,
                           double W, double* out4, hipStream_t s);
int op_tiles(int64_t P);

}  // namespace pic
