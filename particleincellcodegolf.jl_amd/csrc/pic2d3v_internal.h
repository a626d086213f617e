// pic2d3v_internal.h -- descriptors and launchers of the electrostatic path
// of the structured 2D3V module, shared by the host runtime (capi2d3v.cpp)
// and the kernels (kernels_2d3v.hip, kernels_2d3v_mixed.hip).
#pragma once
#include "pic2d_internal.h"

namespace pic {

constexpr int kMaxSpecies23 = 4;
constexpr int kPush23Threads = 256;
constexpr int64_t kMaxP23 = int64_t(1) << 26;  // all species together: 2^26 contributions per cell
constexpr int64_t kMaxSortBuckets = int64_t(1) << 24;
constexpr int64_t kMaxSortCell = 16384;        // particles one cell may hold in a sort

// Grid, box and the step's constants.
struct Grid23 {
    int32_t NX, NY, lx, ly;
    double Lx, Ly;
    double NX_Lx, NY_Ly;  // NX/Lx, NY/Ly (GridParameters, src/PIC2D3V.jl:234-240)
    double dt;
    double B[3];
    int32_t q;            // fraction bits of the signed exact charge sums
    int32_t boris;        // 0: ElectrostaticBoris, 1: ElectromagneticBoris form
    // mixed precision: x = xq*Lx/2^32 (xq = 0 meaning x = Lx), so the cell of
    // xq is (xq - 1) >> sx; a move of v*dt is v*step quanta of 2^-32 of the
    // box, step = dt*2^32/L held as a float hi + lo pair
    int32_t sx, sy;
    float stepx[2], stepy[2];
    // the electrostatic pushes add the number of stencils (a gather and a
    // deposit per particle) that left their chunk's window and read or added
    // through global memory
    unsigned long long* fallback;
};

// mixed precision: shape weights in quanta of 2^-kQ23 of a cell per axis
// (products of two: 2^-44 of a particle's charge)
constexpr int kQ23 = 22;

// One species on the device (SoA).
struct Species23 {
    int64_t P;
    int32_t shape;
    double q_m;           // charge/mass
    double qw;            // (charge*weight)/dV
    double mass, weight;
    double *x, *y, *vx, *vy, *vz;
    // the slot holding caller index i (the particles are regrouped by tile
    // every few steps; diagnostics and downloads run in the caller's order)
    uint32_t* slot;
    // mixed precision: the state (positions in fixed point, velocities as
    // FP32 hi + lo), the Boris constants (v' = v + 2e + ((v+e) + (v+e) x t) x t * bs
    // with e = E*eh), and the species' charge quanta (lo words, then hi words)
    int32_t mixed;
    uint32_t *xq, *yq;
    float *vxh, *vyh, *vzh, *vxl, *vyl, *vzl;
    float bt[3], bs, eh;
    unsigned long long* cnt;
    double qscale;        // qw * 2^(-2 kQ23): the charge of one quantum
};

// Regrouping of held particles by kRegroupTile23 x kRegroupTile23-cell tile.
constexpr int kRegroupTile23 = 8;
constexpr int kRegroupEvery23 = 16;  // steps between regroupings (the electrostatic path: between checks)
constexpr double kRegroupFallback = 0.01;  // electrostatic path: regroup above this fraction of stencils falling back
struct Tiles23 {
    int32_t TX, TY, ntx, nty, ntiles;
    int32_t *count, *start, *cursor;
};

// Diagnostics and binned-field state.
struct Diag23 {
    int32_t buffer, ntskip, ngskip, NXd, NYd;
    int64_t* t;           // device step counter
    double* rows;         // 8 per bin
    double *Exs, *Eys, *phis;  // NXd*NYd per bin
    int64_t cap;          // bins allocated
    double* fpart;        // field-energy chunk partials
    double* kpart[kMaxSpecies23];  // kinetic chunk partials (3P entries per species)
    double* mpart[kMaxSpecies23];  // momentum chunk partials, 6 per chunk
};

// one step's particle loop for one species: gather, Boris, move, deposit
void launch23_push(const Grid23& g, const Species23& s, const double2* E, unsigned long long* racc, hipStream_t st);
// update! of the gathered field and the binned fields
void launch23_update(const Grid23& g, const Diag23& d, const Buffers2D& b, double2* E, int accumulate,
                     hipStream_t st);
// the diagnostics row of bin t/ntskip
void launch23_diag(const Grid23& g, const Diag23& d, const double2* E, const Species23* sp, int ns, hipStream_t st);
// the species sums of the diagnostics: kinetic and momentum chunk partials
void launch23_diag_species(const Species23& s, double* kpart, double* mpart, hipStream_t st);
// t += 1
void launch23_tick(int64_t* t, hipStream_t st);
// xyv (5 x P, column-major) <-> SoA
void launch23_unpack(const double* xyv, const Species23& s, hipStream_t st);
void launch23_pack(const Species23& s, double* xyv, hipStream_t st);
// regrouping: the held particles of s by tile into o (and pid into pido,
// s.slot updated); the held particles back in the caller's order (into o);
// pid = slot = 0 .. P-1
void launch23_regroup(const Grid23& g, const Tiles23& T, const Species23& s, const Species23& o, const uint32_t* pid,
                      uint32_t* pido, hipStream_t st);
void launch23_to_caller(const Species23& s, const Species23& o, hipStream_t st);
// the exclusive scan of the tile counts into start and the cursors
void launch23_regroup_scan(const Tiles23& T, hipStream_t st);
void launch23_identity(uint32_t* pid, uint32_t* slot, int64_t P, hipStream_t st);
// operators
void launch23_deposit_op(const Grid23& g, int shape, const double* x, const double* y, int64_t n, double w,
                         unsigned long long* racc, hipStream_t st);
void launch23_limbs_to_rho(const Grid23& g, unsigned long long* racc, double* rho, hipStream_t st);
void launch23_gather_op(const Grid23& g, int shape, const double2* E, const double* x, const double* y, int64_t n,
                        double* ex, double* ey, hipStream_t st);
void launch23_boris_op(const Grid23& g, double q_m, double* vx, double* vy, double* vz, const double* ex,
                       const double* ey, int64_t n, hipStream_t st);
void launch23_pair_field(const double* Ex, const double* Ey, double2* E, int64_t G, hipStream_t st);
void launch23_split_field(const double2* E, double* Ex, double* Ey, int64_t G, hipStream_t st);

// mixed precision (kernels_2d3v_mixed.hip): the particle loop (gather from
// the FP32 field copy E32, Boris, fixed-point move, exact quanta deposit into
// s.cnt); the species' quanta into rho (read and cleared); xyv <-> the mixed
// state; the mixed state by tile, back to the caller's order, permuted; FP64
// copies of the held state (for the sort's keys)
void launch23m_push(const Grid23& g, const Species23& s, const float2* E32, hipStream_t st);
void launch23m_rho(const Grid23& g, const Species23* sp, int ns, double* rho, hipStream_t st);
void launch23m_unpack(const Grid23& g, const double* xyv, const Species23& s, hipStream_t st);
void launch23m_pack(const Grid23& g, const Species23& s, double* xyv, hipStream_t st);
void launch23m_regroup(const Grid23& g, const Tiles23& T, const Species23& s, const Species23& o,
                       const uint32_t* pid, uint32_t* pido, hipStream_t st);
void launch23m_to_caller(const Species23& s, const Species23& o, hipStream_t st);
void launch23m_permute(const Species23& s, const Species23& o, const uint32_t* perm, hipStream_t st);
void launch23m_to64(const Grid23& g, const Species23& s, hipStream_t st);
// the kinetic and momentum partials of a mixed species (one of each kind per
// kTreeChunk particles); returns the number of chunks
int64_t launch23m_diag_species(const Species23& s, double* kpart, double* mpart, hipStream_t st);
// update! writing the FP32 copy of Exy as well
void launch23_update32(const Grid23& g, const Diag23& d, const Buffers2D& b, double2* E, float2* E32,
                       int accumulate, hipStream_t st);

// sort!(species, dx, dy)
struct Sort23 {
    int64_t* ext;          // [4]: min cx, max cx, min cy, max cy
    uint32_t* bucket;      // [P]
    unsigned long long* vkey;  // [P]
    uint32_t* count;       // [nb]
    uint32_t* start;       // [nb + 1]
    uint32_t* cursor;      // [nb]
    unsigned long long* skey;  // [P] slot keys
    uint32_t* sidx;        // [P] slot particle
    uint32_t* sbucket;     // [P] slot bucket
    uint32_t* perm;        // [P]
    uint32_t* maxcount;    // [1]
};
void launch23_sort_extent(const Species23& s, double dx, double dy, Sort23& w, hipStream_t st);
void launch23_sort_buckets(const Species23& s, double dx, double dy, int64_t ry, int64_t nb, Sort23& w,
                           hipStream_t st);
void launch23_sort_order(const Species23& s, int64_t nb, Sort23& w, hipStream_t st);
void launch23_sort_apply(const Species23& s, const uint32_t* perm, double* const* dst, hipStream_t st);

}  // namespace pic
