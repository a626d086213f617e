// pic2d3v_lg_internal.h -- descriptors and launchers of the Lorenz-gauge
// electromagnetic path of the structured 2D3V module (LorenzGaugeStaggeredField
// LorenzGaugeField and LorenzGaugeSemiImplicitField of src/PIC2D3V.jl),
// shared by the host runtime (capi2d3v.cpp) and the kernels
// (kernels_2d3v_lg.hip).
#pragma once
#include "pic2d3v_internal.h"

namespace pic {

// Field state slots (complex NX x NY arrays), in the order of include/pic2d3v.h.
enum LGSlot {
    kLGPhi0 = 0, kLGPhiM, kLGAx0, kLGAy0, kLGAz0, kLGAxM, kLGAyM, kLGAzM,
    kLGRho0, kLGJx0, kLGJy0, kLGJz0, kLGEx, kLGEy, kLGEz, kLGBx, kLGBy, kLGBz, kLGSlots
};
// LorenzGaugeSemiImplicitField's sources at n-1 (warmup! at -dt), after the 18
enum { kLGRhoM = kLGSlots, kLGJxM, kLGJyM, kLGJzM, kLGSlotsSI };
constexpr int kLGBins = 14;      // binned fields of LorenzGaugeDiagnostics
constexpr int kLGRow = 11;       // doubles per diagnostics row
constexpr int kLGLinePoints = 2048;  // transform points per block (whole lines; 35 KB of LDS at 1024)

struct LG23 {
    int32_t NX, NY, lx, ly;
    double dt, theta;
    double B0[3];
    double2* F;          // kLGSlots x G: the field state
    double2* plus;       // 4 x G: phi+, Ax+, Ay+, Az+
    double* EB;          // 6 x G: the interior of EBxyz (0.0 + real(F), + B0 for B)
    const double* kx;    // NX wavenumbers (FFTHelper)
    const double* ky;    // NY
    unsigned long long* racc;  // 4 grids (rho, Jx, Jy, Jz) x 3 limbs x G
    unsigned long long* umax;  // kMaxSpecies23 x 3: max |v_c| per species (bits)
    int32_t ns;
    int32_t qr;          // fraction bits of the exact charge sums
    int64_t P[kMaxSpecies23];
    double qw[kMaxSpecies23];
    const double2* tw;
    double* bins;        // kLGBins x cap x NXd x NYd
    double* rows;        // cap x kLGRow
    int64_t cap;
    double* fpart;       // chunk partials of the field energy
    double* mfpart;      // 3 x chunk partials of the field momentum
    // LorenzGaugeSemiImplicitField: the deposit grid is the buffered
    // (NX+2b) x (NY+2b) one of the reference's offset arrays
    double theta_s;      // sourceimex theta
    double rtol;         // the convergence test's rtol (atol = 0)
    int32_t b, NXe, NYe;
    double2* src;        // 4 x G: rho+, Jx+, Jy+, Jz+
    unsigned long long* eacc;  // 4 components x 3 limbs x Ge: the ghost cells' exact sums (interior 0)
    double* gridT;       // 4 x Ge, component fastest: the grid before the iteration
    double* gridP;       // 4 x Ge: the iteration's grid
    double* npart;       // 6 x chunk partials of the three norms (sums, then maxima)
    double* res;         // norm(T - P), norm(T), norm(P), converged, stencil outside the buffer
    int32_t* eflag;      // set by a deposit whose stencil leaves the buffered grid
};

// step 1-2 of loop!: the wave update, E and B in the spectrum, their inverse
// transforms and update! (centred: LorenzGaugeField's, after the forward
// transforms of the sources)
void lg_launch_solve(const LG23& L, bool centred, hipStream_t st);
// update! alone, from the state's six fields
void lg_launch_update_eb(const LG23& L, hipStream_t st);
// step 3: gather, Boris, half move (xh, yh), half move (centred: one move);
// per-species maxima
void lg_launch_push(const LG23& L, const Grid23& g, const Species23& s, int sidx, double* xh, double* yh,
                    bool centred, hipStream_t st);
// the charge at (xh, yh) and the currents at (x, y)
void lg_launch_deposit_rho(const LG23& L, const Grid23& g, const Species23& s, const double* xh, const double* yh,
                           hipStream_t st);
void lg_launch_deposit_j(const LG23& L, const Grid23& g, const Species23& s, hipStream_t st);
// the reduction into rho0, J0 and the potentials' buffer copies (step 4;
// centred: the reduction alone)
void lg_launch_finish(const LG23& L, bool centred, hipStream_t st);
// warmup!: advect by h, the velocity maxima, and the reduction (rho0 = 0)
void lg_launch_advect(const LG23& L, const Grid23& g, const Species23& s, double h, hipStream_t st);
void lg_launch_vmax(const LG23& L, const Species23& s, int sidx, bool clear, hipStream_t st);
void lg_launch_warm_finish(const LG23& L, hipStream_t st);
// diagnose!: the row (when t % ntskip == 0), the in-place inverse transforms
// and the binned fields
void lg_launch_diag_row(const LG23& L, const Grid23& g, const Diag23& d, const Species23* sp, int ns,
                        hipStream_t st);
void lg_launch_diag_fields(const LG23& L, const Diag23& d, hipStream_t st);
// LorenzGaugeSemiImplicitField (loop! :993-1100, warmup! :864-871):
// the push from the step's start state s into the work arrays o
// (xw = unimod(x + vx*dt) with the old v, vw = Boris), per-species maxima
void lg_launch_si_push(const LG23& L, const Grid23& g, const Species23& s, const Species23& o, int sidx,
                       hipStream_t st);
// the exact deposits at s's positions into the periodic limbs racc and the
// ghost cells of the buffered grid (eacc): the charge and currents (charge)
// or the currents alone
void lg_launch_si_deposit(const LG23& L, const Grid23& g, const Species23& s, bool charge, hipStream_t st);
// the buffered grid into gridP (an interior cell's limbs are its periodic
// cell's minus its ghost images'; test: with isapprox(gridT, gridP) into
// res) and the periodic reduction into out[0..3] (imaginary parts 0); the
// limbs are cleared
struct LGOut {
    double2* a[4];
};
void lg_launch_si_reduce(const LG23& L, const LGOut& out, bool test, hipStream_t st);
// fft! of rho+, J+, the wave update with the three source weights, E and B
// in the spectrum, their inverse transforms and update!
void lg_launch_si_solve(const LG23& L, hipStream_t st);
// phi- = phi0, phi0 = phi+, A- = A0, A0 = A+
void lg_launch_rotate(const LG23& L, hipStream_t st);
void lg_launch_zero_umax(const LG23& L, hipStream_t st);
int configure_kernels_2d3v_lg();

}  // namespace pic
