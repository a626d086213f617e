// capi2d3v.cpp -- host runtime behind include/pic2d3v.h: the PIC2D3V context
// (species, an electrostatic or Lorenz-gauge field, diagnostics), its device
// buffers, the hipGraphs of one step (with and without the diagnostics row),
// the per-cell sort, warmup! and the operator calls.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/pic2d3v.h"
#include "pic2d3v_internal.h"
#include "trace.h"
#include "pic2d3v_lg_internal.h"

#pragma clang fp contract(off)

using namespace pic;

struct pic2d3v_ctx {
    pic2d3v_config cfg{};
    hipStream_t stream = nullptr;
    Grid23 g{};
    Species23 sp[kMaxSpecies23]{};
    bool uploaded[kMaxSpecies23] = {};
    // regrouping by tile: the other buffer set, the caller's index of each
    // held particle (and its other buffer), the tiles
    Species23 alt[kMaxSpecies23]{};
    uint32_t* pid[kMaxSpecies23] = {};
    uint32_t* pid_alt[kMaxSpecies23] = {};
    bool held_identity[kMaxSpecies23] = {};
    Tiles23 tiles{};
    int64_t since_regroup = 0;
    bool regroup_due = true;  // after an upload: regroup before the next step
    Diag23 d{};
    Buffers2D b{};           // the 2D solve's descriptor and buffers
    double2* E = nullptr;    // the gathered field Exy (interior, (Ex, Ey) per cell)
    // mixed precision: the FP32 copy of Exy the particles gather, the charge
    // density the solve reads
    bool mixed = false;
    float2* E32 = nullptr;
    double* rho = nullptr;
    std::vector<void*> allocs;
    double* tw = nullptr;
    int64_t step = 0;        // t of the next step
    // step graphs per buffer parity: a regroup exchanges every species' state
    // with its other buffer set and flips the parity, so the graphs of both
    // sets stay valid and are captured once each
    int parity = 0;
    hipGraphExec_t g_plain[2] = {}, g_diag[2] = {};
    // Lorenz gauge: the state, the half-step positions, and the step graphs
    // (loop! alone, with diagnose! without and with the row)
    bool lorenz = false;
    bool centred = false;    // LorenzGaugeField (else LorenzGaugeStaggeredField)
    bool semi = false;       // LorenzGaugeSemiImplicitField
    LG23 lg{};
    Species23 wk[kMaxSpecies23]{};  // semi-implicit: the work arrays xyvwork
    double* hres = nullptr;  // pinned: the iteration's test result
    std::vector<int64_t> iters;  // semi-implicit: iterations of the last call's steps
    double* xh[kMaxSpecies23] = {};
    double* yh[kMaxSpecies23] = {};
    hipGraphExec_t g_lg[2][3] = {};
    void* op[8] = {};        // scratch (grown on demand)
    size_t op_bytes[8] = {};
    std::string err;
};

static thread_local std::string g23_create_err;

static void set_err(pic2d3v_ctx* c, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    else g23_create_err = buf;
}

#define HIPCHK(ctx, expr)                                                                   \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess) {                                                             \
            set_err(ctx, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, \
                    __LINE__);                                                              \
            return e_ == hipErrorOutOfMemory ? PIC_ENOMEM : PIC_EHIP;                       \
        }                                                                                   \
    } while (0)

static int ilog2i(int n) {
    int l = 0;
    while ((1 << l) < n) ++l;
    return l;
}

static bool pow2_in(int n, int lo, int hi) { return n >= lo && n <= hi && (n & (n - 1)) == 0; }

// Fraction bits of the signed exact charge sums: the total |charge| of all
// species times 2^q stays below 2^111 (three 38-bit limbs hold 2^113).
static int frac_bits23(int ns, const int64_t* P, const double* w) {
    double total = 0.0;
    for (int s = 0; s < ns; ++s) total += (double)P[s] * std::fabs(w[s]);
    if (!(total > 0.0) || !std::isfinite(total)) return 64;
    const int q = 110 - std::ilogb(total);
    return q > 1100 ? 1100 : q;
}

template <typename T>
static int dalloc(pic2d3v_ctx* c, T** p, size_t count) {
    void* q = nullptr;
    HIPCHK(c, hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T)));
    c->allocs.push_back(q);
    *p = (T*)q;
    return PIC_OK;
}

static int scratch(pic2d3v_ctx* c, int slot, size_t bytes, void** out) {
    if (c->op_bytes[slot] < bytes || !c->op[slot]) {
        if (c->op[slot]) HIPCHK(c, hipFree(c->op[slot]));
        c->op[slot] = nullptr;
        c->op_bytes[slot] = 0;
        HIPCHK(c, hipMalloc(&c->op[slot], std::max<size_t>(bytes, 8)));
        c->op_bytes[slot] = bytes;
    }
    *out = c->op[slot];
    return PIC_OK;
}

static void free_graphs(pic2d3v_ctx* c) {
    for (int q = 0; q < 2; ++q) {
        for (hipGraphExec_t* g : {&c->g_plain[q], &c->g_diag[q], &c->g_lg[q][0], &c->g_lg[q][1], &c->g_lg[q][2]}) {
            if (*g) (void)hipGraphExecDestroy(*g);
            *g = nullptr;
        }
    }
}

static void free_bins(pic2d3v_ctx* c) {
    Diag23& d = c->d;
    for (double** p : {&d.rows, &d.Exs, &d.Eys, &d.phis, &c->lg.bins, &c->lg.rows}) {
        if (*p) (void)hipFree(*p);
        *p = nullptr;
    }
    d.cap = 0;
    c->lg.cap = 0;
}

static void release(pic2d3v_ctx* c) {
    if (!c) return;
    free_graphs(c);
    free_bins(c);
    for (void* p : c->allocs) (void)hipFree(p);
    if (c->hres) (void)hipHostFree(c->hres);
    for (void*& p : c->op)
        if (p) (void)hipFree(p);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

static int validate(const pic2d3v_config* g, std::string& why) {
    char buf[256];
    if (!pow2_in(g->NX, 8, kMax2D) || !pow2_in(g->NY, 8, kMax2D)) {
        snprintf(buf, sizeof buf, "NX, NY must be powers of two in [8, %d] (got %d x %d)", kMax2D, g->NX, g->NY);
        why = buf;
        return PIC_EINVAL;
    }
    if (!(std::isfinite(g->Lx) && g->Lx > 0.0 && std::isfinite(g->Ly) && g->Ly > 0.0)) {
        why = "Lx and Ly must be finite and positive";
        return PIC_EINVAL;
    }
    if (!(std::isfinite(g->dt) && g->dt > 0.0)) {
        why = "dt must be finite and positive";
        return PIC_EINVAL;
    }
    if (!(std::isfinite(g->B0[0]) && std::isfinite(g->B0[1]) && std::isfinite(g->B0[2]))) {
        why = "B0 must be finite";
        return PIC_EINVAL;
    }
    if (g->buffer < 0 || g->buffer > 1024) {
        why = "buffer must be in [0, 1024]";
        return PIC_EINVAL;
    }
    if (g->ntskip < 1) {
        why = "ntskip must be >= 1";
        return PIC_EINVAL;
    }
    if (!pow2_in(g->ngskip, 1, std::min(g->NX, g->NY))) {
        why = "ngskip must be a power of two no larger than NX and NY";
        return PIC_EINVAL;
    }
    if (g->field_update != PIC2D3V_FIELD_ACCUMULATE && g->field_update != PIC2D3V_FIELD_REPLACE) {
        why = "field_update must be PIC2D3V_FIELD_ACCUMULATE or PIC2D3V_FIELD_REPLACE";
        return PIC_EINVAL;
    }
    if (g->boris != PIC2D3V_BORIS_REFERENCE && g->boris != PIC2D3V_BORIS_TEXTBOOK) {
        why = "boris must be PIC2D3V_BORIS_REFERENCE or PIC2D3V_BORIS_TEXTBOOK";
        return PIC_EINVAL;
    }
    if (g->nspecies < 1 || g->nspecies > PIC2D3V_MAX_SPECIES) {
        snprintf(buf, sizeof buf, "nspecies must be in [1, %d]", PIC2D3V_MAX_SPECIES);
        why = buf;
        return PIC_EINVAL;
    }
    int64_t total = 0;
    for (int s = 0; s < g->nspecies; ++s) {
        const pic2d3v_species& q = g->species[s];
        if (q.P < 1 || q.P > kMaxP23) {
            snprintf(buf, sizeof buf, "species %d: P must be in [1, 2^26]", s);
            why = buf;
            return PIC_EINVAL;
        }
        total += q.P;
        if (q.shape < PIC2D3V_NGP || q.shape > PIC2D3V_BSPLINE5) {
            snprintf(buf, sizeof buf, "species %d: shape must be PIC2D3V_NGP, PIC2D3V_AREA or a B-spline", s);
            why = buf;
            return PIC_EINVAL;
        }
        if (!(std::isfinite(q.charge) && std::isfinite(q.mass) && q.mass != 0.0 && std::isfinite(q.weight) &&
              q.weight >= 0.0)) {
            snprintf(buf, sizeof buf, "species %d: charge, mass (nonzero) and weight (>= 0) must be finite", s);
            why = buf;
            return PIC_EINVAL;
        }
    }
    if (total > kMaxP23) {
        why = "the species hold more than 2^26 particles together";
        return PIC_EINVAL;
    }
    if (g->precision != PIC_PREC_FP64 && g->precision != PIC_PREC_MIXED) {
        why = "precision must be PIC_PREC_FP64 or PIC_PREC_MIXED";
        return PIC_EINVAL;
    }
    if (g->precision == PIC_PREC_MIXED && g->field != PIC2D3V_ELECTROSTATIC) {
        why = "mixed precision is provided for the electrostatic field";
        return PIC_EINVAL;
    }
    if (g->field != PIC2D3V_ELECTROSTATIC && g->field != PIC2D3V_LORENZ_STAGGERED && g->field != PIC2D3V_LORENZ &&
        g->field != PIC2D3V_LORENZ_SEMI_IMPLICIT) {
        why = "field must be PIC2D3V_ELECTROSTATIC, PIC2D3V_LORENZ_STAGGERED, PIC2D3V_LORENZ or "
              "PIC2D3V_LORENZ_SEMI_IMPLICIT";
        return PIC_EINVAL;
    }
    if (!std::isfinite(g->theta)) {
        why = "theta must be finite";
        return PIC_EINVAL;
    }
    if (g->field == PIC2D3V_LORENZ_SEMI_IMPLICIT) {
        if (!std::isfinite(g->theta_source) || !(std::isfinite(g->rtol) && g->rtol >= 0.0) || g->maxiters < 0) {
            why = "theta_source must be finite, rtol finite and >= 0, maxiters >= 0";
            return PIC_EINVAL;
        }
        // the deposit grid's axes -(b-1):NX+b hold every stencil cell of a
        // particle in (0, L] only for b >= the shape's upper reach
        for (int s = 0; s < g->nspecies; ++s) {
            const int sh = g->species[s].shape;
            const int n = sh - PIC2D3V_BSPLINE0;
            const int reach = sh == PIC2D3V_NGP ? 0 : (sh == PIC2D3V_AREA ? 1 : ((n & 1) ? (n + 1) / 2 : n / 2 + 1));
            if (g->buffer < reach) {
                snprintf(buf, sizeof buf, "species %d: the semi-implicit deposit grid needs buffer >= %d", s, reach);
                why = buf;
                return PIC_EINVAL;
            }
        }
    }
    return PIC_OK;
}

static int zero_state(pic2d3v_ctx* c) {
    const int64_t G = (int64_t)c->cfg.NX * c->cfg.NY;
    const int64_t zero = 0;
    HIPCHK(c, hipMemsetAsync(c->E, 0, sizeof(double2) * (size_t)G, c->stream));
    HIPCHK(c, hipMemsetAsync(c->b.racc, 0, 3 * sizeof(unsigned long long) * (size_t)G, c->stream));
    HIPCHK(c, hipMemsetAsync(c->b.flag, 0, sizeof(int32_t), c->stream));
    HIPCHK(c, hipMemsetAsync(c->g.fallback, 0, sizeof(unsigned long long), c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d.t, &zero, sizeof zero, hipMemcpyHostToDevice, c->stream));
    if (c->mixed) {
        HIPCHK(c, hipMemsetAsync(c->E32, 0, sizeof(float2) * (size_t)G, c->stream));
        for (int s = 0; s < c->cfg.nspecies; ++s)
            HIPCHK(c, hipMemsetAsync(c->sp[s].cnt, 0, 2 * sizeof(unsigned long long) * (size_t)G, c->stream));
    }
    if (c->lorenz) {
        LG23& L = c->lg;
        HIPCHK(c, hipMemsetAsync(L.F, 0, sizeof(double2) * (c->semi ? kLGSlotsSI : kLGSlots) * (size_t)G, c->stream));
        if (c->semi) {
            const size_t Ge = (size_t)L.NXe * L.NYe;
            HIPCHK(c, hipMemsetAsync(L.src, 0, sizeof(double2) * 4 * (size_t)G, c->stream));
            HIPCHK(c, hipMemsetAsync(L.eacc, 0, sizeof(unsigned long long) * 12 * Ge, c->stream));
            HIPCHK(c, hipMemsetAsync(L.gridT, 0, sizeof(double) * 4 * Ge, c->stream));
            HIPCHK(c, hipMemsetAsync(L.gridP, 0, sizeof(double) * 4 * Ge, c->stream));
            HIPCHK(c, hipMemsetAsync(L.eflag, 0, sizeof(int32_t), c->stream));
        }
        HIPCHK(c, hipMemsetAsync(L.plus, 0, sizeof(double2) * 4 * (size_t)G, c->stream));
        HIPCHK(c, hipMemsetAsync(L.EB, 0, sizeof(double) * 6 * (size_t)G, c->stream));
        HIPCHK(c, hipMemsetAsync(L.racc, 0, sizeof(unsigned long long) * 12 * (size_t)G, c->stream));
        HIPCHK(c, hipMemsetAsync(L.umax, 0, sizeof(unsigned long long) * 3 * kMaxSpecies23, c->stream));
        if (L.cap > 0) {
            HIPCHK(c, hipMemsetAsync(L.rows, 0, sizeof(double) * kLGRow * (size_t)L.cap, c->stream));
            HIPCHK(c, hipMemsetAsync(L.bins, 0, sizeof(double) * kLGBins * (size_t)L.cap * c->d.NXd * c->d.NYd,
                                     c->stream));
        }
    }
    Diag23& d = c->d;
    if (d.cap > 0) {
        const size_t nb = (size_t)d.cap * d.NXd * d.NYd;
        HIPCHK(c, hipMemsetAsync(d.rows, 0, sizeof(double) * 8 * (size_t)d.cap, c->stream));
        HIPCHK(c, hipMemsetAsync(d.Exs, 0, sizeof(double) * nb, c->stream));
        HIPCHK(c, hipMemsetAsync(d.Eys, 0, sizeof(double) * nb, c->stream));
        HIPCHK(c, hipMemsetAsync(d.phis, 0, sizeof(double) * nb, c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->step = 0;
    return PIC_OK;
}

// v as a float hi + lo pair
static void split_f(double v, float* hl) {
    hl[0] = (float)v;
    hl[1] = (float)(v - (double)hl[0]);
}

// The mixed state of a species (and of its regrouping buffers), its Boris
// constants and its charge counters.  Boris as a change, v' = v + 2e + d*bs
// with e = E*eh, m = v + e, d = (m + m x t) x t:
//   ElectrostaticBoris (:39-54): t = B dt/2, eh = q_m dt/2,
//     bs = 2 q_m^2 / (1 + q_m^2 |t|^2);
//   ElectromagneticBoris (:56-69): t = B q_m dt/2, eh = q_m dt/2,
//     bs = 2 / (1 + |t|^2).
static int alloc_mixed(pic2d3v_ctx* c, Species23& sp, Species23& al, int64_t G) {
    const int64_t P = sp.P;
    int rc;
    for (Species23* q : {&sp, &al}) {
        q->mixed = 1;
        if ((rc = dalloc(c, &q->xq, P)) || (rc = dalloc(c, &q->yq, P)) || (rc = dalloc(c, &q->vxh, P)) ||
            (rc = dalloc(c, &q->vyh, P)) || (rc = dalloc(c, &q->vzh, P)) || (rc = dalloc(c, &q->vxl, P)) ||
            (rc = dalloc(c, &q->vyl, P)) || (rc = dalloc(c, &q->vzl, P)))
            return rc;
    }
    if ((rc = dalloc(c, &sp.cnt, 2 * G))) return rc;
    al.cnt = sp.cnt;
    const double dt_2 = c->cfg.dt / 2.0, qm = sp.q_m;
    double t[3], tt = 0.0;
    for (int k = 0; k < 3; ++k) {
        t[k] = c->cfg.B0[k] * (c->cfg.boris == PIC2D3V_BORIS_TEXTBOOK ? dt_2 * qm : dt_2);
        tt += t[k] * t[k];
    }
    const double bs = c->cfg.boris == PIC2D3V_BORIS_TEXTBOOK ? 2.0 / (1.0 + tt) : 2.0 * qm * qm / (1.0 + qm * qm * tt);
    for (Species23* q : {&sp, &al}) {
        for (int k = 0; k < 3; ++k) q->bt[k] = (float)t[k];
        q->bs = (float)bs;
        q->eh = (float)(dt_2 * qm);
        q->qscale = sp.qw * std::ldexp(1.0, -2 * kQ23);
    }
    return PIC_OK;
}

// exchanges the mixed state arrays of a and b
static void swap_mixed(Species23& a, Species23& b) {
    std::swap(a.xq, b.xq);
    std::swap(a.yq, b.yq);
    std::swap(a.vxh, b.vxh);
    std::swap(a.vyh, b.vyh);
    std::swap(a.vzh, b.vzh);
    std::swap(a.vxl, b.vxl);
    std::swap(a.vyl, b.vyl);
    std::swap(a.vzl, b.vzl);
}

extern "C" int pic2d3v_create(pic2d3v_ctx** out, const pic2d3v_config* cfg) {
    PIC_TRACE("pic2d3v_create");
    if (!out || !cfg) {
        g23_create_err = "pic2d3v_create: null argument";
        return PIC_EINVAL;
    }
    *out = nullptr;
    std::string why;
    if (int rc = validate(cfg, why)) {
        g23_create_err = "pic2d3v_create: " + why;
        return rc;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || cfg->device < 0 || cfg->device >= ndev) {
        g23_create_err = "pic2d3v_create: no such HIP device";
        return PIC_EHIP;
    }
    pic2d3v_ctx* c = new pic2d3v_ctx();
    c->cfg = *cfg;
    auto fail = [&](int rc) {
        g23_create_err = c->err.empty() ? std::string("pic2d3v_create failed") : c->err;
        release(c);
        return rc;
    };
    if (hipSetDevice(cfg->device) != hipSuccess) return fail(PIC_EHIP);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return fail(PIC_EHIP);
    configure_kernels_2d();
    const int NX = cfg->NX, NY = cfg->NY, ns = cfg->nspecies;
    const int64_t G = (int64_t)NX * NY;
    c->mixed = cfg->precision == PIC_PREC_MIXED;
    Grid23& g = c->g;
    g.NX = NX;
    g.NY = NY;
    g.lx = ilog2i(NX);
    g.ly = ilog2i(NY);
    g.Lx = cfg->Lx;
    g.Ly = cfg->Ly;
    g.NX_Lx = (double)NX / cfg->Lx;
    g.NY_Ly = (double)NY / cfg->Ly;
    g.dt = cfg->dt;
    for (int k = 0; k < 3; ++k) g.B[k] = cfg->B0[k];
    g.boris = cfg->boris;
    g.sx = 32 - g.lx;
    g.sy = 32 - g.ly;
    split_f(cfg->dt * 4294967296.0 / cfg->Lx, g.stepx);
    split_f(cfg->dt * 4294967296.0 / cfg->Ly, g.stepy);
    // qw = (charge*weight)/dV with dV = (Lx/NX)*(Ly/NY) (loop!, :533, :537)
    const double dV = (cfg->Lx / (double)NX) * (cfg->Ly / (double)NY);
    int64_t Ps[kMaxSpecies23];
    double qws[kMaxSpecies23];
    int rc;
    for (int s = 0; s < ns; ++s) {
        const pic2d3v_species& q = cfg->species[s];
        Species23& sp = c->sp[s];
        sp.P = q.P;
        sp.shape = q.shape;
        sp.q_m = q.charge / q.mass;
        sp.qw = q.charge * q.weight / dV;
        sp.mass = q.mass;
        sp.weight = q.weight;
        Ps[s] = q.P;
        qws[s] = sp.qw;
        if ((rc = dalloc(c, &sp.x, q.P)) || (rc = dalloc(c, &sp.y, q.P)) || (rc = dalloc(c, &sp.vx, q.P)) ||
            (rc = dalloc(c, &sp.vy, q.P)) || (rc = dalloc(c, &sp.vz, q.P)))
            return fail(rc);
        for (double* a : {sp.x, sp.y, sp.vx, sp.vy, sp.vz})
            if (hipMemset(a, 0, sizeof(double) * (size_t)q.P) != hipSuccess) return fail(PIC_EHIP);
        Species23& al = c->alt[s];
        al = sp;
        if ((rc = dalloc(c, &al.x, q.P)) || (rc = dalloc(c, &al.y, q.P)) || (rc = dalloc(c, &al.vx, q.P)) ||
            (rc = dalloc(c, &al.vy, q.P)) || (rc = dalloc(c, &al.vz, q.P)) || (rc = dalloc(c, &sp.slot, q.P)) ||
            (rc = dalloc(c, &c->pid[s], q.P)) || (rc = dalloc(c, &c->pid_alt[s], q.P)))
            return fail(rc);
        al.slot = sp.slot;
        launch23_identity(c->pid[s], sp.slot, q.P, c->stream);
        c->held_identity[s] = true;
        if (cfg->precision == PIC_PREC_MIXED && (rc = alloc_mixed(c, sp, al, G))) return fail(rc);
    }
    {
        Tiles23& T = c->tiles;
        T.TX = std::min(kRegroupTile23, NX);
        T.TY = std::min(kRegroupTile23, NY);
        T.ntx = NX / T.TX;
        T.nty = NY / T.TY;
        T.ntiles = T.ntx * T.nty;
        if ((rc = dalloc(c, &T.count, T.ntiles)) || (rc = dalloc(c, &T.start, T.ntiles + 1)) ||
            (rc = dalloc(c, &T.cursor, T.ntiles)))
            return fail(rc);
        if (hipMemset(T.count, 0, sizeof(int32_t) * T.ntiles) != hipSuccess) return fail(PIC_EHIP);
    }
    g.q = frac_bits23(ns, Ps, qws);
    Buffers2D& b = c->b;
    b.D.NX = NX;
    b.D.NY = NY;
    b.D.lx = g.lx;
    b.D.ly = g.ly;
    b.D.P = 0;
    b.D.q = g.q;
    b.D.kxs = kTwoPi / cfg->Lx;  // kx = 2pi/Lx*m (FFTHelper, :252)
    b.D.kys = kTwoPi / cfg->Ly;
    b.precision = PIC_PREC_FP64;
    if ((rc = dalloc(c, &b.racc, 3 * G)) || (rc = dalloc(c, &b.S, G)) || (rc = dalloc(c, &b.T1, G)) ||
        (rc = dalloc(c, &b.T2, G)) || (rc = dalloc(c, &b.T3, G)) || (rc = dalloc(c, &b.Ex, G)) ||
        (rc = dalloc(c, &b.Ey, G)) || (rc = dalloc(c, &b.phi, G)) || (rc = dalloc(c, &b.flag, 1)) ||
        (rc = dalloc(c, &c->E, G)))
        return fail(rc);
    if (c->mixed && ((rc = dalloc(c, &c->E32, G)) || (rc = dalloc(c, &c->rho, G)))) return fail(rc);
    if ((rc = dalloc(c, &g.fallback, 1))) return fail(rc);
    b.E32 = nullptr;
    const std::vector<double> twp = pass_twiddle_table();
    if ((rc = dalloc(c, &c->tw, twp.size()))) return fail(rc);
    if (hipMemcpy(c->tw, twp.data(), sizeof(double) * twp.size(), hipMemcpyHostToDevice) != hipSuccess)
        return fail(PIC_EHIP);
    b.tw = reinterpret_cast<const double2*>(c->tw);
    Diag23& d = c->d;
    d.buffer = cfg->buffer;
    d.ntskip = cfg->ntskip;
    d.ngskip = cfg->ngskip;
    d.NXd = NX / cfg->ngskip;
    d.NYd = NY / cfg->ngskip;
    const int64_t nf = 2 * (int64_t)(NX + 2 * cfg->buffer) * (NY + 2 * cfg->buffer);
    if ((rc = dalloc(c, &d.t, 1)) || (rc = dalloc(c, &d.fpart, (nf + kTreeChunk - 1) / kTreeChunk))) return fail(rc);
    for (int s = 0; s < ns; ++s) {
        const int64_t P = cfg->species[s].P;
        if ((rc = dalloc(c, &d.kpart[s], (3 * P + kTreeChunk - 1) / kTreeChunk)) ||
            (rc = dalloc(c, &d.mpart[s], 6 * ((P + kTreeChunk - 1) / kTreeChunk))))
            return fail(rc);
    }
    if (cfg->field == PIC2D3V_LORENZ_STAGGERED || cfg->field == PIC2D3V_LORENZ ||
        cfg->field == PIC2D3V_LORENZ_SEMI_IMPLICIT) {
        c->lorenz = true;
        c->centred = cfg->field == PIC2D3V_LORENZ;
        c->semi = cfg->field == PIC2D3V_LORENZ_SEMI_IMPLICIT;
        configure_kernels_2d3v_lg();
        LG23& L = c->lg;
        L.NX = NX;
        L.NY = NY;
        L.lx = g.lx;
        L.ly = g.ly;
        L.dt = cfg->dt;
        L.theta = cfg->theta;
        for (int k = 0; k < 3; ++k) L.B0[k] = cfg->B0[k];
        L.ns = ns;
        L.qr = g.q;  // the charge's fraction bits: sum_s P_s |qw_s|, as the electrostatic path
        for (int s = 0; s < ns; ++s) {
            L.P[s] = Ps[s];
            L.qw[s] = qws[s];
        }
        L.tw = b.tw;
        double *kx = nullptr, *ky = nullptr;
        const int64_t nf6 = 6 * (int64_t)(NX + 2 * cfg->buffer) * (NY + 2 * cfg->buffer);
        if ((rc = dalloc(c, &L.F, (c->semi ? kLGSlotsSI : kLGSlots) * G)) || (rc = dalloc(c, &L.plus, 4 * G)) ||
            (rc = dalloc(c, &L.EB, 6 * G)) ||
            (rc = dalloc(c, &kx, NX)) || (rc = dalloc(c, &ky, NY)) || (rc = dalloc(c, &L.racc, 12 * G)) ||
            (rc = dalloc(c, &L.umax, 3 * kMaxSpecies23)) ||
            (rc = dalloc(c, &L.fpart, (nf6 + kTreeChunk - 1) / kTreeChunk)) ||
            (rc = dalloc(c, &L.mfpart, 3 * ((G + kTreeChunk - 1) / kTreeChunk))))
            return fail(rc);
        for (int s = 0; s < ns; ++s)
            if ((rc = dalloc(c, &c->xh[s], cfg->species[s].P)) || (rc = dalloc(c, &c->yh[s], cfg->species[s].P)))
                return fail(rc);
        if (c->semi) {
            L.theta_s = cfg->theta_source;
            L.rtol = cfg->rtol;
            L.b = cfg->buffer;
            L.NXe = NX + 2 * cfg->buffer;
            L.NYe = NY + 2 * cfg->buffer;
            const int64_t Ge = (int64_t)L.NXe * L.NYe;
            if ((rc = dalloc(c, &L.src, 4 * G)) || (rc = dalloc(c, &L.eacc, 12 * Ge)) ||
                (rc = dalloc(c, &L.gridT, 4 * Ge)) || (rc = dalloc(c, &L.gridP, 4 * Ge)) ||
                (rc = dalloc(c, &L.npart, 6 * ((4 * Ge + kTreeChunk - 1) / kTreeChunk))) ||
                (rc = dalloc(c, &L.res, 8)) || (rc = dalloc(c, &L.eflag, 1)))
                return fail(rc);
            for (int s = 0; s < ns; ++s) {
                Species23& w = c->wk[s];
                w = c->sp[s];
                const int64_t P = cfg->species[s].P;
                if ((rc = dalloc(c, &w.x, P)) || (rc = dalloc(c, &w.y, P)) || (rc = dalloc(c, &w.vx, P)) ||
                    (rc = dalloc(c, &w.vy, P)) || (rc = dalloc(c, &w.vz, P)))
                    return fail(rc);
            }
            if (hipHostMalloc((void**)&c->hres, 8 * sizeof(double), hipHostMallocDefault) != hipSuccess)
                return fail(PIC_ENOMEM);
        }
        // FFTHelper (:251-254): kx = 2pi/Lx * vcat(0:NX/2-1, -NX/2:-1), ky likewise
        std::vector<double> hx(NX), hy(NY);
        const double sx = kTwoPi / cfg->Lx, sy = kTwoPi / cfg->Ly;
        for (int i = 0; i < NX; ++i) hx[i] = sx * (double)(i < NX / 2 ? i : i - NX);
        for (int j = 0; j < NY; ++j) hy[j] = sy * (double)(j < NY / 2 ? j : j - NY);
        if (hipMemcpy(kx, hx.data(), sizeof(double) * NX, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(ky, hy.data(), sizeof(double) * NY, hipMemcpyHostToDevice) != hipSuccess)
            return fail(PIC_EHIP);
        L.kx = kx;
        L.ky = ky;
    }
    if ((rc = zero_state(c))) return fail(rc);
    *out = c;
    return PIC_OK;
}

extern "C" int pic2d3v_destroy(pic2d3v_ctx* c) {
    PIC_TRACE("pic2d3v_destroy");
    if (!c) return PIC_EINVAL;
    (void)hipSetDevice(c->cfg.device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    release(c);
    return PIC_OK;
}

extern "C" const char* pic2d3v_last_error(const pic2d3v_ctx* c) {
    return c ? c->err.c_str() : g23_create_err.c_str();
}

static int finish(pic2d3v_ctx* c) {
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PIC_OK;
}

static int check_species(pic2d3v_ctx* c, const char* who, int32_t s) {
    if (s < 0 || s >= c->cfg.nspecies) {
        set_err(c, "%s: no species %d", who, (int)s);
        return PIC_EINVAL;
    }
    return PIC_OK;
}

// ---- state ------------------------------------------------------------------

extern "C" int pic2d3v_upload(pic2d3v_ctx* c, int32_t s, const double* xyv) {
    PIC_TRACE("pic2d3v_upload");
    if (!c) return PIC_EINVAL;
    if (int rc = check_species(c, "pic2d3v_upload", s)) return rc;
    if (!xyv) {
        set_err(c, "pic2d3v_upload: xyv must not be NULL");
        return PIC_EINVAL;
    }
    HIPCHK(c, hipSetDevice(c->cfg.device));
    const Species23& sp = c->sp[s];
    void* st = nullptr;
    int rc;
    if ((rc = scratch(c, 0, sizeof(double) * 5 * (size_t)sp.P, &st))) return rc;
    HIPCHK(c, hipMemcpyAsync(st, xyv, sizeof(double) * 5 * (size_t)sp.P, hipMemcpyHostToDevice, c->stream));
    if (c->mixed) launch23m_unpack(c->g, (const double*)st, sp, c->stream);
    else launch23_unpack((const double*)st, sp, c->stream);
    launch23_identity(c->pid[s], sp.slot, sp.P, c->stream);
    if ((rc = finish(c))) return rc;
    c->uploaded[s] = true;
    c->held_identity[s] = true;
    c->regroup_due = true;
    return PIC_OK;
}

extern "C" int pic2d3v_download(pic2d3v_ctx* c, int32_t s, double* xyv) {
    PIC_TRACE("pic2d3v_download");
    if (!c) return PIC_EINVAL;
    if (int rc = check_species(c, "pic2d3v_download", s)) return rc;
    if (!xyv) {
        set_err(c, "pic2d3v_download: xyv must not be NULL");
        return PIC_EINVAL;
    }
    HIPCHK(c, hipSetDevice(c->cfg.device));
    const Species23& sp = c->sp[s];
    void* st = nullptr;
    int rc;
    if ((rc = scratch(c, 0, sizeof(double) * 5 * (size_t)sp.P, &st))) return rc;
    if (c->mixed) launch23m_pack(c->g, sp, (double*)st, c->stream);
    else launch23_pack(sp, (double*)st, c->stream);
    HIPCHK(c, hipMemcpyAsync(xyv, st, sizeof(double) * 5 * (size_t)sp.P, hipMemcpyDeviceToHost, c->stream));
    return finish(c);
}

extern "C" int pic2d3v_reset(pic2d3v_ctx* c) {
    PIC_TRACE("pic2d3v_reset");
    if (!c) return PIC_EINVAL;
    HIPCHK(c, hipSetDevice(c->cfg.device));
    return zero_state(c);
}

extern "C" int pic2d3v_field(pic2d3v_ctx* c, double* Ex, double* Ey) {
    PIC_TRACE("pic2d3v_field");
    if (!c) return PIC_EINVAL;
    HIPCHK(c, hipSetDevice(c->cfg.device));
    const int64_t G = (int64_t)c->cfg.NX * c->cfg.NY;
    if (c->lorenz) {
        if (Ex) HIPCHK(c, hipMemcpyAsync(Ex, c->lg.EB, sizeof(double) * (size_t)G, hipMemcpyDeviceToHost, c->stream));
        if (Ey)
            HIPCHK(c, hipMemcpyAsync(Ey, c->lg.EB + G, sizeof(double) * (size_t)G, hipMemcpyDeviceToHost, c->stream));
        return finish(c);
    }
    void *a = nullptr, *b = nullptr;
    int rc;
    if ((rc = scratch(c, 1, sizeof(double) * (size_t)G, &a)) || (rc = scratch(c, 2, sizeof(double) * (size_t)G, &b)))
        return rc;
    launch23_split_field(c->E, (double*)a, (double*)b, G, c->stream);
    if (Ex) HIPCHK(c, hipMemcpyAsync(Ex, a, sizeof(double) * (size_t)G, hipMemcpyDeviceToHost, c->stream));
    if (Ey) HIPCHK(c, hipMemcpyAsync(Ey, b, sizeof(double) * (size_t)G, hipMemcpyDeviceToHost, c->stream));
    return finish(c);
}

// ---- sort!(species, dx, dy) -------------------------------------------------

// The state arrays (FP64 or mixed) of `from` copied into those of `to`, on
// the context's stream.
static int copy_state(pic2d3v_ctx* c, const Species23& from, const Species23& to) {
    const size_t P = (size_t)from.P;
    if (c->mixed) {
        const void* src[8] = {from.xq, from.yq, from.vxh, from.vyh, from.vzh, from.vxl, from.vyl, from.vzl};
        void* dst[8] = {to.xq, to.yq, to.vxh, to.vyh, to.vzh, to.vxl, to.vyl, to.vzl};
        for (int k = 0; k < 8; ++k) HIPCHK(c, hipMemcpyAsync(dst[k], src[k], 4 * P, hipMemcpyDeviceToDevice, c->stream));
        return PIC_OK;
    }
    const void* src[5] = {from.x, from.y, from.vx, from.vy, from.vz};
    void* dst[5] = {to.x, to.y, to.vx, to.vy, to.vz};
    for (int k = 0; k < 5; ++k) HIPCHK(c, hipMemcpyAsync(dst[k], src[k], 8 * P, hipMemcpyDeviceToDevice, c->stream));
    return PIC_OK;
}

// Regroups every species' held particles by tile (see launch23_regroup) into
// the other buffer set and flips the parity of the step graphs.
static int regroup(pic2d3v_ctx* c) {
    for (int s = 0; s < c->cfg.nspecies; ++s) {
        Species23& sp = c->sp[s];
        Species23& al = c->alt[s];
        if (c->mixed) {
            launch23m_regroup(c->g, c->tiles, sp, al, c->pid[s], c->pid_alt[s], c->stream);
            swap_mixed(sp, al);
            std::swap(c->pid[s], c->pid_alt[s]);
            c->held_identity[s] = false;
            continue;
        }
        launch23_regroup(c->g, c->tiles, sp, al, c->pid[s], c->pid_alt[s], c->stream);
        std::swap(sp.x, al.x);
        std::swap(sp.y, al.y);
        std::swap(sp.vx, al.vx);
        std::swap(sp.vy, al.vy);
        std::swap(sp.vz, al.vz);
        std::swap(c->pid[s], c->pid_alt[s]);
        c->held_identity[s] = false;
    }
    HIPCHK(c, hipGetLastError());
    c->parity ^= 1;
    c->since_regroup = 0;
    c->regroup_due = false;
    HIPCHK(c, hipMemsetAsync(c->g.fallback, 0, sizeof(unsigned long long), c->stream));
    return PIC_OK;
}

// Before a step: regroup when an upload asked for it; otherwise every
// kRegroupEvery23 steps, on the Lorenz-gauge paths always, on the
// electrostatic path only when more than kRegroupFallback of the stencils
// pushed since the last check fell back to global memory (the pushes count
// them).
static int maybe_regroup(pic2d3v_ctx* c) {
    if (c->regroup_due) return regroup(c);
    if (c->since_regroup < kRegroupEvery23) return PIC_OK;
    if (!c->lorenz) {
        unsigned long long fb = 0;
        HIPCHK(c, hipMemcpyAsync(&fb, c->g.fallback, sizeof fb, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        int64_t P = 0;
        for (int s = 0; s < c->cfg.nspecies; ++s) P += c->sp[s].P;
        if ((double)fb <= kRegroupFallback * 2.0 * (double)P * (double)c->since_regroup) {
            HIPCHK(c, hipMemsetAsync(c->g.fallback, 0, sizeof fb, c->stream));
            c->since_regroup = 0;
            return PIC_OK;
        }
    }
    return regroup(c);
}

// The held particles of species s back in the caller's order (through the
// other buffer set and copied back, so the arrays the graphs hold stay).
static int to_caller(pic2d3v_ctx* c, int32_t s) {
    if (c->held_identity[s]) return PIC_OK;
    Species23& sp = c->sp[s];
    Species23& al = c->alt[s];
    if (c->mixed) launch23m_to_caller(sp, al, c->stream);
    else launch23_to_caller(sp, al, c->stream);
    int rc;
    if ((rc = copy_state(c, al, sp))) return rc;
    launch23_identity(c->pid[s], sp.slot, sp.P, c->stream);
    HIPCHK(c, hipGetLastError());
    c->held_identity[s] = true;
    return PIC_OK;
}

extern "C" int pic2d3v_sort(pic2d3v_ctx* c, int32_t s, double dx, double dy, int64_t* perm) {
    PIC_TRACE("pic2d3v_sort");
    if (!c) return PIC_EINVAL;
    if (int rc = check_species(c, "pic2d3v_sort", s)) return rc;
    if (!(std::isfinite(dx) && dx > 0.0 && std::isfinite(dy) && dy > 0.0)) {
        set_err(c, "pic2d3v_sort: dx and dy must be finite and positive");
        return PIC_EINVAL;
    }
    if (!c->uploaded[s]) {
        set_err(c, "pic2d3v_sort: upload species %d first", (int)s);
        return PIC_ESTATE;
    }
    HIPCHK(c, hipSetDevice(c->cfg.device));
    int rc;
    // sort!(species) permutes the caller's order: from that order
    if ((rc = to_caller(c, s))) return rc;
    // mixed precision: the keys from FP64 copies of the state, the
    // permutation applied to the mixed arrays
    if (c->mixed) launch23m_to64(c->g, c->sp[s], c->stream);
    const Species23& sp = c->sp[s];
    const int64_t P = sp.P;
    Sort23 w{};
    void* p = nullptr;
    // slot 3: ext, maxcount; 4: per-particle arrays; 5: per-bucket arrays; 0: applied particles
    if ((rc = scratch(c, 3, 64, &p))) return rc;
    w.ext = (int64_t*)p;
    w.maxcount = (uint32_t*)((char*)p + 32);
    const int64_t init[4] = {LLONG_MAX, LLONG_MIN, LLONG_MAX, LLONG_MIN};
    HIPCHK(c, hipMemcpyAsync(w.ext, init, sizeof init, hipMemcpyHostToDevice, c->stream));
    launch23_sort_extent(sp, dx, dy, w, c->stream);
    int64_t ext[4];
    HIPCHK(c, hipMemcpyAsync(ext, w.ext, sizeof ext, hipMemcpyDeviceToHost, c->stream));
    if ((rc = finish(c))) return rc;
    const int64_t rx = ext[1] - ext[0] + 1, ry = ext[3] - ext[2] + 1;
    if (!(rx >= 1 && ry >= 1 && rx <= kMaxSortBuckets && ry <= kMaxSortBuckets && rx * ry <= kMaxSortBuckets)) {
        set_err(c, "pic2d3v_sort: the particles span more than 2^24 cells of %g x %g", dx, dy);
        return PIC_EINVAL;
    }
    const int64_t nb = rx * ry;
    // per particle: bucket, vkey, skey, sidx, sbucket, perm
    const size_t per = (size_t)P;
    if ((rc = scratch(c, 4, per * (4 + 8 + 8 + 4 + 4 + 4), &p))) return rc;
    char* q = (char*)p;
    w.vkey = (unsigned long long*)q;
    w.skey = (unsigned long long*)(q + 8 * per);
    w.bucket = (uint32_t*)(q + 16 * per);
    w.sidx = (uint32_t*)(q + 20 * per);
    w.sbucket = (uint32_t*)(q + 24 * per);
    w.perm = (uint32_t*)(q + 28 * per);
    if ((rc = scratch(c, 5, sizeof(uint32_t) * (3 * (size_t)nb + 1), &p))) return rc;
    w.count = (uint32_t*)p;
    w.cursor = w.count + nb;
    w.start = w.cursor + nb;
    HIPCHK(c, hipMemsetAsync(w.count, 0, sizeof(uint32_t) * (size_t)nb, c->stream));
    launch23_sort_buckets(sp, dx, dy, ry, nb, w, c->stream);
    uint32_t mx = 0;
    HIPCHK(c, hipMemcpyAsync(&mx, w.maxcount, sizeof mx, hipMemcpyDeviceToHost, c->stream));
    if ((rc = finish(c))) return rc;
    if ((int64_t)mx > kMaxSortCell) {
        set_err(c, "pic2d3v_sort: %u particles share one cell (at most %lld)", mx, (long long)kMaxSortCell);
        return PIC_EINVAL;
    }
    launch23_sort_order(sp, nb, w, c->stream);
    if (c->mixed) {
        launch23m_permute(c->sp[s], c->alt[s], w.perm, c->stream);
        if ((rc = copy_state(c, c->alt[s], c->sp[s]))) return rc;
    } else {
        if ((rc = scratch(c, 0, sizeof(double) * 5 * per, &p))) return rc;
        double* dst[5];
        for (int k = 0; k < 5; ++k) dst[k] = (double*)p + k * per;
        launch23_sort_apply(sp, w.perm, dst, c->stream);
        double* cols[5] = {sp.x, sp.y, sp.vx, sp.vy, sp.vz};
        for (int k = 0; k < 5; ++k)
            HIPCHK(c, hipMemcpyAsync(cols[k], dst[k], sizeof(double) * per, hipMemcpyDeviceToDevice, c->stream));
    }
    std::vector<uint32_t> hp;
    if (perm) {
        hp.resize(per);
        HIPCHK(c, hipMemcpyAsync(hp.data(), w.perm, sizeof(uint32_t) * per, hipMemcpyDeviceToHost, c->stream));
    }
    if ((rc = finish(c))) return rc;
    for (size_t k = 0; k < hp.size(); ++k) perm[k] = hp[k];
    return PIC_OK;
}

// ---- run --------------------------------------------------------------------

// Room for bins 0 .. nbins-1 of the rows and the binned fields (grown
// geometrically, old bins kept, new ones zero).
static int ensure_bins_lg(pic2d3v_ctx* c, int64_t nbins) {
    LG23& L = c->lg;
    if (nbins <= L.cap) return PIC_OK;
    const int64_t cap = std::max<int64_t>(nbins, std::max<int64_t>(16, 2 * L.cap));
    const size_t per = (size_t)c->d.NXd * c->d.NYd;
    double *nr = nullptr, *nb = nullptr;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    hipError_t e1 = hipMalloc(&nr, sizeof(double) * kLGRow * (size_t)cap);
    hipError_t e2 = hipMalloc(&nb, sizeof(double) * kLGBins * per * (size_t)cap);
    if (e1 != hipSuccess || e2 != hipSuccess) {
        if (nr) (void)hipFree(nr);
        if (nb) (void)hipFree(nb);
        set_err(c, "pic2d3v_run: cannot allocate %lld field bins", (long long)cap);
        return PIC_ENOMEM;
    }
    HIPCHK(c, hipMemsetAsync(nr, 0, sizeof(double) * kLGRow * (size_t)cap, c->stream));
    HIPCHK(c, hipMemsetAsync(nb, 0, sizeof(double) * kLGBins * per * (size_t)cap, c->stream));
    if (L.cap > 0) {
        HIPCHK(c, hipMemcpyAsync(nr, L.rows, sizeof(double) * kLGRow * (size_t)L.cap, hipMemcpyDeviceToDevice,
                                 c->stream));
        for (int a = 0; a < kLGBins; ++a)
            HIPCHK(c, hipMemcpyAsync(nb + (size_t)a * cap * per, L.bins + (size_t)a * L.cap * per,
                                     sizeof(double) * per * (size_t)L.cap, hipMemcpyDeviceToDevice, c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (L.rows) HIPCHK(c, hipFree(L.rows));
    if (L.bins) HIPCHK(c, hipFree(L.bins));
    L.rows = nr;
    L.bins = nb;
    L.cap = cap;
    free_graphs(c);  // the graphs hold the old pointers
    return PIC_OK;
}

static int ensure_bins(pic2d3v_ctx* c, int64_t nbins) {
    if (c->lorenz) return ensure_bins_lg(c, nbins);
    Diag23& d = c->d;
    if (nbins <= d.cap) return PIC_OK;
    const int64_t cap = std::max<int64_t>(nbins, std::max<int64_t>(16, 2 * d.cap));
    const size_t per = (size_t)d.NXd * d.NYd;
    double* nr = nullptr;
    double* nf[3] = {nullptr, nullptr, nullptr};
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMalloc(&nr, sizeof(double) * 8 * (size_t)cap));
    for (double*& p : nf) {
        hipError_t e = hipMalloc(&p, sizeof(double) * per * (size_t)cap);
        if (e != hipSuccess) {
            (void)hipFree(nr);
            for (double* z : nf)
                if (z) (void)hipFree(z);
            set_err(c, "pic2d3v_run: cannot allocate %lld field bins", (long long)cap);
            return PIC_ENOMEM;
        }
    }
    HIPCHK(c, hipMemsetAsync(nr, 0, sizeof(double) * 8 * (size_t)cap, c->stream));
    for (double* p : nf) HIPCHK(c, hipMemsetAsync(p, 0, sizeof(double) * per * (size_t)cap, c->stream));
    if (d.cap > 0) {
        HIPCHK(c, hipMemcpyAsync(nr, d.rows, sizeof(double) * 8 * (size_t)d.cap, hipMemcpyDeviceToDevice, c->stream));
        double* old[3] = {d.Exs, d.Eys, d.phis};
        for (int k = 0; k < 3; ++k)
            HIPCHK(c, hipMemcpyAsync(nf[k], old[k], sizeof(double) * per * (size_t)d.cap, hipMemcpyDeviceToDevice,
                                     c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    free_bins(c);
    d.rows = nr;
    d.Exs = nf[0];
    d.Eys = nf[1];
    d.phis = nf[2];
    d.cap = cap;
    free_graphs(c);  // the graphs hold the old pointers
    return PIC_OK;
}

// One Lorenz-gauge step: loop! (staggered :875-979, centred :665-745) and,
// variant 1 or 2, diagnose! (variant 2 with the row).
static void enqueue_step_lg(pic2d3v_ctx* c, int variant, hipStream_t s) {
    const int ns = c->cfg.nspecies;
    const LG23& L = c->lg;
    const bool cen = c->centred;
    if (!cen) lg_launch_solve(L, false, s);
    for (int k = 0; k < ns; ++k) lg_launch_push(L, c->g, c->sp[k], k, c->xh[k], c->yh[k], cen, s);
    for (int k = 0; k < ns; ++k) {
        const Species23& sp = c->sp[k];
        lg_launch_deposit_rho(L, c->g, sp, cen ? sp.x : c->xh[k], cen ? sp.y : c->yh[k], s);
    }
    for (int k = 0; k < ns; ++k) lg_launch_deposit_j(L, c->g, c->sp[k], s);
    lg_launch_finish(L, cen, s);
    if (cen) lg_launch_solve(L, true, s);
    if (variant == 2) lg_launch_diag_row(L, c->g, c->d, c->sp, ns, s);
    if (variant >= 1) lg_launch_diag_fields(L, c->d, s);
    launch23_tick(c->d.t, s);
}

static void enqueue_step(pic2d3v_ctx* c, bool diag, hipStream_t s) {
    const int acc = c->cfg.field_update == PIC2D3V_FIELD_ACCUMULATE ? 1 : 0;
    if (c->mixed) {
        for (int k = 0; k < c->cfg.nspecies; ++k) launch23m_push(c->g, c->sp[k], c->E32, s);
        launch23m_rho(c->g, c->sp, c->cfg.nspecies, c->rho, s);
        launch2d_solve(c->b, 2, c->rho, true, s);
        launch23_update32(c->g, c->d, c->b, c->E, c->E32, acc, s);
    } else {
        for (int k = 0; k < c->cfg.nspecies; ++k) launch23_push(c->g, c->sp[k], c->E, c->b.racc, s);
        launch2d_solve(c->b, 3, nullptr, true, s);
        launch23_update(c->g, c->d, c->b, c->E, acc, s);
    }
    if (diag) launch23_diag(c->g, c->d, c->E, c->sp, c->cfg.nspecies, s);
    launch23_tick(c->d.t, s);
}

// variant: electrostatic 0/1 = without/with the row; Lorenz gauge 0 = loop!
// alone, 1/2 = with diagnose! without/with the row
static int capture(pic2d3v_ctx* c, int variant, hipGraphExec_t* out) {
    hipGraph_t g = nullptr;
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    if (c->lorenz) enqueue_step_lg(c, variant, c->stream);
    else enqueue_step(c, variant != 0, c->stream);
    HIPCHK(c, hipStreamEndCapture(c->stream, &g));
    hipError_t e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIPCHK(c, e);
    return PIC_OK;
}

static int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

// One LorenzGaugeSemiImplicitField step, loop! (:993-1100): iterations from
// the step's start state (the particles stay unchanged, the pushes go to the
// work arrays) until iteration k > maxiters or isapprox(grid before, grid
// after) (:1000), each read back here; then x, v = xw, vw (the arrays swap),
// the potentials' rotation, and diag 1/2 as in enqueue_step_lg.
static int step_si(pic2d3v_ctx* c, int diag) {
    LG23& L = c->lg;
    const int ns = c->cfg.nspecies;
    const int64_t G = (int64_t)c->cfg.NX * c->cfg.NY;
    hipStream_t st = c->stream;
    LGOut out{};
    for (int a = 0; a < 4; ++a) out.a[a] = L.src + (size_t)a * G;
    int64_t k = 0;
    for (;;) {
        ++k;
        lg_launch_zero_umax(L, st);
        for (int s = 0; s < ns; ++s) lg_launch_si_push(L, c->g, c->sp[s], c->wk[s], s, st);
        for (int s = 0; s < ns; ++s) lg_launch_si_deposit(L, c->g, c->wk[s], false, st);
        lg_launch_si_reduce(L, out, true, st);
        lg_launch_si_solve(L, st);
        std::swap(L.gridT, L.gridP);  // the iteration's grid is the next one's "before"
        HIPCHK(c, hipMemcpyAsync(c->hres, L.res, 5 * sizeof(double), hipMemcpyDeviceToHost, st));
        HIPCHK(c, hipStreamSynchronize(st));
        if (c->hres[4] != 0.0) {
            set_err(c, "a deposit stencil reached outside the buffered grid (the reference's BoundsError); "
                       "raise buffer");
            return PIC_EINVAL;
        }
        if (k > c->cfg.maxiters || c->hres[3] != 0.0) break;
    }
    c->iters.push_back(k);
    for (int s = 0; s < ns; ++s) {
        std::swap(c->sp[s].x, c->wk[s].x);
        std::swap(c->sp[s].y, c->wk[s].y);
        std::swap(c->sp[s].vx, c->wk[s].vx);
        std::swap(c->sp[s].vy, c->wk[s].vy);
        std::swap(c->sp[s].vz, c->wk[s].vz);
    }
    lg_launch_rotate(L, st);
    if (diag == 2) lg_launch_diag_row(L, c->g, c->d, c->sp, ns, st);
    if (diag >= 1) lg_launch_diag_fields(L, c->d, st);
    launch23_tick(c->d.t, st);
    HIPCHK(c, hipGetLastError());
    return PIC_OK;
}

extern "C" int pic2d3v_run(pic2d3v_ctx* c, int64_t nsteps, double* rows) {
    PIC_TRACE("pic2d3v_run");
    if (!c) return PIC_EINVAL;
    if (nsteps < 0) {
        set_err(c, "pic2d3v_run: nsteps must be >= 0");
        return PIC_EINVAL;
    }
    for (int s = 0; s < c->cfg.nspecies; ++s)
        if (!c->uploaded[s]) {
            set_err(c, "pic2d3v_run: upload species %d first", s);
            return PIC_ESTATE;
        }
    if (c->centred) {
        set_err(c, "pic2d3v_run: LorenzGaugeField has no diagnose! (the reference's reads field.rho0, which the "
                   "type lacks); use pic2d3v_loop");
        return PIC_ESTATE;
    }
    HIPCHK(c, hipSetDevice(c->cfg.device));
    const int64_t nt = c->cfg.ntskip;
    const int64_t t0 = c->step, t1 = c->step + nsteps;
    int rc;
    if ((rc = ensure_bins(c, cdiv(t1, nt)))) return rc;
    if (c->semi) {
        c->iters.clear();
        for (int64_t t = t0; t < t1; ++t) {
            if ((rc = maybe_regroup(c))) return rc;
            if ((rc = step_si(c, t % nt == 0 ? 2 : 1))) return rc;
            c->step = t + 1;
            ++c->since_regroup;
        }
    } else {
        for (int64_t t = t0; t < t1; ++t) {
            if ((rc = maybe_regroup(c))) return rc;
            ++c->since_regroup;
            const bool diag = t % nt == 0;
            const int q = c->parity;
            hipGraphExec_t& ge = c->lorenz ? c->g_lg[q][diag ? 2 : 1] : (diag ? c->g_diag[q] : c->g_plain[q]);
            if (!ge && (rc = capture(c, c->lorenz ? (diag ? 2 : 1) : (diag ? 1 : 0), &ge))) return rc;
            HIPCHK(c, hipGraphLaunch(ge, c->stream));
        }
    }
    if ((rc = finish(c))) return rc;
    c->step = t1;
    const int64_t r0 = cdiv(t0, nt), r1 = cdiv(t1, nt);
    const int rl = c->lorenz ? kLGRow : 8;
    const double* dr = c->lorenz ? c->lg.rows : c->d.rows;
    if (rows && r1 > r0)
        HIPCHK(c, hipMemcpy(rows, dr + rl * r0, sizeof(double) * rl * (size_t)(r1 - r0), hipMemcpyDeviceToHost));
    int32_t flag = 0;
    HIPCHK(c, hipMemcpy(&flag, c->b.flag, sizeof flag, hipMemcpyDeviceToHost));
    if (flag) {
        set_err(c, "pic2d3v_run: the field solve met a non-finite value");
        return PIC_ENONFINITE;
    }
    return PIC_OK;
}

extern "C" int pic2d3v_fields(pic2d3v_ctx* c, double* Exs, double* Eys, double* phis, int64_t nbins) {
    PIC_TRACE("pic2d3v_fields");
    if (!c) return PIC_EINVAL;
    // bins exist for the steps pic2d3v_run took (pic2d3v_loop writes none)
    const int64_t have = std::min<int64_t>(cdiv(c->step, c->cfg.ntskip), c->lorenz ? c->lg.cap : c->d.cap);
    if (nbins < 0 || nbins > have) {
        set_err(c, "pic2d3v_fields: nbins must be in [0, %lld]", (long long)have);
        return PIC_EINVAL;
    }
    if (nbins == 0) return PIC_OK;
    HIPCHK(c, hipSetDevice(c->cfg.device));
    const size_t n = sizeof(double) * (size_t)c->d.NXd * c->d.NYd * (size_t)nbins;
    if (c->lorenz) {
        const size_t blk = (size_t)c->d.NXd * c->d.NYd * (size_t)c->lg.cap;
        if (Exs) HIPCHK(c, hipMemcpy(Exs, c->lg.bins + PIC2D3V_LG_BIN_EX * blk, n, hipMemcpyDeviceToHost));
        if (Eys) HIPCHK(c, hipMemcpy(Eys, c->lg.bins + PIC2D3V_LG_BIN_EY * blk, n, hipMemcpyDeviceToHost));
        if (phis) HIPCHK(c, hipMemcpy(phis, c->lg.bins + PIC2D3V_LG_BIN_PHI * blk, n, hipMemcpyDeviceToHost));
        return PIC_OK;
    }
    if (Exs) HIPCHK(c, hipMemcpy(Exs, c->d.Exs, n, hipMemcpyDeviceToHost));
    if (Eys) HIPCHK(c, hipMemcpy(Eys, c->d.Eys, n, hipMemcpyDeviceToHost));
    if (phis) HIPCHK(c, hipMemcpy(phis, c->d.phis, n, hipMemcpyDeviceToHost));
    return PIC_OK;
}

// ---- Lorenz gauge: warmup!, loop! alone, the state, the bins ----------------

static int need_lorenz(pic2d3v_ctx* c, const char* who) {
    if (!c->lorenz) {
        set_err(c, "%s: the context's field is not a Lorenz-gauge field", who);
        return PIC_ESTATE;
    }
    return PIC_OK;
}

static int need_semi(pic2d3v_ctx* c, const char* who) {
    if (!c->semi) {
        set_err(c, "%s: the context's field is not PIC2D3V_LORENZ_SEMI_IMPLICIT", who);
        return PIC_ESTATE;
    }
    return PIC_OK;
}

static int need_uploaded(pic2d3v_ctx* c, const char* who) {
    for (int s = 0; s < c->cfg.nspecies; ++s)
        if (!c->uploaded[s]) {
            set_err(c, "%s: upload species %d first", who, s);
            return PIC_ESTATE;
        }
    return PIC_OK;
}

// warmup!(::LorenzGaugeSemiImplicitField) (:864-871): three passes of the
// four-weight warmup! (:746-775) -- advect by h, deposit the charge and
// currents, advect by -h -- into rho-, J- (h = -dt), rho0, J0 (+dt) and the
// grid rhoJs+ (+dt; its periodic reduction, rho+ and J+, is a temporary),
// then advect! by -dt.
static int warmup_si(pic2d3v_ctx* c) {
    LG23& L = c->lg;
    const int ns = c->cfg.nspecies;
    const int64_t G = (int64_t)c->cfg.NX * c->cfg.NY;
    const double dt = c->cfg.dt;
    hipStream_t st = c->stream;
    for (int s = 0; s < ns; ++s) lg_launch_vmax(L, c->sp[s], s, s == 0, st);
    for (int pass = 0; pass < 3; ++pass) {
        const double h = pass == 0 ? -dt : dt;
        const int first = pass == 0 ? kLGRhoM : kLGRho0;
        LGOut out{};
        for (int a = 0; a < 4; ++a) out.a[a] = pass == 2 ? L.src + (size_t)a * G : L.F + (size_t)(first + a) * G;
        for (int s = 0; s < ns; ++s) lg_launch_advect(L, c->g, c->sp[s], h, st);
        for (int s = 0; s < ns; ++s) lg_launch_si_deposit(L, c->g, c->sp[s], true, st);
        lg_launch_si_reduce(L, out, false, st);
        for (int s = 0; s < ns; ++s) lg_launch_advect(L, c->g, c->sp[s], -h, st);
    }
    std::swap(L.gridT, L.gridP);
    for (int s = 0; s < ns; ++s) lg_launch_advect(L, c->g, c->sp[s], -dt, st);
    int rc;
    if ((rc = finish(c))) return rc;
    int32_t flag = 0;
    HIPCHK(c, hipMemcpy(&flag, L.eflag, sizeof flag, hipMemcpyDeviceToHost));
    if (flag) {
        set_err(c, "pic2d3v_warmup: a deposit stencil reached outside the buffered grid (the reference's "
                   "BoundsError); raise buffer");
        return PIC_EINVAL;
    }
    return PIC_OK;
}

extern "C" int pic2d3v_warmup(pic2d3v_ctx* c) {
    PIC_TRACE("pic2d3v_warmup");
    if (!c) return PIC_EINVAL;
    if (!c->lorenz || c->centred) return PIC_OK;  // warmup!(::AbstractField) = field
    if (int rc = need_uploaded(c, "pic2d3v_warmup")) return rc;
    HIPCHK(c, hipSetDevice(c->cfg.device));
    if (c->semi) return warmup_si(c);
    const int ns = c->cfg.nspecies;
    const double h0 = -c->cfg.dt / 2.0, h1 = c->cfg.dt / 2.0;
    for (int k = 0; k < ns; ++k) lg_launch_advect(c->lg, c->g, c->sp[k], h0, c->stream);
    for (int k = 0; k < ns; ++k) lg_launch_advect(c->lg, c->g, c->sp[k], h1, c->stream);
    for (int k = 0; k < ns; ++k) lg_launch_vmax(c->lg, c->sp[k], k, k == 0, c->stream);
    for (int k = 0; k < ns; ++k) lg_launch_deposit_j(c->lg, c->g, c->sp[k], c->stream);
    lg_launch_warm_finish(c->lg, c->stream);
    return finish(c);
}

extern "C" int pic2d3v_loop(pic2d3v_ctx* c, int64_t nsteps) {
    PIC_TRACE("pic2d3v_loop");
    if (!c) return PIC_EINVAL;
    if (int rc = need_lorenz(c, "pic2d3v_loop")) return rc;
    if (nsteps < 0) {
        set_err(c, "pic2d3v_loop: nsteps must be >= 0");
        return PIC_EINVAL;
    }
    if (int rc = need_uploaded(c, "pic2d3v_loop")) return rc;
    HIPCHK(c, hipSetDevice(c->cfg.device));
    int rc;
    if (c->semi) {
        c->iters.clear();
        for (int64_t t = 0; t < nsteps; ++t) {
            if ((rc = maybe_regroup(c))) return rc;
            if ((rc = step_si(c, 0))) return rc;
            c->step += 1;
            ++c->since_regroup;
        }
        return finish(c);
    }
    for (int64_t t = 0; t < nsteps; ++t) {
        if ((rc = maybe_regroup(c))) return rc;
        ++c->since_regroup;
        hipGraphExec_t& ge = c->g_lg[c->parity][0];
        if (!ge && (rc = capture(c, 0, &ge))) return rc;
        HIPCHK(c, hipGraphLaunch(ge, c->stream));
    }
    if ((rc = finish(c))) return rc;
    c->step += nsteps;
    return PIC_OK;
}

static int check_slot(pic2d3v_ctx* c, const char* who, int32_t which, int32_t n) {
    if (which < 0 || which >= n) {
        set_err(c, "%s: no slot %d", who, (int)which);
        return PIC_EINVAL;
    }
    return PIC_OK;
}

extern "C" int pic2d3v_lg_get(pic2d3v_ctx* c, int32_t which, double* z) {
    PIC_TRACE("pic2d3v_lg_get");
    if (!c) return PIC_EINVAL;
    if (int rc = need_lorenz(c, "pic2d3v_lg_get")) return rc;
    if (int rc = check_slot(c, "pic2d3v_lg_get", which, c->semi ? PIC2D3V_LG_NSTATE_SI : PIC2D3V_LG_NSTATE))
        return rc;
    if (!z) {
        set_err(c, "pic2d3v_lg_get: z must not be NULL");
        return PIC_EINVAL;
    }
    HIPCHK(c, hipSetDevice(c->cfg.device));
    const size_t G = (size_t)c->cfg.NX * c->cfg.NY;
    HIPCHK(c, hipMemcpyAsync(z, c->lg.F + which * G, sizeof(double2) * G, hipMemcpyDeviceToHost, c->stream));
    return finish(c);
}

extern "C" int pic2d3v_lg_set(pic2d3v_ctx* c, int32_t which, const double* z) {
    PIC_TRACE("pic2d3v_lg_set");
    if (!c) return PIC_EINVAL;
    if (int rc = need_lorenz(c, "pic2d3v_lg_set")) return rc;
    if (int rc = check_slot(c, "pic2d3v_lg_set", which, c->semi ? PIC2D3V_LG_NSTATE_SI : PIC2D3V_LG_NSTATE))
        return rc;
    if (!z) {
        set_err(c, "pic2d3v_lg_set: z must not be NULL");
        return PIC_EINVAL;
    }
    HIPCHK(c, hipSetDevice(c->cfg.device));
    const size_t G = (size_t)c->cfg.NX * c->cfg.NY;
    HIPCHK(c, hipMemcpyAsync(c->lg.F + which * G, z, sizeof(double2) * G, hipMemcpyHostToDevice, c->stream));
    // the centred and semi-implicit loops gather EBxyz before they solve:
    // setting a field sets what the next step gathers
    if ((c->centred || c->semi) && which >= PIC2D3V_LG_EX && which <= PIC2D3V_LG_BZ)
        lg_launch_update_eb(c->lg, c->stream);
    return finish(c);
}

extern "C" int pic2d3v_lg_grid(pic2d3v_ctx* c, double* out) {
    PIC_TRACE("pic2d3v_lg_grid");
    if (!c) return PIC_EINVAL;
    if (int rc = need_semi(c, "pic2d3v_lg_grid")) return rc;
    if (!out) {
        set_err(c, "pic2d3v_lg_grid: out must not be NULL");
        return PIC_EINVAL;
    }
    HIPCHK(c, hipSetDevice(c->cfg.device));
    const size_t n = 4 * (size_t)c->lg.NXe * c->lg.NYe;
    HIPCHK(c, hipMemcpyAsync(out, c->lg.gridT, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    return finish(c);
}

extern "C" int pic2d3v_lg_set_grid(pic2d3v_ctx* c, const double* in) {
    PIC_TRACE("pic2d3v_lg_set_grid");
    if (!c) return PIC_EINVAL;
    if (int rc = need_semi(c, "pic2d3v_lg_set_grid")) return rc;
    if (!in) {
        set_err(c, "pic2d3v_lg_set_grid: in must not be NULL");
        return PIC_EINVAL;
    }
    HIPCHK(c, hipSetDevice(c->cfg.device));
    const size_t n = 4 * (size_t)c->lg.NXe * c->lg.NYe;
    HIPCHK(c, hipMemcpyAsync(c->lg.gridT, in, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    return finish(c);
}

extern "C" int pic2d3v_lg_iterations(pic2d3v_ctx* c, int64_t* out, int64_t n) {
    PIC_TRACE("pic2d3v_lg_iterations");
    if (!c) return PIC_EINVAL;
    if (int rc = need_semi(c, "pic2d3v_lg_iterations")) return rc;
    if (n < 0 || n > (int64_t)c->iters.size()) {
        set_err(c, "pic2d3v_lg_iterations: n must be in [0, %lld]", (long long)c->iters.size());
        return PIC_EINVAL;
    }
    if (n > 0 && !out) {
        set_err(c, "pic2d3v_lg_iterations: out must not be NULL");
        return PIC_EINVAL;
    }
    for (int64_t k = 0; k < n; ++k) out[k] = c->iters[c->iters.size() - (size_t)n + (size_t)k];
    return PIC_OK;
}

extern "C" int pic2d3v_lg_bins(pic2d3v_ctx* c, int32_t which, double* out, int64_t nbins) {
    PIC_TRACE("pic2d3v_lg_bins");
    if (!c) return PIC_EINVAL;
    if (int rc = need_lorenz(c, "pic2d3v_lg_bins")) return rc;
    if (int rc = check_slot(c, "pic2d3v_lg_bins", which, PIC2D3V_LG_NBINS)) return rc;
    const int64_t have = std::min<int64_t>(cdiv(c->step, c->cfg.ntskip), c->lg.cap);
    if (nbins < 0 || nbins > have) {
        set_err(c, "pic2d3v_lg_bins: nbins must be in [0, %lld]", (long long)have);
        return PIC_EINVAL;
    }
    if (nbins == 0) return PIC_OK;
    if (!out) {
        set_err(c, "pic2d3v_lg_bins: out must not be NULL");
        return PIC_EINVAL;
    }
    HIPCHK(c, hipSetDevice(c->cfg.device));
    const size_t per = (size_t)c->d.NXd * c->d.NYd;
    HIPCHK(c, hipMemcpy(out, c->lg.bins + (size_t)which * c->lg.cap * per, sizeof(double) * per * (size_t)nbins,
                        hipMemcpyDeviceToHost));
    return PIC_OK;
}

// ---- operators ----------------------------------------------------------------

static int h2d(pic2d3v_ctx* c, int slot, const double* src, int64_t n, double** dev) {
    void* p = nullptr;
    int rc = scratch(c, slot, sizeof(double) * (size_t)n, &p);
    if (rc) return rc;
    if (src && n > 0)
        HIPCHK(c, hipMemcpyAsync(p, src, sizeof(double) * (size_t)n, hipMemcpyHostToDevice, c->stream));
    *dev = (double*)p;
    return PIC_OK;
}

static int d2h(pic2d3v_ctx* c, double* dst, const double* dev, int64_t n) {
    if (n > 0) HIPCHK(c, hipMemcpyAsync(dst, dev, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost, c->stream));
    return PIC_OK;
}

static bool shape_ok(int32_t shape) { return shape >= PIC2D3V_NGP && shape <= PIC2D3V_BSPLINE5; }

extern "C" int pic2d3v_deposit(pic2d3v_ctx* c, int32_t shape, const double* x, const double* y, int64_t n, double w,
                               double* rho) {
    PIC_TRACE("pic2d3v_deposit");
    if (!c) return PIC_EINVAL;
    if (!shape_ok(shape) || !x || !y || !rho || n < 0 || n > kMaxP23 || !std::isfinite(w)) {
        set_err(c, "pic2d3v_deposit: bad argument");
        return PIC_EINVAL;
    }
    HIPCHK(c, hipSetDevice(c->cfg.device));
    const int64_t G = (int64_t)c->cfg.NX * c->cfg.NY;
    double *dx, *dy, *dr;
    int rc;
    if ((rc = h2d(c, 1, x, n, &dx)) || (rc = h2d(c, 2, y, n, &dy)) || (rc = h2d(c, 3, nullptr, G, &dr))) return rc;
    Grid23 g = c->g;
    g.q = frac_bits23(1, &n, &w);
    // the run's accumulator is clear between steps (the solve empties it)
    if (n > 0) launch23_deposit_op(g, shape, dx, dy, n, w, c->b.racc, c->stream);
    launch23_limbs_to_rho(g, c->b.racc, dr, c->stream);
    if ((rc = d2h(c, rho, dr, G))) return rc;
    return finish(c);
}

extern "C" int pic2d3v_gather(pic2d3v_ctx* c, int32_t shape, const double* Ex, const double* Ey, const double* x,
                              const double* y, int64_t n, double* ex, double* ey) {
    PIC_TRACE("pic2d3v_gather");
    if (!c) return PIC_EINVAL;
    if (!shape_ok(shape) || !Ex || !Ey || !x || !y || !ex || !ey || n < 0) {
        set_err(c, "pic2d3v_gather: bad argument");
        return PIC_EINVAL;
    }
    HIPCHK(c, hipSetDevice(c->cfg.device));
    const int64_t G = (int64_t)c->cfg.NX * c->cfg.NY;
    double *dEx, *dEy, *dx, *dy, *dex, *dey, *dE;
    int rc;
    if ((rc = h2d(c, 0, Ex, G, &dEx)) || (rc = h2d(c, 1, Ey, G, &dEy)) || (rc = h2d(c, 2, x, n, &dx)) ||
        (rc = h2d(c, 3, y, n, &dy)) || (rc = h2d(c, 4, nullptr, n, &dex)) || (rc = h2d(c, 5, nullptr, n, &dey)) ||
        (rc = h2d(c, 6, nullptr, 2 * G, &dE)))
        return rc;
    launch23_pair_field(dEx, dEy, (double2*)dE, G, c->stream);
    if (n > 0) launch23_gather_op(c->g, shape, (const double2*)dE, dx, dy, n, dex, dey, c->stream);
    if ((rc = d2h(c, ex, dex, n)) || (rc = d2h(c, ey, dey, n))) return rc;
    return finish(c);
}

extern "C" int pic2d3v_boris(pic2d3v_ctx* c, double* vx, double* vy, double* vz, const double* ex, const double* ey,
                             int64_t n, double q_m) {
    PIC_TRACE("pic2d3v_boris");
    if (!c) return PIC_EINVAL;
    if (!vx || !vy || !vz || !ex || !ey || n < 0 || !std::isfinite(q_m)) {
        set_err(c, "pic2d3v_boris: bad argument");
        return PIC_EINVAL;
    }
    HIPCHK(c, hipSetDevice(c->cfg.device));
    double *a, *b, *cc, *dex, *dey;
    int rc;
    if ((rc = h2d(c, 0, vx, n, &a)) || (rc = h2d(c, 1, vy, n, &b)) || (rc = h2d(c, 2, vz, n, &cc)) ||
        (rc = h2d(c, 3, ex, n, &dex)) || (rc = h2d(c, 4, ey, n, &dey)))
        return rc;
    if (n > 0) launch23_boris_op(c->g, q_m, a, b, cc, dex, dey, n, c->stream);
    if ((rc = d2h(c, vx, a, n)) || (rc = d2h(c, vy, b, n)) || (rc = d2h(c, vz, cc, n))) return rc;
    return finish(c);
}

extern "C" int pic2d3v_fieldsolve(pic2d3v_ctx* c, const double* rho, double* Ex, double* Ey, double* phi) {
    PIC_TRACE("pic2d3v_fieldsolve");
    if (!c) return PIC_EINVAL;
    if (!rho || !Ex || !Ey) {
        set_err(c, "pic2d3v_fieldsolve: rho, Ex, Ey must not be NULL");
        return PIC_EINVAL;
    }
    HIPCHK(c, hipSetDevice(c->cfg.device));
    const int64_t G = (int64_t)c->cfg.NX * c->cfg.NY;
    double *dr, *ex, *ey, *ph;
    int rc;
    if ((rc = h2d(c, 0, rho, G, &dr)) || (rc = h2d(c, 1, nullptr, G, &ex)) || (rc = h2d(c, 2, nullptr, G, &ey)) ||
        (rc = h2d(c, 3, nullptr, G, &ph)))
        return rc;
    Buffers2D b = c->b;
    b.Ex = ex;
    b.Ey = ey;
    b.phi = ph;
    b.E32 = nullptr;
    HIPCHK(c, hipMemsetAsync(b.flag, 0, sizeof(int32_t), c->stream));
    launch2d_solve(b, 2, dr, phi != nullptr, c->stream);
    if ((rc = d2h(c, Ex, ex, G)) || (rc = d2h(c, Ey, ey, G))) return rc;
    if (phi && (rc = d2h(c, phi, ph, G))) return rc;
    if ((rc = finish(c))) return rc;
    int32_t flag = 0;
    HIPCHK(c, hipMemcpy(&flag, b.flag, sizeof flag, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemset(b.flag, 0, sizeof(int32_t)));
    if (flag) {
        set_err(c, "pic2d3v_fieldsolve: non-finite field");
        return PIC_ENONFINITE;
    }
    return PIC_OK;
}
