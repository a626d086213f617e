// capi2d.cpp -- host runtime behind include/pic2d.h: the 2D3V context, its
// device buffers, the hipGraphs of one time step (with and without the
// diagnostics), and the operator calls.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/pic2d.h"
#include "pic2d_internal.h"
#include "trace.h"

using namespace pic;

struct pic2d_ctx {
    pic2d_config cfg{};
    hipStream_t stream = nullptr;
    Buffers2D b{};
    std::vector<void*> allocs;
    double* tw = nullptr;
    bool uploaded = false;
    int64_t step = 0;                // steps since upload
    // step graphs per buffer parity of the particle arrays
    hipGraphExec_t g_plain[2] = {nullptr, nullptr};
    hipGraphExec_t g_diag[2] = {nullptr, nullptr};
    int par = 0;                     // which buffer set holds the particles
    int64_t since_sort = 0;          // steps since the last tile sort
    uint32_t *xq[2] = {}, *yq[2] = {}, *pid[2] = {};
    float *vx[2] = {}, *vy[2] = {}, *vz[2] = {};
    float *vxl[2] = {}, *vyl[2] = {}, *vzl[2] = {};
    double *X64[2] = {}, *Y64[2] = {}, *VX64[2] = {}, *VY64[2] = {}, *VZ64[2] = {};
    // operator scratch (grown on demand)
    void* op[8] = {};
    size_t op_bytes[8] = {};
    std::string err;
};

static thread_local std::string g2_create_err;

static void set_err(pic2d_ctx* c, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    else g2_create_err = buf;
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

// P*w*2^q < 2^125 (the exact per-cell sums cannot wrap)
static int frac_bits(int64_t P, double w) {
    const double total = (double)P * std::fabs(w);
    if (!(total > 0.0) || !std::isfinite(total)) return 64;
    const int q = 124 - std::ilogb(total);
    return q > 1100 ? 1100 : q;
}

template <typename T>
static int dalloc(pic2d_ctx* c, T** p, size_t count) {
    void* q = nullptr;
    HIPCHK(c, hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T)));
    c->allocs.push_back(q);
    *p = (T*)q;
    return PIC_OK;
}

static int scratch(pic2d_ctx* c, int slot, size_t bytes, void** out) {
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

static void free_graphs(pic2d_ctx* c) {
    for (int k = 0; k < 2; ++k) {
        if (c->g_plain[k]) (void)hipGraphExecDestroy(c->g_plain[k]);
        if (c->g_diag[k]) (void)hipGraphExecDestroy(c->g_diag[k]);
        c->g_plain[k] = c->g_diag[k] = nullptr;
    }
}

// Points the particle arrays of the launch descriptors at buffer set par.
static void set_par(pic2d_ctx* c, int par) {
    Buffers2D& b = c->b;
    c->par = par;
    b.X = c->X64[par];
    b.Y = c->Y64[par];
    b.VX = c->VX64[par];
    b.VY = c->VY64[par];
    b.VZ = c->VZ64[par];
    b.X_alt = c->X64[1 - par];
    b.Y_alt = c->Y64[1 - par];
    b.VX_alt = c->VX64[1 - par];
    b.VY_alt = c->VY64[1 - par];
    b.VZ_alt = c->VZ64[1 - par];
    b.Xq = c->xq[par];
    b.Yq = c->yq[par];
    b.VXf = c->vx[par];
    b.VYf = c->vy[par];
    b.VZf = c->vz[par];
    b.VXl = c->vxl[par];
    b.VYl = c->vyl[par];
    b.VZl = c->vzl[par];
    b.pid = c->pid[par];
    b.Xq_alt = c->xq[1 - par];
    b.Yq_alt = c->yq[1 - par];
    b.VXf_alt = c->vx[1 - par];
    b.VYf_alt = c->vy[1 - par];
    b.VZf_alt = c->vz[1 - par];
    b.VXl_alt = c->vxl[1 - par];
    b.VYl_alt = c->vyl[1 - par];
    b.VZl_alt = c->vzl[1 - par];
    b.pid_alt = c->pid[1 - par];
}

// Regroups the particles by tile (into the other buffer set).
static int tile_sort(pic2d_ctx* c) {
    launch2d_tile_sort(c->b, c->stream);
    HIPCHK(c, hipGetLastError());
    set_par(c, 1 - c->par);
    c->since_sort = 0;
    return PIC_OK;
}

static void release(pic2d_ctx* c) {
    if (!c) return;
    free_graphs(c);
    for (void* p : c->allocs) (void)hipFree(p);
    for (void*& p : c->op)
        if (p) (void)hipFree(p);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

static int validate(const pic2d_config* g, std::string& why) {
    char buf[256];
    if (!pow2_in(g->NX, 8, kMax2D) || !pow2_in(g->NY, 8, kMax2D)) {
        snprintf(buf, sizeof buf, "NX, NY must be powers of two in [8, %d] (got %d x %d)", kMax2D, g->NX, g->NY);
        why = buf;
        return PIC_EINVAL;
    }
    if (g->precision != PIC_PREC_FP64 && g->precision != PIC_PREC_MIXED) {
        why = "precision must be PIC_PREC_FP64 or PIC_PREC_MIXED";
        return PIC_EINVAL;
    }
    const int64_t pmax = g->precision == PIC_PREC_MIXED ? kMaxP2D32 : kMaxP2D64;
    if (g->P < 1 || g->P > pmax) {
        snprintf(buf, sizeof buf, "P must be in [1, %lld] for this precision (got %lld)", (long long)pmax,
                 (long long)g->P);
        why = buf;
        return PIC_EINVAL;
    }
    if (g->pairing != PIC2D_PAIR_REFERENCE && g->pairing != PIC2D_PAIR_TEXTBOOK) {
        why = "pairing must be PIC2D_PAIR_REFERENCE or PIC2D_PAIR_TEXTBOOK";
        return PIC_EINVAL;
    }
    if (!(std::isfinite(g->dt) && g->dt > 0.0) || !(std::isfinite(g->w) && g->w > 0.0) || !std::isfinite(g->B0)) {
        why = "dt and w must be finite and positive, B0 finite";
        return PIC_EINVAL;
    }
    if (g->ns < 1) {
        why = "ns must be >= 1";
        return PIC_EINVAL;
    }
    return PIC_OK;
}

static bool in_unit_box(const double* x, const double* y, int64_t P, int64_t* bad) {
    for (int64_t p = 0; p < P; ++p)
        if (!(x[p] > 0.0 && x[p] <= 1.0 && y[p] > 0.0 && y[p] <= 1.0)) {
            *bad = p;
            return false;
        }
    return true;
}

#define CHECK_BOX(c, who, x, y, P)                                                            \
    do {                                                                                      \
        int64_t bad_ = 0;                                                                     \
        if (!in_unit_box(x, y, P, &bad_)) {                                                   \
            set_err(c, "%s: particle %lld lies outside (0, 1]^2", who, (long long)bad_);      \
            return PIC_EINVAL;                                                                \
        }                                                                                     \
    } while (0)

extern "C" int pic2d_create(pic2d_ctx** out, const pic2d_config* cfg) {
    PIC_TRACE("pic2d_create");
    if (!out || !cfg) {
        g2_create_err = "pic2d_create: null argument";
        return PIC_EINVAL;
    }
    *out = nullptr;
    std::string why;
    if (int rc = validate(cfg, why)) {
        g2_create_err = "pic2d_create: " + why;
        return rc;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || cfg->device < 0 || cfg->device >= ndev) {
        g2_create_err = "pic2d_create: no such HIP device";
        return PIC_EHIP;
    }
    pic2d_ctx* c = new pic2d_ctx();
    c->cfg = *cfg;
    auto fail = [&](int rc) {
        g2_create_err = c->err.empty() ? std::string("pic2d_create failed") : c->err;
        release(c);
        return rc;
    };
    if (hipSetDevice(cfg->device) != hipSuccess) return fail(PIC_EHIP);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return fail(PIC_EHIP);
    configure_kernels_2d();
    Buffers2D& b = c->b;
    const int NX = cfg->NX, NY = cfg->NY;
    const int64_t G = (int64_t)NX * NY, P = cfg->P;
    b.D.NX = NX;
    b.D.NY = NY;
    b.D.lx = ilog2i(NX);
    b.D.ly = ilog2i(NY);
    b.D.P = P;
    b.D.q = frac_bits(P, cfg->w);
    b.D.pairing = cfg->pairing;
    b.D.sx = 32 - b.D.lx;
    b.D.sy = 32 - b.D.ly;
    b.D.w = cfg->w;
    b.D.rho_scale = cfg->w * std::ldexp(1.0, -2 * kQuanta2D);
    b.D.kxs = kTwoPi / 1.0;  // the script's unit box: kx = 2pi*m
    b.D.kys = kTwoPi / 1.0;
    b.precision = cfg->precision;
    b.dt = cfg->dt;
    b.B0 = cfg->B0;
    {
        const int64_t bl = (P + kPush2DThreads - 1) / kPush2DThreads;
        b.blocks_push = (int)std::min<int64_t>(bl, 8192);
    }
    int rc;
    // both precisions hold the particles by tile (two buffer sets, the
    // caller's index of each held particle)
    for (int k = 0; k < 2; ++k) {
        if ((rc = dalloc(c, &c->pid[k], P))) return fail(rc);
        if (cfg->precision == PIC_PREC_MIXED) {
            if ((rc = dalloc(c, &c->xq[k], P)) || (rc = dalloc(c, &c->yq[k], P)) || (rc = dalloc(c, &c->vx[k], P)) ||
                (rc = dalloc(c, &c->vy[k], P)) || (rc = dalloc(c, &c->vz[k], P)) || (rc = dalloc(c, &c->vxl[k], P)) ||
                (rc = dalloc(c, &c->vyl[k], P)) || (rc = dalloc(c, &c->vzl[k], P)))
                return fail(rc);
        } else {
            if ((rc = dalloc(c, &c->X64[k], P)) || (rc = dalloc(c, &c->Y64[k], P)) ||
                (rc = dalloc(c, &c->VX64[k], P)) || (rc = dalloc(c, &c->VY64[k], P)) ||
                (rc = dalloc(c, &c->VZ64[k], P)))
                return fail(rc);
        }
    }
    if (cfg->precision != PIC_PREC_MIXED && (rc = dalloc(c, &b.slot, P))) return fail(rc);
    if ((rc = dalloc(c, &b.racc, (cfg->precision == PIC_PREC_MIXED ? 1 : 3) * G))) return fail(rc);
    set_par(c, 0);
    {
        Tiles2D& T = b.T;
        T.TX = std::min(kTile2D, NX);
        T.TY = std::min(kTile2D, NY);
        T.ntx = NX / T.TX;
        T.nty = NY / T.TY;
        T.ntiles = T.ntx * T.nty;
        const int64_t per_tile = (P + T.ntiles - 1) / T.ntiles;
        T.splits = (int)std::max<int64_t>(1, std::min<int64_t>(64, (per_tile + kTileBlockP - 1) / kTileBlockP));
        if ((rc = dalloc(c, &T.count, T.ntiles)) || (rc = dalloc(c, &T.start, T.ntiles + 1)) ||
            (rc = dalloc(c, &T.cursor, T.ntiles)))
            return fail(rc);
        if (hipMemset(T.count, 0, sizeof(int32_t) * T.ntiles) != hipSuccess) return fail(PIC_EHIP);
    }
    if ((rc = dalloc(c, &b.S, G)) || (rc = dalloc(c, &b.T1, G)) || (rc = dalloc(c, &b.T2, G)) ||
        (rc = dalloc(c, &b.T3, G)) || (rc = dalloc(c, &b.Ex, G)) || (rc = dalloc(c, &b.Ey, G)) ||
        (rc = dalloc(c, &b.phi, G)) || (rc = dalloc(c, &b.E32, G)) ||
        (rc = dalloc(c, &b.gpart, (G + kTreeChunk - 1) / kTreeChunk)) ||
        (rc = dalloc(c, &b.ppart, 6 * ((P + kTreeChunk - 1) / kTreeChunk))) || (rc = dalloc(c, &b.drow, 1)) ||
        (rc = dalloc(c, &b.flag, 1)))
        return fail(rc);
    const std::vector<double> twp = pass_twiddle_table();
    if ((rc = dalloc(c, &c->tw, twp.size()))) return fail(rc);
    if (hipMemcpy(c->tw, twp.data(), sizeof(double) * twp.size(), hipMemcpyHostToDevice) != hipSuccess)
        return fail(PIC_EHIP);
    b.tw = reinterpret_cast<const double2*>(c->tw);
    const size_t rb = (cfg->precision == PIC_PREC_MIXED ? 1 : 3) * (size_t)G * sizeof(unsigned long long);
    if (hipMemset(b.racc, 0, rb) != hipSuccess || hipMemset(b.flag, 0, sizeof(int32_t)) != hipSuccess ||
        hipMemset(b.Ex, 0, G * sizeof(double)) != hipSuccess || hipMemset(b.Ey, 0, G * sizeof(double)) != hipSuccess ||
        hipMemset(b.E32, 0, G * sizeof(float2)) != hipSuccess || hipDeviceSynchronize() != hipSuccess)
        return fail(PIC_EHIP);
    *out = c;
    return PIC_OK;
}

extern "C" int pic2d_destroy(pic2d_ctx* c) {
    PIC_TRACE("pic2d_destroy");
    if (!c) return PIC_EINVAL;
    (void)hipSetDevice(c->cfg.device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    release(c);
    return PIC_OK;
}

extern "C" const char* pic2d_last_error(const pic2d_ctx* c) {
    return c ? c->err.c_str() : g2_create_err.c_str();
}

// ---- state ------------------------------------------------------------------

extern "C" int pic2d_upload(pic2d_ctx* c, const double* x, const double* y, const double* vx, const double* vy,
                            const double* vz, const double* Ex, const double* Ey) {
    PIC_TRACE("pic2d_upload");
    if (!c) return PIC_EINVAL;
    if (!x || !y || !vx || !vy || !vz) {
        set_err(c, "pic2d_upload: particle arrays must not be NULL");
        return PIC_EINVAL;
    }
    HIPCHK(c, hipSetDevice(c->cfg.device));
    const int64_t P = c->cfg.P, G = (int64_t)c->cfg.NX * c->cfg.NY;
    CHECK_BOX(c, "pic2d_upload", x, y, P);
    Buffers2D& b = c->b;
    if (c->cfg.precision == PIC_PREC_MIXED) {
        std::vector<uint32_t> q(P);
        std::vector<float> f(P);
        auto fx = [&](const double* z) {
            for (int64_t p = 0; p < P; ++p) q[p] = (uint32_t)(uint64_t)std::llrint(std::ldexp(z[p], 32));
        };
        fx(x);
        HIPCHK(c, hipMemcpy(b.Xq, q.data(), 4 * P, hipMemcpyHostToDevice));
        fx(y);
        HIPCHK(c, hipMemcpy(b.Yq, q.data(), 4 * P, hipMemcpyHostToDevice));
        // v = hi + lo: hi = v rounded to FP32, lo = the rest rounded to FP32
        const double* vs[3] = {vx, vy, vz};
        float* ds[3] = {b.VXf, b.VYf, b.VZf};
        float* dl[3] = {b.VXl, b.VYl, b.VZl};
        std::vector<float> fl(P);
        for (int k = 0; k < 3; ++k) {
            for (int64_t p = 0; p < P; ++p) {
                f[p] = (float)vs[k][p];
                fl[p] = (float)(vs[k][p] - (double)f[p]);
            }
            HIPCHK(c, hipMemcpy(ds[k], f.data(), 4 * P, hipMemcpyHostToDevice));
            HIPCHK(c, hipMemcpy(dl[k], fl.data(), 4 * P, hipMemcpyHostToDevice));
        }
        for (int64_t p = 0; p < P; ++p) q[p] = (uint32_t)p;
        HIPCHK(c, hipMemcpy(b.pid, q.data(), 4 * P, hipMemcpyHostToDevice));
    } else {
        HIPCHK(c, hipMemcpy(b.X, x, 8 * P, hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(b.Y, y, 8 * P, hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(b.VX, vx, 8 * P, hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(b.VY, vy, 8 * P, hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(b.VZ, vz, 8 * P, hipMemcpyHostToDevice));
        std::vector<uint32_t> q(P);
        for (int64_t p = 0; p < P; ++p) q[p] = (uint32_t)p;
        HIPCHK(c, hipMemcpy(b.pid, q.data(), 4 * P, hipMemcpyHostToDevice));
    }
    std::vector<double> ex(G, 0.0), ey(G, 0.0);
    if (Ex) std::copy(Ex, Ex + G, ex.begin());
    if (Ey) std::copy(Ey, Ey + G, ey.begin());
    std::vector<float2> e32(G);
    for (int64_t k = 0; k < G; ++k) e32[k] = make_float2((float)ex[k], (float)ey[k]);
    HIPCHK(c, hipMemcpy(b.Ex, ex.data(), 8 * G, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(b.Ey, ey.data(), 8 * G, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(b.E32, e32.data(), sizeof(float2) * G, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemset(b.flag, 0, sizeof(int32_t)));
    {
        int rc = tile_sort(c);
        if (rc) return rc;
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    c->step = 0;
    c->uploaded = true;
    return PIC_OK;
}

extern "C" int pic2d_download(pic2d_ctx* c, double* x, double* y, double* vx, double* vy, double* vz, double* Ex,
                              double* Ey) {
    PIC_TRACE("pic2d_download");
    if (!c) return PIC_EINVAL;
    if (!c->uploaded) {
        set_err(c, "pic2d_download: nothing uploaded");
        return PIC_ESTATE;
    }
    HIPCHK(c, hipSetDevice(c->cfg.device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const int64_t P = c->cfg.P, G = (int64_t)c->cfg.NX * c->cfg.NY;
    Buffers2D& b = c->b;
    if (c->cfg.precision == PIC_PREC_MIXED) {
        // held particles are grouped by tile; pid gives each one's caller index
        std::vector<uint32_t> q(P), id(P);
        std::vector<float> f(P);
        HIPCHK(c, hipMemcpy(id.data(), b.pid, 4 * P, hipMemcpyDeviceToHost));
        auto tx = [&](uint32_t* src, double* dst) -> int {
            HIPCHK(c, hipMemcpy(q.data(), src, 4 * P, hipMemcpyDeviceToHost));
            for (int64_t p = 0; p < P; ++p) dst[id[p]] = q[p] ? std::ldexp((double)q[p], -32) : 1.0;
            return PIC_OK;
        };
        int rc;
        if (x && (rc = tx(b.Xq, x))) return rc;
        if (y && (rc = tx(b.Yq, y))) return rc;
        float* ss[3] = {b.VXf, b.VYf, b.VZf};
        float* sl[3] = {b.VXl, b.VYl, b.VZl};
        double* ds[3] = {vx, vy, vz};
        std::vector<float> fl(P);
        for (int k = 0; k < 3; ++k) {
            if (!ds[k]) continue;
            HIPCHK(c, hipMemcpy(f.data(), ss[k], 4 * P, hipMemcpyDeviceToHost));
            HIPCHK(c, hipMemcpy(fl.data(), sl[k], 4 * P, hipMemcpyDeviceToHost));
            for (int64_t p = 0; p < P; ++p) ds[k][id[p]] = (double)f[p] + (double)fl[p];
        }
    } else {
        // held by tile too: back to the caller's order through pid
        std::vector<uint32_t> id(P);
        std::vector<double> f(P);
        HIPCHK(c, hipMemcpy(id.data(), b.pid, 4 * P, hipMemcpyDeviceToHost));
        const double* ss[5] = {b.X, b.Y, b.VX, b.VY, b.VZ};
        double* ds[5] = {x, y, vx, vy, vz};
        for (int k = 0; k < 5; ++k) {
            if (!ds[k]) continue;
            HIPCHK(c, hipMemcpy(f.data(), ss[k], 8 * P, hipMemcpyDeviceToHost));
            for (int64_t p = 0; p < P; ++p) ds[k][id[p]] = f[p];
        }
    }
    if (Ex) HIPCHK(c, hipMemcpy(Ex, b.Ex, 8 * G, hipMemcpyDeviceToHost));
    if (Ey) HIPCHK(c, hipMemcpy(Ey, b.Ey, 8 * G, hipMemcpyDeviceToHost));
    return PIC_OK;
}

// ---- run --------------------------------------------------------------------

static void enqueue_step(pic2d_ctx* c, bool diag, hipStream_t s) {
    launch2d_push(c->b, s);
    launch2d_solve(c->b, c->cfg.precision == PIC_PREC_MIXED ? 1 : 0, nullptr, false, s);
    if (diag) launch2d_diag(c->b, s);
}

static int capture(pic2d_ctx* c, bool diag, hipGraphExec_t* out) {
    hipGraph_t g = nullptr;
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    enqueue_step(c, diag, c->stream);
    HIPCHK(c, hipStreamEndCapture(c->stream, &g));
    hipError_t e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIPCHK(c, e);
    return PIC_OK;
}

extern "C" int pic2d_run(pic2d_ctx* c, int64_t nsteps, double* K) {
    PIC_TRACE("pic2d_run");
    if (!c) return PIC_EINVAL;
    if (nsteps < 0) {
        set_err(c, "pic2d_run: nsteps must be >= 0");
        return PIC_EINVAL;
    }
    if (!c->uploaded) {
        set_err(c, "pic2d_run: upload particles first");
        return PIC_ESTATE;
    }
    HIPCHK(c, hipSetDevice(c->cfg.device));
    const int64_t ns = c->cfg.ns;
    const int64_t t0 = c->step, t1 = c->step + nsteps;
    const int64_t rows = t1 / ns - t0 / ns;
    Buffers2D& b = c->b;
    int rc;
    if (rows > b.diag_cap) {
        if (b.diag) {
            HIPCHK(c, hipFree(b.diag));
            b.diag = nullptr;
            b.diag_cap = 0;
        }
        HIPCHK(c, hipMalloc(&b.diag, sizeof(double) * 5 * (size_t)rows));
        b.diag_cap = rows;
        free_graphs(c);  // the graphs hold the old pointer
    }
    const int64_t zero = 0;
    HIPCHK(c, hipMemcpyAsync(b.drow, &zero, sizeof zero, hipMemcpyHostToDevice, c->stream));
    const int64_t every = c->cfg.sort_every > 0 ? c->cfg.sort_every : kSortEvery2D;
    for (int64_t t = t0 + 1; t <= t1; ++t) {
        if (c->since_sort >= every && (rc = tile_sort(c))) return rc;
        const int k = c->par;
        if (!c->g_plain[k] && (rc = capture(c, false, &c->g_plain[k]))) return rc;
        if (!c->g_diag[k] && (rc = capture(c, true, &c->g_diag[k]))) return rc;
        HIPCHK(c, hipGraphLaunch(t % ns == 0 ? c->g_diag[k] : c->g_plain[k], c->stream));
        ++c->since_sort;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->step = t1;
    if (K && rows > 0) HIPCHK(c, hipMemcpy(K, b.diag, sizeof(double) * 5 * rows, hipMemcpyDeviceToHost));
    int32_t flag = 0;
    HIPCHK(c, hipMemcpy(&flag, b.flag, sizeof flag, hipMemcpyDeviceToHost));
    if (flag) {
        set_err(c, "pic2d_run: the field became non-finite");
        return PIC_ENONFINITE;
    }
    return PIC_OK;
}

// ---- operators (FP64) -------------------------------------------------------

static int need_fp64_ops(pic2d_ctx* c, const char* who) {
    if (c->cfg.precision != PIC_PREC_FP64) {
        set_err(c, "%s: operators need a PIC_PREC_FP64 context", who);
        return PIC_EINVAL;
    }
    return PIC_OK;
}

static int h2d(pic2d_ctx* c, int slot, const double* src, int64_t n, double** dev) {
    void* p = nullptr;
    int rc = scratch(c, slot, sizeof(double) * (size_t)n, &p);
    if (rc) return rc;
    if (src) HIPCHK(c, hipMemcpyAsync(p, src, sizeof(double) * (size_t)n, hipMemcpyHostToDevice, c->stream));
    *dev = (double*)p;
    return PIC_OK;
}

static int d2h(pic2d_ctx* c, double* dst, const double* dev, int64_t n) {
    HIPCHK(c, hipMemcpyAsync(dst, dev, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost, c->stream));
    return PIC_OK;
}

static int finish(pic2d_ctx* c) {
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PIC_OK;
}

extern "C" int pic2d_deposit(pic2d_ctx* c, const double* x, const double* y, int64_t P, double w, double* rho) {
    PIC_TRACE("pic2d_deposit");
    if (!c) return PIC_EINVAL;
    if (int rc = need_fp64_ops(c, "pic2d_deposit")) return rc;
    if (!x || !y || !rho || P < 0 || P > kMaxP2D64 || !(std::isfinite(w) && w > 0.0)) {
        set_err(c, "pic2d_deposit: bad argument");
        return PIC_EINVAL;
    }
    CHECK_BOX(c, "pic2d_deposit", x, y, P);
    HIPCHK(c, hipSetDevice(c->cfg.device));
    const int64_t G = (int64_t)c->cfg.NX * c->cfg.NY;
    double *dx, *dy, *dr;
    int rc;
    if ((rc = h2d(c, 0, x, P, &dx)) || (rc = h2d(c, 1, y, P, &dy)) || (rc = h2d(c, 2, nullptr, G, &dr))) return rc;
    Buffers2D b = c->b;
    b.D.w = w;
    b.D.q = frac_bits(P, w);
    // the run's accumulator is free between runs (the solve leaves it cleared)
    if (c->cfg.precision == PIC_PREC_FP64) launch2d_deposit_op(b, dx, dy, P, dr, c->stream);
    if ((rc = d2h(c, rho, dr, G))) return rc;
    return finish(c);
}

extern "C" int pic2d_fieldsolve(pic2d_ctx* c, const double* rho, double* Ex, double* Ey, double* phi) {
    PIC_TRACE("pic2d_fieldsolve");
    if (!c) return PIC_EINVAL;
    if (int rc = need_fp64_ops(c, "pic2d_fieldsolve")) return rc;
    if (!rho || !Ex || !Ey) {
        set_err(c, "pic2d_fieldsolve: rho, Ex, Ey must not be NULL");
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
    if (flag) {
        set_err(c, "pic2d_fieldsolve: non-finite field");
        return PIC_ENONFINITE;
    }
    return PIC_OK;
}

extern "C" int pic2d_gather(pic2d_ctx* c, const double* Ex, const double* Ey, const double* x, const double* y,
                            int64_t P, double* ex, double* ey) {
    PIC_TRACE("pic2d_gather");
    if (!c) return PIC_EINVAL;
    if (int rc = need_fp64_ops(c, "pic2d_gather")) return rc;
    if (!Ex || !Ey || !x || !y || !ex || !ey || P < 0) {
        set_err(c, "pic2d_gather: bad argument");
        return PIC_EINVAL;
    }
    CHECK_BOX(c, "pic2d_gather", x, y, P);
    HIPCHK(c, hipSetDevice(c->cfg.device));
    const int64_t G = (int64_t)c->cfg.NX * c->cfg.NY;
    double *dEx, *dEy, *dx, *dy, *dex, *dey;
    int rc;
    if ((rc = h2d(c, 0, Ex, G, &dEx)) || (rc = h2d(c, 1, Ey, G, &dEy)) || (rc = h2d(c, 2, x, P, &dx)) ||
        (rc = h2d(c, 3, y, P, &dy)) || (rc = h2d(c, 4, nullptr, P, &dex)) || (rc = h2d(c, 5, nullptr, P, &dey)))
        return rc;
    if (P > 0) launch2d_gather_op(c->b, dEx, dEy, dx, dy, P, dex, dey, c->stream);
    if ((rc = d2h(c, ex, dex, P)) || (rc = d2h(c, ey, dey, P))) return rc;
    return finish(c);
}

extern "C" int pic2d_boris(pic2d_ctx* c, double* vx, double* vy, double* vz, const double* ex, const double* ey,
                           int64_t P, double dt, double B0) {
    PIC_TRACE("pic2d_boris");
    if (!c) return PIC_EINVAL;
    if (int rc = need_fp64_ops(c, "pic2d_boris")) return rc;
    if (!vx || !vy || !vz || !ex || !ey || P < 0) {
        set_err(c, "pic2d_boris: bad argument");
        return PIC_EINVAL;
    }
    HIPCHK(c, hipSetDevice(c->cfg.device));
    double *a, *b2, *cc, *dex, *dey;
    int rc;
    if ((rc = h2d(c, 0, vx, P, &a)) || (rc = h2d(c, 1, vy, P, &b2)) || (rc = h2d(c, 2, vz, P, &cc)) ||
        (rc = h2d(c, 3, ex, P, &dex)) || (rc = h2d(c, 4, ey, P, &dey)))
        return rc;
    if (P > 0) launch2d_boris_op(a, b2, cc, dex, dey, P, dt, B0, c->stream);
    if ((rc = d2h(c, vx, a, P)) || (rc = d2h(c, vy, b2, P)) || (rc = d2h(c, vz, cc, P))) return rc;
    return finish(c);
}

extern "C" int pic2d_push(pic2d_ctx* c, double* x, double* y, double* vx, double* vy, double* vz, int64_t P,
                          const double* Ex, const double* Ey, double dt, double B0) {
    PIC_TRACE("pic2d_push");
    if (!c) return PIC_EINVAL;
    if (int rc = need_fp64_ops(c, "pic2d_push")) return rc;
    if (!x || !y || !vx || !vy || !vz || !Ex || !Ey || P < 0) {
        set_err(c, "pic2d_push: bad argument");
        return PIC_EINVAL;
    }
    CHECK_BOX(c, "pic2d_push", x, y, P);
    HIPCHK(c, hipSetDevice(c->cfg.device));
    const int64_t G = (int64_t)c->cfg.NX * c->cfg.NY;
    double *dx, *dy, *dvx, *dvy, *dvz, *dEx, *dEy;
    int rc;
    if ((rc = h2d(c, 0, x, P, &dx)) || (rc = h2d(c, 1, y, P, &dy)) || (rc = h2d(c, 2, vx, P, &dvx)) ||
        (rc = h2d(c, 3, vy, P, &dvy)) || (rc = h2d(c, 4, vz, P, &dvz)) || (rc = h2d(c, 5, Ex, G, &dEx)) ||
        (rc = h2d(c, 6, Ey, G, &dEy)))
        return rc;
    if (P > 0) launch2d_push_op(c->b, dx, dy, dvx, dvy, dvz, P, dEx, dEy, dt, B0, c->stream);
    if ((rc = d2h(c, x, dx, P)) || (rc = d2h(c, y, dy, P)) || (rc = d2h(c, vx, dvx, P)) ||
        (rc = d2h(c, vy, dvy, P)) || (rc = d2h(c, vz, dvz, P)))
        return rc;
    return finish(c);
}

extern "C" int pic2d_diagnostics(pic2d_ctx* c, const double* Ex, const double* Ey, const double* vx, const double* vy,
                                 int64_t P, double w, double* K) {
    PIC_TRACE("pic2d_diagnostics");
    if (!c) return PIC_EINVAL;
    if (!Ex || !Ey || !vx || !vy || !K || P < 1) {
        set_err(c, "pic2d_diagnostics: bad argument");
        return PIC_EINVAL;
    }
    HIPCHK(c, hipSetDevice(c->cfg.device));
    const int64_t G = (int64_t)c->cfg.NX * c->cfg.NY;
    double *dEx, *dEy, *dvx, *dvy, *gp, *pp, *out;
    int rc;
    if ((rc = h2d(c, 0, Ex, G, &dEx)) || (rc = h2d(c, 1, Ey, G, &dEy)) || (rc = h2d(c, 2, vx, P, &dvx)) ||
        (rc = h2d(c, 3, vy, P, &dvy)) || (rc = h2d(c, 4, nullptr, (G + kTreeChunk - 1) / kTreeChunk, &gp)) ||
        (rc = h2d(c, 5, nullptr, 3 * ((P + kTreeChunk - 1) / kTreeChunk), &pp)) ||
        (rc = h2d(c, 6, nullptr, 5, &out)))
        return rc;
    launch2d_diag_op(dEx, dEy, G, dvx, dvy, P, w, gp, pp, out, nullptr, c->stream);
    if ((rc = d2h(c, K, out, 5))) return rc;
    return finish(c);
}
