// capi_em1d.cpp -- host runtime behind include/picem1d.h: the 1D
// electromagnetic context (fixed-point or explicit scheme), its device
// buffers, the hipGraph of one step (every fixed-point sweep predicated on the
// device), and the operator calls.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/picem1d.h"
#include "picem1d_internal.h"
#include "trace.h"

using namespace pic;

struct picem1d_ctx {
    picem1d_config cfg{};
    hipStream_t stream = nullptr;
    EM1D s{};
    std::vector<void*> allocs;
    double* tw = nullptr;
    bool uploaded = false;
    int64_t step = 0;
    hipGraphExec_t graph[2] = {nullptr, nullptr};  // without / with the moment fields
    void* op[12] = {};
    size_t op_bytes[12] = {};
    std::string err;
};

static thread_local std::string gem_create_err;

static void set_err(picem1d_ctx* c, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    else gem_create_err = buf;
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

template <typename T>
static int dalloc(picem1d_ctx* c, T** p, size_t count) {
    void* q = nullptr;
    HIPCHK(c, hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T)));
    c->allocs.push_back(q);
    *p = (T*)q;
    return PIC_OK;
}

static int scratch(picem1d_ctx* c, int slot, size_t bytes, void** out) {
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

static void free_graphs(picem1d_ctx* c) {
    for (auto& g : c->graph) {
        if (g) (void)hipGraphExecDestroy(g);
        g = nullptr;
    }
}

static void release(picem1d_ctx* c) {
    if (!c) return;
    free_graphs(c);
    if (c->s.out) (void)hipFree(c->s.out);
    if (c->s.sweeps_out) (void)hipFree(c->s.sweeps_out);
    for (void* p : c->allocs) (void)hipFree(p);
    for (void*& p : c->op)
        if (p) (void)hipFree(p);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

static int validate(const picem1d_config* g, std::string& why) {
    char buf[256];
    if (g->N < 8 || g->N > kMaxNEM || (g->N & (g->N - 1))) {
        snprintf(buf, sizeof buf, "N must be a power of two in [8, %d] (got %d)", kMaxNEM, g->N);
        why = buf;
        return PIC_EINVAL;
    }
    if (g->P < 1 || g->P > kMaxPEM) {
        why = "P must be in [1, 2^26]";
        return PIC_EINVAL;
    }
    if (g->scheme != PICEM1D_FIXED_POINT && g->scheme != PICEM1D_EXPLICIT) {
        why = "scheme must be PICEM1D_FIXED_POINT (0) or PICEM1D_EXPLICIT (1)";
        return PIC_EINVAL;
    }
    if (g->scheme == PICEM1D_FIXED_POINT && (g->max_sweeps < 1 || g->max_sweeps > kEMMaxSweeps)) {
        snprintf(buf, sizeof buf, "max_sweeps must be in [1, %d]", kEMMaxSweeps);
        why = buf;
        return PIC_EINVAL;
    }
    if (!(std::isfinite(g->eps) && g->eps > 0.0) || !(std::isfinite(g->dt) && g->dt > 0.0) || !std::isfinite(g->w) ||
        !(std::isfinite(g->rtol) && g->rtol >= 0.0)) {
        why = "eps and dt must be finite and positive, w finite, rtol finite and >= 0";
        return PIC_EINVAL;
    }
    if (g->sort != 0 && g->sort != 1) {
        why = "sort must be 0 or 1";
        return PIC_EINVAL;
    }
    return PIC_OK;
}

static int ilog2i(int n) {
    int l = 0;
    while ((1 << l) < n) ++l;
    return l;
}

extern "C" int picem1d_create(picem1d_ctx** out, const picem1d_config* cfg) {
    PIC_TRACE("picem1d_create");
    if (!out || !cfg) {
        gem_create_err = "picem1d_create: null argument";
        return PIC_EINVAL;
    }
    *out = nullptr;
    std::string why;
    if (int rc = validate(cfg, why)) {
        gem_create_err = "picem1d_create: " + why;
        return rc;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || cfg->device < 0 || cfg->device >= ndev) {
        gem_create_err = "picem1d_create: no such HIP device";
        return PIC_EHIP;
    }
    picem1d_ctx* c = new picem1d_ctx();
    c->cfg = *cfg;
    auto fail = [&](int rc) {
        gem_create_err = c->err.empty() ? std::string("picem1d_create failed") : c->err;
        release(c);
        return rc;
    };
    if (hipSetDevice(cfg->device) != hipSuccess) return fail(PIC_EHIP);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return fail(PIC_EHIP);
    configure_kernels_em1d();
    EM1D& s = c->s;
    const int N = cfg->N;
    const int64_t P = cfg->P;
    s.N = N;
    s.lN = ilog2i(N);
    s.max_sweeps = cfg->max_sweeps;
    s.P = P;
    s.eps = cfg->eps;
    s.dt = cfg->dt;
    s.w = cfg->w;
    s.rtol = cfg->rtol;
    s.tile = P <= (int64_t(1) << 18) ? 256 : 4096;
    int64_t nb = 1;
    while (nb < std::max<int64_t>(N, P / 16)) nb <<= 1;
    s.nb = (int32_t)std::min<int64_t>(nb, int64_t(1) << 20);
    int rc;
    double** parts[] = {&s.x, &s.v1, &s.v2, &s.v3, &s.X, &s.V1, &s.V2, &s.V3, &s.c};
    const int nparts = cfg->scheme == PICEM1D_EXPLICIT ? 5 : 9;  // the explicit step keeps X only
    for (int k = 0; k < nparts; ++k)
        if ((rc = dalloc(c, parts[k], P))) return fail(rc);
    double** grids[] = {&s.Ex, &s.Ey, &s.Ez, &s.By, &s.Bz, &s.Fx, &s.Fy, &s.Fz,
                        &s.Cy, &s.Cz, &s.ex, &s.ey, &s.ez, &s.by, &s.bz};
    for (double** p : grids)
        if ((rc = dalloc(c, p, N))) return fail(rc);
    if ((rc = dalloc(c, &s.racc, 30 * (size_t)N)) || (rc = dalloc(c, &s.umax, 8)) ||
        (rc = dalloc(c, &s.ctl, kEMMaxSweeps + 1)) || (rc = dalloc(c, &s.t, 2)))
        return fail(rc);
    if (cfg->sort) {
        const size_t B = (size_t)s.nb;
        if ((rc = dalloc(c, &s.count, B)) || (rc = dalloc(c, &s.cursor, B)) || (rc = dalloc(c, &s.start, B + 1)) ||
            (rc = dalloc(c, &s.bsum, 256)) || (rc = dalloc(c, &s.bucket, P)) || (rc = dalloc(c, &s.sidx, P)) ||
            (rc = dalloc(c, &s.sbucket, P)) || (rc = dalloc(c, &s.perm, P)) || (rc = dalloc(c, &s.vkey, P)) ||
            (rc = dalloc(c, &s.skey, P)) || (rc = dalloc(c, &s.tmp, 4 * (size_t)P)))
            return fail(rc);
    }
    const std::vector<double> twp = pass_twiddle_table();
    if ((rc = dalloc(c, &c->tw, twp.size()))) return fail(rc);
    if (hipMemcpy(c->tw, twp.data(), sizeof(double) * twp.size(), hipMemcpyHostToDevice) != hipSuccess)
        return fail(PIC_EHIP);
    s.tw = reinterpret_cast<const double2*>(c->tw);
    if (hipMemset(s.racc, 0, sizeof(unsigned long long) * 30 * (size_t)N) != hipSuccess ||
        hipMemset(s.umax, 0, sizeof(unsigned long long) * 8) != hipSuccess ||
        hipDeviceSynchronize() != hipSuccess)
        return fail(PIC_EHIP);
    *out = c;
    return PIC_OK;
}

extern "C" int picem1d_destroy(picem1d_ctx* c) {
    PIC_TRACE("picem1d_destroy");
    if (!c) return PIC_EINVAL;
    (void)hipSetDevice(c->cfg.device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    release(c);
    return PIC_OK;
}

extern "C" const char* picem1d_last_error(const picem1d_ctx* c) {
    return c ? c->err.c_str() : gem_create_err.c_str();
}

static int finish(picem1d_ctx* c) {
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PIC_OK;
}

// ---- state --------------------------------------------------------------------

extern "C" int picem1d_upload(picem1d_ctx* c, const double* x, const double* v, const double* E, const double* B) {
    PIC_TRACE("picem1d_upload");
    if (!c) return PIC_EINVAL;
    if (!x || !v) {
        set_err(c, "picem1d_upload: x and v must not be NULL");
        return PIC_EINVAL;
    }
    HIPCHK(c, hipSetDevice(c->cfg.device));
    EM1D& s = c->s;
    const int N = s.N;
    const size_t P = (size_t)s.P;
    HIPCHK(c, hipMemcpyAsync(s.x, x, 8 * P, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(s.v1, v, 8 * P, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(s.v2, v + P, 8 * P, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(s.v3, v + 2 * P, 8 * P, hipMemcpyHostToDevice, c->stream));
    std::vector<double> e(3 * (size_t)N, 0.0), b(2 * (size_t)N, 0.0);
    if (E) std::copy(E, E + 3 * N, e.begin());
    if (B) std::copy(B, B + 2 * N, b.begin());
    else std::fill(b.begin() + N, b.end(), 1.0);  // Bz = Z(N) .+ 1 (:6)
    HIPCHK(c, hipMemcpy(s.Ex, e.data(), 8 * (size_t)N, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(s.Ey, e.data() + N, 8 * (size_t)N, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(s.Ez, e.data() + 2 * N, 8 * (size_t)N, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(s.By, b.data(), 8 * (size_t)N, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(s.Bz, b.data() + N, 8 * (size_t)N, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemset(s.Fz, 0, 8 * (size_t)N));
    int32_t ctl[kEMMaxSweeps + 1] = {1};
    const int64_t t[2] = {0, 0};
    HIPCHK(c, hipMemcpy(s.ctl, ctl, sizeof ctl, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(s.t, t, sizeof t, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemset(s.umax, 0, sizeof(unsigned long long) * 8));
    if (int rc = finish(c)) return rc;
    c->step = 0;
    c->uploaded = true;
    return PIC_OK;
}

extern "C" int picem1d_download(picem1d_ctx* c, double* x, double* v, double* E, double* B) {
    PIC_TRACE("picem1d_download");
    if (!c) return PIC_EINVAL;
    if (!c->uploaded) {
        set_err(c, "picem1d_download: nothing uploaded");
        return PIC_ESTATE;
    }
    HIPCHK(c, hipSetDevice(c->cfg.device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    EM1D& s = c->s;
    const size_t P = (size_t)s.P, N = (size_t)s.N;
    if (x) HIPCHK(c, hipMemcpy(x, s.x, 8 * P, hipMemcpyDeviceToHost));
    if (v) {
        HIPCHK(c, hipMemcpy(v, s.v1, 8 * P, hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(v + P, s.v2, 8 * P, hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(v + 2 * P, s.v3, 8 * P, hipMemcpyDeviceToHost));
    }
    if (E) {
        HIPCHK(c, hipMemcpy(E, s.Ex, 8 * N, hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(E + N, s.Ey, 8 * N, hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(E + 2 * N, s.Ez, 8 * N, hipMemcpyDeviceToHost));
    }
    if (B) {
        HIPCHK(c, hipMemcpy(B, s.By, 8 * N, hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(B + N, s.Bz, 8 * N, hipMemcpyDeviceToHost));
    }
    return PIC_OK;
}

// ---- run ------------------------------------------------------------------------

static int capture(picem1d_ctx* c, bool moments, hipGraphExec_t* out) {
    hipGraph_t g = nullptr;
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    if (c->cfg.scheme == PICEM1D_EXPLICIT) {
        emx_launch_step(c->s, c->cfg.sort != 0, moments, c->stream);
    } else {
        for (int sw = 0; sw < c->s.max_sweeps; ++sw) em_launch_sweep(c->s, sw, c->stream);
        em_launch_end(c->s, c->cfg.sort != 0, moments, c->stream);
    }
    HIPCHK(c, hipStreamEndCapture(c->stream, &g));
    hipError_t e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIPCHK(c, e);
    return PIC_OK;
}

extern "C" int picem1d_run(picem1d_ctx* c, int64_t nsteps, double* fields, int32_t* sweeps) {
    PIC_TRACE("picem1d_run");
    if (!c) return PIC_EINVAL;
    if (nsteps < 0) {
        set_err(c, "picem1d_run: nsteps must be >= 0");
        return PIC_EINVAL;
    }
    if (!c->uploaded) {
        set_err(c, "picem1d_run: upload first");
        return PIC_ESTATE;
    }
    if (nsteps == 0) return PIC_OK;
    HIPCHK(c, hipSetDevice(c->cfg.device));
    EM1D& s = c->s;
    const int N = s.N;
    if (nsteps > s.out_cap) {
        free_graphs(c);
        if (s.out) HIPCHK(c, hipFree(s.out));
        if (s.sweeps_out) HIPCHK(c, hipFree(s.sweeps_out));
        s.out = nullptr;
        s.sweeps_out = nullptr;
        s.out_cap = 0;
        HIPCHK(c, hipMalloc(&s.out, sizeof(double) * 12 * (size_t)N * (size_t)nsteps));
        HIPCHK(c, hipMalloc(&s.sweeps_out, sizeof(int32_t) * (size_t)nsteps));
        s.out_cap = nsteps;
    }
    const bool moments = fields != nullptr;
    hipGraphExec_t& g = c->graph[moments ? 1 : 0];
    int rc;
    if (!g && (rc = capture(c, moments, &g))) return rc;
    const int64_t base = c->step;
    HIPCHK(c, hipMemcpyAsync(s.t + 1, &base, sizeof base, hipMemcpyHostToDevice, c->stream));
    for (int64_t k = 0; k < nsteps; ++k) HIPCHK(c, hipGraphLaunch(g, c->stream));
    if ((rc = finish(c))) return rc;
    c->step += nsteps;
    if (sweeps) HIPCHK(c, hipMemcpy(sweeps, s.sweeps_out, sizeof(int32_t) * (size_t)nsteps, hipMemcpyDeviceToHost));
    if (fields) {
        // device slots are [step][12][N]; the caller's layout is N x nsteps x 12
        std::vector<double> h(12 * (size_t)N * (size_t)nsteps);
        HIPCHK(c, hipMemcpy(h.data(), s.out, sizeof(double) * h.size(), hipMemcpyDeviceToHost));
        for (int64_t t = 0; t < nsteps; ++t)
            for (int k = 0; k < 12; ++k)
                std::memcpy(fields + (size_t)N * ((size_t)t + (size_t)nsteps * k), h.data() + ((size_t)t * 12 + k) * N,
                            sizeof(double) * (size_t)N);
    }
    return PIC_OK;
}

// ---- operators --------------------------------------------------------------------

static int h2d(picem1d_ctx* c, int slot, const double* src, int64_t n, double** dev) {
    void* p = nullptr;
    int rc = scratch(c, slot, sizeof(double) * (size_t)std::max<int64_t>(n, 1), &p);
    if (rc) return rc;
    if (src && n > 0)
        HIPCHK(c, hipMemcpyAsync(p, src, sizeof(double) * (size_t)n, hipMemcpyHostToDevice, c->stream));
    *dev = (double*)p;
    return PIC_OK;
}

static int d2h(picem1d_ctx* c, double* dst, const double* dev, int64_t n) {
    if (n > 0) HIPCHK(c, hipMemcpyAsync(dst, dev, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost, c->stream));
    return PIC_OK;
}

extern "C" int picem1d_deposit(picem1d_ctx* c, const double* y, const double* u, int64_t n, double w, double* rho) {
    PIC_TRACE("picem1d_deposit");
    if (!c) return PIC_EINVAL;
    if (!y || !rho || n < 0 || n > kMaxPEM || !std::isfinite(w)) {
        set_err(c, "picem1d_deposit: bad argument");
        return PIC_EINVAL;
    }
    HIPCHK(c, hipSetDevice(c->cfg.device));
    const int N = c->s.N;
    double *dy, *du = nullptr, *dr;
    void* m = nullptr;
    int rc;
    if ((rc = h2d(c, 0, y, n, &dy)) || (u && (rc = h2d(c, 1, u, n, &du))) || (rc = h2d(c, 2, nullptr, N, &dr)) ||
        (rc = scratch(c, 3, 64, &m)))
        return rc;
    HIPCHK(c, hipMemsetAsync(m, 0, 64, c->stream));
    em_launch_deposit_op(c->s, dy, du, n, w, dr, (unsigned long long*)m, c->stream);
    if ((rc = d2h(c, rho, dr, N))) return rc;
    return finish(c);
}

extern "C" int picem1d_fields(picem1d_ctx* c, const double* rho, const double* Jy, const double* Jz, const double* Fx,
                              const double* Fy, const double* Fz, const double* Cy, const double* Cz, double* E,
                              double* B, double* e3, double* b2) {
    PIC_TRACE("picem1d_fields");
    if (!c) return PIC_EINVAL;
    if (!rho || !Jy || !Jz || !Fx || !Fy || !Fz || !Cy || !Cz || !E || !B || !e3 || !b2) {
        set_err(c, "picem1d_fields: arguments must not be NULL");
        return PIC_EINVAL;
    }
    HIPCHK(c, hipSetDevice(c->cfg.device));
    const int N = c->s.N;
    double *dr, *djy, *djz, *buf;
    int rc;
    // slot 6: Ex Ey Ez By Bz Fx Fy Fz Cy Cz ex ey ez by bz (15 N)
    if ((rc = h2d(c, 4, rho, N, &dr)) || (rc = h2d(c, 5, Jy, N, &djy)) || (rc = h2d(c, 7, Jz, N, &djz)) ||
        (rc = h2d(c, 6, nullptr, 15 * (int64_t)N, &buf)))
        return rc;
    EM1D s = c->s;
    double** f[15] = {&s.Ex, &s.Ey, &s.Ez, &s.By, &s.Bz, &s.Fx, &s.Fy, &s.Fz, &s.Cy, &s.Cz,
                      &s.ex, &s.ey, &s.ez, &s.by, &s.bz};
    for (int k = 0; k < 15; ++k) *f[k] = buf + (size_t)k * N;
    const double* src[10] = {E, E + N, E + 2 * N, B, B + N, Fx, Fy, Fz, Cy, Cz};
    for (int k = 0; k < 10; ++k)
        HIPCHK(c, hipMemcpyAsync(*f[k], src[k], 8 * (size_t)N, hipMemcpyHostToDevice, c->stream));
    em_launch_fields_op(s, dr, djy, djz, c->stream);
    double* dst[10] = {E, E + N, E + 2 * N, B, B + N, e3, e3 + N, e3 + 2 * N, b2, b2 + N};
    const double* from[10] = {s.Ex, s.Ey, s.Ez, s.By, s.Bz, s.ex, s.ey, s.ez, s.by, s.bz};
    for (int k = 0; k < 10; ++k)
        if ((rc = d2h(c, dst[k], from[k], N))) return rc;
    return finish(c);
}

extern "C" int picem1d_push(picem1d_ctx* c, const double* cc, const double* e3, const double* b2, const double* u,
                            int64_t n, double* v) {
    PIC_TRACE("picem1d_push");
    if (!c) return PIC_EINVAL;
    if (!cc || !e3 || !b2 || !u || !v || n < 0) {
        set_err(c, "picem1d_push: bad argument");
        return PIC_EINVAL;
    }
    HIPCHK(c, hipSetDevice(c->cfg.device));
    const int N = c->s.N;
    double *dc, *du, *dv, *fe;
    int rc;
    if ((rc = h2d(c, 8, cc, n, &dc)) || (rc = h2d(c, 9, u, 3 * n, &du)) || (rc = h2d(c, 10, nullptr, 3 * n, &dv)) ||
        (rc = h2d(c, 11, nullptr, 5 * (int64_t)N, &fe)))
        return rc;
    HIPCHK(c, hipMemcpyAsync(fe, e3, 24 * (size_t)N, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(fe + 3 * N, b2, 16 * (size_t)N, hipMemcpyHostToDevice, c->stream));
    EM1D s = c->s;
    s.ex = fe;
    s.ey = fe + N;
    s.ez = fe + 2 * N;
    s.by = fe + 3 * N;
    s.bz = fe + 4 * N;
    em_launch_push_op(s, dc, du, n, dv, c->stream);
    if ((rc = d2h(c, v, dv, 3 * n))) return rc;
    return finish(c);
}

// ---- operators of the explicit scheme -------------------------------------------------

extern "C" int picem1d_explicit_fields(picem1d_ctx* c, const double* rho, const double* Jy, const double* Jz,
                                       double* E, double* B, double* b2) {
    PIC_TRACE("picem1d_explicit_fields");
    if (!c) return PIC_EINVAL;
    if (!rho || !Jy || !Jz || !E || !B || !b2) {
        set_err(c, "picem1d_explicit_fields: arguments must not be NULL");
        return PIC_EINVAL;
    }
    HIPCHK(c, hipSetDevice(c->cfg.device));
    const int N = c->s.N;
    double *dr, *djy, *djz, *buf;
    int rc;
    // slot 6: Ex Ey Ez By Bz by bz (7 N)
    if ((rc = h2d(c, 4, rho, N, &dr)) || (rc = h2d(c, 5, Jy, N, &djy)) || (rc = h2d(c, 7, Jz, N, &djz)) ||
        (rc = h2d(c, 6, nullptr, 15 * (int64_t)N, &buf)))
        return rc;
    EM1D s = c->s;
    double** f[7] = {&s.Ex, &s.Ey, &s.Ez, &s.By, &s.Bz, &s.by, &s.bz};
    for (int k = 0; k < 7; ++k) *f[k] = buf + (size_t)k * N;
    HIPCHK(c, hipMemcpyAsync(s.Ex, E, 24 * (size_t)N, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(s.By, B, 16 * (size_t)N, hipMemcpyHostToDevice, c->stream));
    emx_launch_fields_op(s, dr, djy, djz, c->stream);
    if ((rc = d2h(c, E, s.Ex, 3 * (int64_t)N)) || (rc = d2h(c, B, s.By, 2 * (int64_t)N)) ||
        (rc = d2h(c, b2, s.by, 2 * (int64_t)N)))
        return rc;
    return finish(c);
}

extern "C" int picem1d_explicit_push(picem1d_ctx* c, const double* x, const double* E, const double* b2,
                                     const double* u, int64_t n, double* v) {
    PIC_TRACE("picem1d_explicit_push");
    if (!c) return PIC_EINVAL;
    if (!x || !E || !b2 || !u || !v || n < 0) {
        set_err(c, "picem1d_explicit_push: bad argument");
        return PIC_EINVAL;
    }
    HIPCHK(c, hipSetDevice(c->cfg.device));
    const int N = c->s.N;
    double *dx, *du, *dv, *fe;
    int rc;
    if ((rc = h2d(c, 8, x, n, &dx)) || (rc = h2d(c, 9, u, 3 * n, &du)) || (rc = h2d(c, 10, nullptr, 3 * n, &dv)) ||
        (rc = h2d(c, 11, nullptr, 5 * (int64_t)N, &fe)))
        return rc;
    HIPCHK(c, hipMemcpyAsync(fe, E, 24 * (size_t)N, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(fe + 3 * N, b2, 16 * (size_t)N, hipMemcpyHostToDevice, c->stream));
    EM1D s = c->s;
    s.Ex = fe;
    s.Ey = fe + N;
    s.Ez = fe + 2 * N;
    s.by = fe + 3 * N;
    s.bz = fe + 4 * N;
    emx_launch_push_op(s, dx, du, n, dv, c->stream);
    if ((rc = d2h(c, v, dv, 3 * n))) return rc;
    return finish(c);
}
