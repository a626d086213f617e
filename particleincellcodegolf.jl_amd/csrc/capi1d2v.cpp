// capi1d2v.cpp -- host runtime behind include/pic1d2v.h: the magnetised
// 1D2V context (src/NGP1D2V.jl, src/NGP1D2V2S.jl), its device buffers, the
// captured step graphs and the operator calls.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/pic1d2v.h"
#include "pic1d2v_internal.h"
#include "trace.h"

using namespace pic;

namespace {
constexpr int64_t kSteps12 = 8;  // steps in the long step graph
}

struct pic1d2v_ctx {
    pic1d2v_config cfg{};
    hipStream_t stream = nullptr;
    Buffers12 b{};
    std::vector<void*> allocs;
    double* tw = nullptr;
    bool uploaded[2] = {false, false};
    bool primed = false;     // the accumulator holds the charge of the current positions
    int64_t step = 0;        // steps since the upload
    hipGraphExec_t g_one = nullptr, g_many = nullptr;
    void* op[4] = {};
    size_t op_bytes[4] = {};
    std::string err;
};

static thread_local std::string g12_create_err;

static void set_err(pic1d2v_ctx* c, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    else g12_create_err = buf;
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

// Fraction bits of the exact per-cell sums: the total charge magnitude
// (ns*P*w*max|q|) stays below 2^125 on the 2^-q grid.
static int frac_bits(int64_t n, double a) {
    const double total = (double)n * std::fabs(a);
    if (!(total > 0.0) || !std::isfinite(total)) return 64;
    const int q = 124 - std::ilogb(total);
    return q > 1100 ? 1100 : q;
}

template <typename T>
static int dalloc(pic1d2v_ctx* c, T** p, size_t count) {
    void* q = nullptr;
    HIPCHK(c, hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T)));
    c->allocs.push_back(q);
    *p = static_cast<T*>(q);
    return PIC_OK;
}

static int grow_op(pic1d2v_ctx* c, int i, size_t bytes) {
    if (c->op_bytes[i] >= bytes && c->op[i]) return PIC_OK;
    if (c->op[i]) HIPCHK(c, hipFree(c->op[i]));
    c->op[i] = nullptr;
    c->op_bytes[i] = 0;
    HIPCHK(c, hipMalloc(&c->op[i], std::max<size_t>(bytes, 8)));
    c->op_bytes[i] = bytes;
    return PIC_OK;
}

static void drop_graphs(pic1d2v_ctx* c) {
    if (c->g_one) (void)hipGraphExecDestroy(c->g_one);
    if (c->g_many) (void)hipGraphExecDestroy(c->g_many);
    c->g_one = c->g_many = nullptr;
}

static const char* validate(const pic1d2v_config& k) {
    if (k.N < 8 || k.N > kMaxN || (k.N & (k.N - 1)) != 0) return "N must be a power of two in [8, 4096]";
    if (k.shape != PIC_SHAPE_GAUSS6 && k.shape != PIC_SHAPE_GAUSS7)
        return "shape must be PIC_SHAPE_GAUSS6 or PIC_SHAPE_GAUSS7 (the scripts' stencil)";
    if (k.nspecies != 1 && k.nspecies != 2) return "nspecies must be 1 or 2";
    if (k.bin < 1) return "bin must be >= 1";
    if (k.P < 1 || k.P > (int64_t)1 << 30) return "P must be in [1, 2^30]";
    for (int s = 0; s < k.nspecies; ++s)
        if (!std::isfinite(k.q[s]) || !std::isfinite(k.q_m[s]) || !std::isfinite(k.mass[s]))
            return "q, q_m and mass must be finite";
    if (!(k.w >= 0.0) || !std::isfinite(k.w)) return "w must be finite and >= 0";
    if (!(k.n0 > 0.0) || !std::isfinite(k.n0)) return "n0 must be finite and > 0";
    if (!std::isfinite(k.B0) || !std::isfinite(k.dt)) return "B0 and dt must be finite";
    if (k.precision != PIC_PREC_FP64 && k.precision != PIC_PREC_MIXED) return "unknown precision";
    if (k.precision == PIC_PREC_MIXED && !(k.dt > 0.0)) return "mixed precision needs dt > 0";
    return nullptr;
}

extern "C" {

const char* pic1d2v_last_error(const pic1d2v_ctx* c) { return c ? c->err.c_str() : g12_create_err.c_str(); }

int pic1d2v_destroy(pic1d2v_ctx* c) {
    PIC_TRACE("pic1d2v_destroy");
    if (!c) return PIC_OK;
    (void)hipSetDevice(c->cfg.device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    drop_graphs(c);
    for (void* p : c->allocs) (void)hipFree(p);
    if (c->b.Es) (void)hipFree(c->b.Es);
    if (c->b.D) (void)hipFree(c->b.D);
    for (void* p : c->op)
        if (p) (void)hipFree(p);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return PIC_OK;
}

int pic1d2v_create(pic1d2v_ctx** out, const pic1d2v_config* cfg) {
    PIC_TRACE("pic1d2v_create");
    if (!out || !cfg) {
        set_err(nullptr, "pic1d2v_create: null argument");
        return PIC_EINVAL;
    }
    *out = nullptr;
    if (const char* why = validate(*cfg)) {
        set_err(nullptr, "pic1d2v_create: %s", why);
        return PIC_EINVAL;
    }
    pic1d2v_ctx* c = new pic1d2v_ctx();
    c->cfg = *cfg;
    auto fail = [&](int rc) {
        g12_create_err = c->err;
        pic1d2v_destroy(c);
        return rc;
    };
    if (hipSetDevice(cfg->device) != hipSuccess) {
        set_err(c, "pic1d2v_create: cannot use device %d", cfg->device);
        return fail(PIC_EHIP);
    }
    if (configure_kernels_1d2v() != 0) {
        set_err(c, "could not raise the dynamic LDS limit of the 1D2V kernels");
        return fail(PIC_EHIP);
    }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        set_err(c, "pic1d2v_create: hipStreamCreateWithFlags failed");
        return fail(PIC_EHIP);
    }
    hipDeviceProp_t prop;
    const int numCU = hipGetDeviceProperties(&prop, cfg->device) == hipSuccess && prop.multiProcessorCount > 0
                          ? prop.multiProcessorCount
                          : 256;
    Buffers12& b = c->b;
    const int ns = cfg->nspecies;
    b.N = cfg->N;
    b.log2N = ilog2i(cfg->N);
    b.R = cfg->shape;
    b.ns = ns;
    b.bin = cfg->bin;
    b.P = cfg->P;
    double qmax = 0.0;
    for (int s = 0; s < 2; ++s) {
        b.q[s] = s < ns ? cfg->q[s] : 0.0;
        b.q_m[s] = s < ns ? cfg->q_m[s] : 0.0;
        b.mass[s] = s < ns ? cfg->mass[s] : 0.0;
        if (s < ns) qmax = std::max(qmax, std::fabs(cfg->q[s]));
    }
    b.w = cfg->w;
    b.n0 = cfg->n0;
    b.B0 = cfg->B0;
    b.dt = cfg->dt;
    b.qbits = frac_bits((int64_t)ns * cfg->P, cfg->w * qmax);
    const size_t P = (size_t)cfg->P, NP = (size_t)ns * P, N = (size_t)cfg->N;
    int rc;
    b.precision = cfg->precision;
    b.nchunks = ((int64_t)P + kTreeChunk - 1) / kTreeChunk;
    if (b.precision == PIC_PREC_FP64) {
        if ((rc = dalloc(c, &b.x, NP)) || (rc = dalloc(c, &b.vx, NP)) || (rc = dalloc(c, &b.vy, NP))) return fail(rc);
    } else {
        // the mixed state; velocities in steps of the 2^-32 position grid
        b.vstep = cfg->dt * 4294967296.0;
        b.inv_vstep = 1.0 / b.vstep;
        b.escale2 = (cfg->dt * cfg->dt) * 4.0;
        for (int s = 0; s < 2; ++s) {
            b.kick_half[s] = (float)(b.q_m[s] / 2.0);
            b.t3[s] = (float)(cfg->B0 * (cfg->dt / 2.0) * b.q_m[s]);
        }
        if ((rc = dalloc(c, &b.xq, NP)) || (rc = dalloc(c, &b.vxh, NP)) || (rc = dalloc(c, &b.vxl, NP)) ||
            (rc = dalloc(c, &b.vyh, NP)) || (rc = dalloc(c, &b.vyl, NP)) || (rc = dalloc(c, &b.E32, N)) ||
            (rc = dalloc(c, &b.racc32, 2 * N)))
            return fail(rc);
        unsigned char* part = nullptr;
        if ((rc = dalloc(c, &part, 96 * (size_t)b.nchunks))) return fail(rc);  // 6 x 128 bits per chunk
        b.dpart32 = part;
        if (hipMemset(b.racc32, 0, 16 * N) != hipSuccess || hipMemset(b.E32, 0, 4 * N) != hipSuccess) {
            set_err(c, "pic1d2v_create: hipMemset failed");
            return fail(PIC_EHIP);
        }
    }
    if ((rc = dalloc(c, &b.racc, 6 * N))) return fail(rc);
    if ((rc = dalloc(c, &b.E, N))) return fail(rc);
    if ((rc = dalloc(c, &b.dpart, 3 * (size_t)b.nchunks))) return fail(rc);
    if ((rc = dalloc(c, &b.st, 1))) return fail(rc);
    if (hipMemset(b.racc, 0, 48 * N) != hipSuccess || hipMemset(b.E, 0, 8 * N) != hipSuccess ||
        hipMemset(b.st, 0, sizeof(State12)) != hipSuccess) {
        set_err(c, "pic1d2v_create: hipMemset failed");
        return fail(PIC_EHIP);
    }
    const std::vector<double> twp = pass_twiddle_table();
    if ((rc = dalloc(c, &c->tw, twp.size()))) return fail(rc);
    if (hipMemcpy(c->tw, twp.data(), sizeof(double) * twp.size(), hipMemcpyHostToDevice) != hipSuccess) {
        set_err(c, "pic1d2v_create: twiddle upload failed");
        return fail(PIC_EHIP);
    }
    b.tw = reinterpret_cast<const double2*>(c->tw);
    // tiles: runs of one species, about two per compute unit over all particles
    const int64_t target = 2 * (int64_t)numCU;
    int64_t tileP = ((int64_t)NP + target - 1) / target;
    tileP = std::min<int64_t>(std::max<int64_t>(64, (tileP + 63) / 64 * 64), (int64_t)1 << 20);
    int block = 1024;
    while (block > 64 && block / 2 >= tileP) block /= 2;
    std::vector<Tile12> tiles;
    for (int s = 0; s < ns; ++s)
        for (int64_t t0 = 0; t0 < (int64_t)P; t0 += tileP) {
            Tile12 T{};
            T.species = s;
            T.start = (int64_t)s * (int64_t)P + t0;
            T.count = std::min<int64_t>(tileP, (int64_t)P - t0);
            tiles.push_back(T);
        }
    if ((rc = dalloc(c, &b.tiles, tiles.size()))) return fail(rc);
    if (hipMemcpy(b.tiles, tiles.data(), sizeof(Tile12) * tiles.size(), hipMemcpyHostToDevice) != hipSuccess) {
        set_err(c, "pic1d2v_create: tile upload failed");
        return fail(PIC_EHIP);
    }
    b.ntiles = (int32_t)tiles.size();
    b.block = block;
    *out = c;
    return PIC_OK;
}

int pic1d2v_upload(pic1d2v_ctx* c, int32_t species, const double* x, const double* vx, const double* vy) {
    PIC_TRACE("pic1d2v_upload");
    if (!c) return PIC_EINVAL;
    if (species < 0 || species >= c->cfg.nspecies || !x || !vx || !vy) {
        set_err(c, "pic1d2v_upload: bad species or null array");
        return PIC_EINVAL;
    }
    HIPCHK(c, hipSetDevice(c->cfg.device));
    const size_t P = (size_t)c->cfg.P, off = (size_t)species * P;
    Buffers12& b = c->b;
    if (b.precision == PIC_PREC_FP64) {
        HIPCHK(c, hipMemcpyAsync(b.x + off, x, 8 * P, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(b.vx + off, vx, 8 * P, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(b.vy + off, vy, 8 * P, hipMemcpyHostToDevice, c->stream));
    } else {
        // x mod 1 on the 2^-32 grid (2^32 wraps to 0); v * dt * 2^32 as hi + lo
        std::vector<uint32_t> q(P);
        std::vector<float> h(P), l(P);
        for (size_t j = 0; j < P; ++j) {
            const double f = x[j] - std::floor(x[j]);
            q[j] = (uint32_t)(uint64_t)(int64_t)std::nearbyint(f * 4294967296.0);
        }
        HIPCHK(c, hipMemcpy(b.xq + off, q.data(), 4 * P, hipMemcpyHostToDevice));
        const double* vs[2] = {vx, vy};
        float* dh[2] = {b.vxh, b.vyh};
        float* dl[2] = {b.vxl, b.vyl};
        for (int k = 0; k < 2; ++k) {
            for (size_t j = 0; j < P; ++j) {
                const double u = vs[k][j] * b.vstep;
                h[j] = (float)u;
                l[j] = (float)(u - (double)h[j]);
            }
            HIPCHK(c, hipMemcpy(dh[k] + off, h.data(), 4 * P, hipMemcpyHostToDevice));
            HIPCHK(c, hipMemcpy(dl[k] + off, l.data(), 4 * P, hipMemcpyHostToDevice));
        }
        HIPCHK(c, hipMemsetAsync(b.racc32, 0, 16 * (size_t)c->cfg.N, c->stream));
    }
    // the run restarts: step 0, no field bins or diagnostics, charge to be deposited
    HIPCHK(c, hipMemsetAsync(c->b.st, 0, sizeof(State12), c->stream));
    HIPCHK(c, hipMemsetAsync(c->b.racc, 0, 48 * (size_t)c->cfg.N, c->stream));
    if (c->b.Es) HIPCHK(c, hipMemsetAsync(c->b.Es, 0, 8 * (size_t)c->b.cap_bins * c->cfg.N, c->stream));
    if (c->b.D) HIPCHK(c, hipMemsetAsync(c->b.D, 0, 40 * (size_t)c->b.cap_bins, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->uploaded[species] = true;
    c->primed = false;
    c->step = 0;
    return PIC_OK;
}

int pic1d2v_download(pic1d2v_ctx* c, int32_t species, double* x, double* vx, double* vy) {
    PIC_TRACE("pic1d2v_download");
    if (!c) return PIC_EINVAL;
    if (species < 0 || species >= c->cfg.nspecies) {
        set_err(c, "pic1d2v_download: bad species");
        return PIC_EINVAL;
    }
    HIPCHK(c, hipSetDevice(c->cfg.device));
    const size_t P = (size_t)c->cfg.P, off = (size_t)species * P;
    const Buffers12& b = c->b;
    if (b.precision == PIC_PREC_FP64) {
        if (x) HIPCHK(c, hipMemcpyAsync(x, b.x + off, 8 * P, hipMemcpyDeviceToHost, c->stream));
        if (vx) HIPCHK(c, hipMemcpyAsync(vx, b.vx + off, 8 * P, hipMemcpyDeviceToHost, c->stream));
        if (vy) HIPCHK(c, hipMemcpyAsync(vy, b.vy + off, 8 * P, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return PIC_OK;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (x) {
        std::vector<uint32_t> q(P);
        HIPCHK(c, hipMemcpy(q.data(), b.xq + off, 4 * P, hipMemcpyDeviceToHost));
        for (size_t j = 0; j < P; ++j) x[j] = (double)q[j] * 0x1p-32;
    }
    const float* sh[2] = {b.vxh, b.vyh};
    const float* sl[2] = {b.vxl, b.vyl};
    double* dst[2] = {vx, vy};
    std::vector<float> h(P), l(P);
    for (int k = 0; k < 2; ++k) {
        if (!dst[k]) continue;
        HIPCHK(c, hipMemcpy(h.data(), sh[k] + off, 4 * P, hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(l.data(), sl[k] + off, 4 * P, hipMemcpyDeviceToHost));
        for (size_t j = 0; j < P; ++j) dst[k][j] = ((double)h[j] + (double)l[j]) * b.inv_vstep;
    }
    return PIC_OK;
}

// Room for the bins of steps up to `steps`; existing bins are kept.
static int ensure_bins(pic1d2v_ctx* c, int64_t steps) {
    Buffers12& b = c->b;
    const int64_t need = (steps + b.bin - 1) / b.bin;
    if (need <= b.cap_bins && b.Es) return PIC_OK;
    const int64_t cap = std::max<int64_t>(need, std::max<int64_t>(2 * b.cap_bins, 16));
    const size_t N = (size_t)b.N;
    double *Es = nullptr, *D = nullptr;
    HIPCHK(c, hipMalloc(&Es, 8 * N * (size_t)cap));
    HIPCHK(c, hipMalloc(&D, 40 * (size_t)cap));
    HIPCHK(c, hipMemsetAsync(Es, 0, 8 * N * (size_t)cap, c->stream));
    HIPCHK(c, hipMemsetAsync(D, 0, 40 * (size_t)cap, c->stream));
    if (b.Es) {
        HIPCHK(c, hipMemcpyAsync(Es, b.Es, 8 * N * (size_t)b.cap_bins, hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(D, b.D, 40 * (size_t)b.cap_bins, hipMemcpyDeviceToDevice, c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (b.Es) (void)hipFree(b.Es);
    if (b.D) (void)hipFree(b.D);
    b.Es = Es;
    b.D = D;
    b.cap_bins = cap;
    drop_graphs(c);  // the graphs hold the old pointers
    return PIC_OK;
}

static void enqueue_step(pic1d2v_ctx* c) {
    launch12_solve(c->b, c->stream);
    launch12_push(c->b, c->stream);
    launch12_diag(c->b, c->stream);
}

static int capture(pic1d2v_ctx* c, int64_t steps, hipGraphExec_t* out) {
    hipGraph_t g = nullptr;
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    for (int64_t i = 0; i < steps; ++i) enqueue_step(c);
    HIPCHK(c, hipStreamEndCapture(c->stream, &g));
    const hipError_t e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        set_err(c, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
        return PIC_EHIP;
    }
    return PIC_OK;
}

int pic1d2v_run(pic1d2v_ctx* c, int64_t nsteps, double* D) {
    PIC_TRACE("pic1d2v_run");
    if (!c) return PIC_EINVAL;
    if (nsteps < 0) {
        set_err(c, "pic1d2v_run: nsteps < 0");
        return PIC_EINVAL;
    }
    for (int s = 0; s < c->cfg.nspecies; ++s)
        if (!c->uploaded[s]) {
            set_err(c, "pic1d2v_run: species %d has not been uploaded", s);
            return PIC_ESTATE;
        }
    if (nsteps == 0) return PIC_OK;
    HIPCHK(c, hipSetDevice(c->cfg.device));
    int rc = ensure_bins(c, c->step + nsteps);
    if (rc) return rc;
    if (!c->primed) {
        launch12_deposit(c->b, c->stream);
        c->primed = true;
    }
    int64_t left = nsteps;
    if (left >= kSteps12 && !c->g_many && (rc = capture(c, kSteps12, &c->g_many))) return rc;
    if (!c->g_one && (rc = capture(c, 1, &c->g_one))) return rc;
    while (left >= kSteps12) {
        HIPCHK(c, hipGraphLaunch(c->g_many, c->stream));
        left -= kSteps12;
    }
    while (left-- > 0) HIPCHK(c, hipGraphLaunch(c->g_one, c->stream));
    State12 st{};
    HIPCHK(c, hipMemcpyAsync(&st, c->b.st, sizeof st, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipGetLastError());
    const int64_t row0 = c->step / c->b.bin;
    c->step += nsteps;
    if (st.step != c->step) {
        set_err(c, "pic1d2v_run: step bookkeeping lost (%lld of %lld)", (long long)st.step, (long long)c->step);
        return PIC_ESTATE;
    }
    if (st.err & 1) {
        set_err(c, "pic1d2v_run: non-finite charge or field");
        return PIC_ENONFINITE;
    }
    const int64_t row1 = c->step / c->b.bin;
    if (D && row1 > row0)
        HIPCHK(c, hipMemcpy(D, c->b.D + 5 * row0, 40 * (size_t)(row1 - row0), hipMemcpyDeviceToHost));
    return PIC_OK;
}

int pic1d2v_fields(pic1d2v_ctx* c, double* Es, int64_t nbins) {
    PIC_TRACE("pic1d2v_fields");
    if (!c) return PIC_EINVAL;
    const int64_t have = (c->step + c->b.bin - 1) / c->b.bin;
    if (!Es || nbins < 0 || nbins > have) {
        set_err(c, "pic1d2v_fields: null array or nbins > %lld bins run", (long long)have);
        return PIC_EINVAL;
    }
    if (nbins == 0) return PIC_OK;
    HIPCHK(c, hipSetDevice(c->cfg.device));
    HIPCHK(c, hipMemcpy(Es, c->b.Es, 8 * (size_t)c->cfg.N * (size_t)nbins, hipMemcpyDeviceToHost));
    return PIC_OK;
}

int pic1d2v_deposit(pic1d2v_ctx* c, const double* x, double* rho) {
    PIC_TRACE("pic1d2v_deposit");
    if (!c) return PIC_EINVAL;
    if (!x || !rho) {
        set_err(c, "pic1d2v_deposit: null array");
        return PIC_EINVAL;
    }
    HIPCHK(c, hipSetDevice(c->cfg.device));
    const size_t NP = (size_t)c->cfg.nspecies * (size_t)c->cfg.P, N = (size_t)c->cfg.N;
    int rc;
    if ((rc = grow_op(c, 0, 8 * NP)) || (rc = grow_op(c, 1, 48 * N)) || (rc = grow_op(c, 2, 8 * N))) return rc;
    HIPCHK(c, hipMemcpyAsync(c->op[0], x, 8 * NP, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemsetAsync(c->op[1], 0, 48 * N, c->stream));
    Buffers12 o = c->b;
    o.precision = PIC_PREC_FP64;  // the operators are FP64 in every context
    o.x = (double*)c->op[0];
    o.racc = (unsigned long long*)c->op[1];
    launch12_deposit(o, c->stream);
    launch12_rho(o, (double*)c->op[2], c->stream);
    HIPCHK(c, hipMemcpyAsync(rho, c->op[2], 8 * N, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipGetLastError());
    return PIC_OK;
}

int pic1d2v_gather(pic1d2v_ctx* c, const double* E, const double* x, int64_t n, double* Ej) {
    PIC_TRACE("pic1d2v_gather");
    if (!c) return PIC_EINVAL;
    if (!E || !x || !Ej || n < 1) {
        set_err(c, "pic1d2v_gather: null array or n < 1");
        return PIC_EINVAL;
    }
    HIPCHK(c, hipSetDevice(c->cfg.device));
    const size_t N = (size_t)c->cfg.N;
    int rc;
    if ((rc = grow_op(c, 0, 8 * (size_t)n)) || (rc = grow_op(c, 2, 8 * N)) || (rc = grow_op(c, 3, 8 * (size_t)n)))
        return rc;
    HIPCHK(c, hipMemcpyAsync(c->op[0], x, 8 * (size_t)n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->op[2], E, 8 * N, hipMemcpyHostToDevice, c->stream));
    launch12_op_gather(c->b, (const double*)c->op[2], (const double*)c->op[0], n, (double*)c->op[3], c->stream);
    HIPCHK(c, hipMemcpyAsync(Ej, c->op[3], 8 * (size_t)n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipGetLastError());
    return PIC_OK;
}

int pic1d2v_boris(pic1d2v_ctx* c, double* vx, double* vy, const double* Ej, int64_t n, double B, double dt,
                  double q_m) {
    PIC_TRACE("pic1d2v_boris");
    if (!c) return PIC_EINVAL;
    if (!vx || !vy || !Ej || n < 1) {
        set_err(c, "pic1d2v_boris: null array or n < 1");
        return PIC_EINVAL;
    }
    HIPCHK(c, hipSetDevice(c->cfg.device));
    const size_t bytes = 8 * (size_t)n;
    int rc;
    if ((rc = grow_op(c, 0, bytes)) || (rc = grow_op(c, 2, bytes)) || (rc = grow_op(c, 3, bytes))) return rc;
    HIPCHK(c, hipMemcpyAsync(c->op[0], vx, bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->op[2], vy, bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->op[3], Ej, bytes, hipMemcpyHostToDevice, c->stream));
    launch12_op_boris((double*)c->op[0], (double*)c->op[2], (const double*)c->op[3], n, B, dt, q_m, c->stream);
    HIPCHK(c, hipMemcpyAsync(vx, c->op[0], bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(vy, c->op[2], bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipGetLastError());
    return PIC_OK;
}

}  // extern "C"
