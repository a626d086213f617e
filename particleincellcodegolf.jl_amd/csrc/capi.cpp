// capi.cpp -- host runtime behind include/pic.h: contexts, device buffers,
// instance/tile tables, hipGraph capture of the step, and the operator calls.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <thread>
#include <vector>

#include "../../include/pic.h"
#include "pic_internal.h"
#include "trace.h"

using namespace pic;

namespace {

constexpr int64_t kGraphSteps = 8;  // steps in the long step graph

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

}  // namespace

struct pic_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<pic_config> cfg;
    int ninst = 0;
    int R = 7, scheme = 1, precision = 0, max_sweeps = 10;
    int maxN = 0;
    int64_t Ptot = 0;
    int numCU = 256;
    std::vector<InstDesc> hinst;
    std::vector<TileDesc> htiles;
    std::vector<ChunkDesc> hchunks;
    RunBuffers rb{};
    std::vector<void*> allocs;
    // growable scratch for the operators
    DevBuf op_a, op_b, op_c, op_d, op_e, op_f, op_part, op_dpart, op_flag;
    double* tw = nullptr;
    // captured step graphs by sweep slots per step: one step, and kGraphSteps steps
    struct StepGraphs {
        hipGraphExec_t one = nullptr, many = nullptr;
    };
    std::map<int, StepGraphs> graphs;
    bool primed = false;       // the accumulator holds the next step's first charge
    int sweeps_seen = 0;       // most sweeps any step has needed (0: none run yet)
    int64_t steps_done = 0;
    // persistent runs of a small single instance (pic_set_run_path)
    int run_path = PIC_RUN_AUTO;
    bool small_ok = false;         // the instance fits the persistent kernels
    int64_t small_K = 0;           // their particle workgroups (Gaussian)
    unsigned long long* small_gacc = nullptr;
    uint32_t* small_bar = nullptr;
    // particle tiling of the push launches (pic_set_tiling / pic_autotune)
    int64_t tileP = 0;
    DevBuf tiles_buf, dpart_buf, snap_buf;
    std::string err;
    // pic_run_async: the run's host thread, and whether it is still in flight
    std::thread async_thread;
    std::atomic<bool> busy{false};
    int async_rc = 0;
};

static thread_local std::string g_create_err;

// Every call but pic_sync, pic_destroy and the getters of fixed settings is
// refused while a pic_run_async run is in flight on the context (the run's
// thread owns c->err until pic_sync, so nothing is written there).
static int in_flight(pic_ctx*) { return PIC_ESTATE; }

static void set_err(pic_ctx* c, const char* fmt, ...) {
    if (!c) return;
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    c->err = buf;
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

// Fraction bits of the exact per-cell accumulator: P*w*2^q < 2^125.
static int fxp_frac_bits(int64_t P, double w) {
    double total = (double)P * std::fabs(w);
    if (!(total > 0.0) || !std::isfinite(total)) return 64;
    int e = std::ilogb(total);
    int q = 124 - e;
    return q > 1100 ? 1100 : q;
}

// W[m] = exp(-2 pi i m / 4096), m < 2048, from cos/sin on [0, pi/4] and exact
// quarter-wave symmetries; (re, im) pairs.
static void twiddles4096(std::vector<double>& out) {
    const int N = kMaxN, h = N / 2, q = N / 4, e = N / 8;
    const double twopi = 6.283185307179586;
    out.assign(2 * h, 0.0);
    for (int m = 0; m < h; ++m) {
        double c, s;
        if (m <= e) {
            double th = twopi * (double)m / (double)N;
            c = std::cos(th);
            s = std::sin(th);
        } else if (m <= q) {
            double th = twopi * (double)(q - m) / (double)N;
            c = std::sin(th);
            s = std::cos(th);
        } else {
            int mm = h - m;
            if (mm <= e) {
                double th = twopi * (double)mm / (double)N;
                c = -std::cos(th);
                s = std::sin(th);
            } else {
                double th = twopi * (double)(q - mm) / (double)N;
                c = -std::sin(th);
                s = std::cos(th);
            }
        }
        out[2 * m] = c;
        out[2 * m + 1] = -s;
    }
}

// The per-stage table of the radix-2 transforms plus the division factors of
// the 1D solve.  Stage s (1..12) uses W_{2^s}^k = W[k << (12 - s)] for
// k < 2^(s-1), stored at entry 2^(s-1) - 1 + k (a pass of stages s0+1..s0+M
// reads lanes lo of entries 2^(s0+t) - 1 + jlo*2^s0 + lo, contiguous across
// the lanes of a wave).  From double kXiOffset on: 1/(2 pi k) for
// k = 1..kMaxN/2, the factor of mode k' = +-k in xi = rho^ / (2 pi i k').
std::vector<double> pic::pass_twiddle_table() {
    std::vector<double> tw;
    twiddles4096(tw);
    const int L = 12;  // log2 of the 4096-point table's transform
    std::vector<double> out(kXiOffset + kMaxN / 2 + 1, 0.0);
    for (int s = 1; s <= L; ++s)
        for (int k = 0; k < (1 << (s - 1)); ++k) {
            const int e = (1 << (s - 1)) - 1 + k;
            out[2 * e] = tw[2 * (k << (L - s))];
            out[2 * e + 1] = tw[2 * (k << (L - s)) + 1];
        }
    for (int k = 1; k <= kMaxN / 2; ++k) out[kXiOffset + k] = 1.0 / (kTwoPi * (double)k);
    return out;
}

template <typename T>
static int dalloc(pic_ctx* c, T** p, size_t count) {
    size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    void* q = nullptr;
    HIPCHK(c, hipMalloc(&q, bytes));
    c->allocs.push_back(q);
    *p = static_cast<T*>(q);
    return PIC_OK;
}

static int grow(pic_ctx* c, DevBuf& b, size_t bytes) {
    if (b.bytes >= bytes && b.p) return PIC_OK;
    if (b.p) {
        HIPCHK(c, hipFree(b.p));
        b.p = nullptr;
        b.bytes = 0;
    }
    HIPCHK(c, hipMalloc(&b.p, std::max<size_t>(bytes, 64)));
    b.bytes = std::max<size_t>(bytes, 64);
    return PIC_OK;
}

static void drop_graphs(pic_ctx* c) {
    for (auto& e : c->graphs) {
        if (e.second.many) (void)hipGraphExecDestroy(e.second.many);
        if (e.second.one) (void)hipGraphExecDestroy(e.second.one);
    }
    c->graphs.clear();
}

static const char* validate(const pic_config& k) {
    if (k.N < 8 || k.N > kMaxN || (k.N & (k.N - 1)) != 0) return "N must be a power of two in [8, 4096]";
    if (k.P < 1 || k.P > (int64_t)1 << 31) return "P must be in [1, 2^31]";
    if (k.shape != PIC_SHAPE_NGP && k.shape != PIC_SHAPE_AREA && k.shape != PIC_SHAPE_GAUSS6 &&
        k.shape != PIC_SHAPE_GAUSS7)
        return "shape must be PIC_SHAPE_NGP, PIC_SHAPE_AREA, PIC_SHAPE_GAUSS6 or PIC_SHAPE_GAUSS7";
    if (k.scheme != PIC_SCHEME_LEAPFROG && k.scheme != PIC_SCHEME_MIDPOINT && k.scheme != PIC_SCHEME_SIMPSON)
        return "unknown scheme";
    if (k.scheme == PIC_SCHEME_MIDPOINT && k.shape != PIC_SHAPE_GAUSS6 && k.shape != PIC_SHAPE_GAUSS7)
        return "the midpoint scheme needs a Gaussian shape (as in the scripts)";
    if (k.scheme == PIC_SCHEME_LEAPFROG && k.shape == PIC_SHAPE_AREA)
        return "the area shape is provided for the Simpson scheme";
    if (k.scheme == PIC_SCHEME_SIMPSON && k.shape == PIC_SHAPE_NGP)
        return "the Simpson scheme needs the area or a Gaussian shape (as in the scripts)";
    if (k.scheme == PIC_SCHEME_SIMPSON && k.N > 1024)
        return "the Simpson scheme is provided for N <= 1024 (its 3 x N convergence test is one reduction chunk)";
    if (k.precision != PIC_PREC_FP64 && k.precision != PIC_PREC_MIXED) return "unknown precision";
    if (k.precision == PIC_PREC_MIXED && k.scheme == PIC_SCHEME_LEAPFROG)
        return "mixed precision is provided for the midpoint and Simpson schemes";
    if (!(k.w >= 0.0) || !std::isfinite(k.w)) return "w must be finite and >= 0";
    if (!std::isfinite(k.dt) || !std::isfinite(k.W) || !(k.W > 0.0)) return "dt must be finite, W finite and > 0";
    if (k.scheme != PIC_SCHEME_LEAPFROG && (k.max_sweeps < 1 || k.max_sweeps > 64))
        return "max_sweeps must be in [1, 64]";
    if (!(k.rtol >= 0.0)) return "rtol must be >= 0";
    if (k.precision == PIC_PREC_MIXED && !mixed_available()) return "mixed precision kernels are not built";
    return nullptr;
}

extern "C" {

int pic_version(void) { return 1; }

const char* pic_strerror(int code) {
    switch (code) {
        case PIC_OK: return "ok";
        case PIC_EINVAL: return "invalid argument";
        case PIC_EHIP: return "HIP runtime error";
        case PIC_ENOMEM: return "out of device memory";
        case PIC_ENONFINITE: return "non-finite charge or field";
        case PIC_ESTATE: return "call out of order";
        default: return "unknown error";
    }
}

// pic_last_error(NULL) reports why the last pic_create / pic_create_batch on
// this thread failed.
// While a pic_run_async run is in flight its thread owns ctx->err, and the
// refusal is reported instead.
const char* pic_last_error(const pic_ctx* ctx) {
    if (ctx && ctx->busy.load()) return "a pic_run_async run is in flight; call pic_sync first";
    return ctx ? ctx->err.c_str() : g_create_err.c_str();
}

// Default tiling: every tile of every instance holds at most tileP particles,
// so the launch is balanced, with enough tiles to cover the chip.  FP64: as
// many tiles as fit at once (LDS-bound); mixed: two 1024-thread blocks per CU,
// so the per-tile partials (N u64 each) stay small.
static void default_tiling(const pic_ctx* c, int64_t* tileP, int* block) {
    const int lds_push = c->scheme == PIC_SCHEME_SIMPSON
                             ? (c->precision == 0 ? simpson_push_lds_bytes(c->maxN) : simpson32_push_lds_bytes(c->maxN))
                             : push_lds_bytes(c->maxN, c->precision);
    const int wg_per_cu = std::max(1, std::min(4, (160 * 1024) / std::max(lds_push, 1)));
    const int64_t target = (int64_t)c->numCU * (c->precision == 0 ? wg_per_cu : 2);
    *tileP = (c->Ptot + target - 1) / target;
    *block = (c->precision == 0 && wg_per_cu >= 2) ? 512 : 1024;
}

// Lays the particles of every instance out in tiles of at most tileP (rounded
// up to a multiple of 64, within [64, 2^20]) run by blocks of `block` threads
// (halved while a block would be twice the tile).  Only the launch layout
// changes: the run's results do not depend on it.
static int layout_tiles(pic_ctx* c, int64_t tileP, int block) {
    tileP = std::max<int64_t>(64, ((tileP + 63) / 64) * 64);
    tileP = std::min<int64_t>(tileP, (int64_t)1 << 20);
    while (block > 64 && block / 2 >= tileP) block /= 2;
    drop_graphs(c);
    c->tileP = tileP;
    c->htiles.clear();
    for (int i = 0; i < c->ninst; ++i) {
        InstDesc& d = c->hinst[i];
        d.tile0 = (int32_t)c->htiles.size();
        const int64_t nt = (d.P + tileP - 1) / tileP;
        d.ntiles = (int32_t)nt;
        for (int64_t t = 0; t < nt; ++t) {
            TileDesc T{};
            T.inst = i;
            T.local = (int32_t)t;
            T.start = t * tileP;
            T.count = std::min<int64_t>(tileP, d.P - t * tileP);
            c->htiles.push_back(T);
        }
    }
    RunBuffers& b = c->rb;
    b.block_push = block;
    b.ntiles = (int32_t)c->htiles.size();
    int rc;
    if ((rc = grow(c, c->tiles_buf, sizeof(TileDesc) * (size_t)b.ntiles))) return rc;
    // FP64: a double pair per diagnostics chunk; mixed midpoint: a 128-bit pair per tile
    if ((rc = grow(c, c->dpart_buf, sizeof(double) * 4 * (size_t)std::max(b.nchunks, b.ntiles)))) return rc;
    b.tiles = (TileDesc*)c->tiles_buf.p;
    b.dpart = (double*)c->dpart_buf.p;
    HIPCHK(c, hipMemcpy(b.insts, c->hinst.data(), sizeof(InstDesc) * c->ninst, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(b.tiles, c->htiles.data(), sizeof(TileDesc) * b.ntiles, hipMemcpyHostToDevice));
    return PIC_OK;
}

static int build_tables(pic_ctx* c) {
    const int n = c->ninst;
    c->hinst.assign(n, InstDesc{});
    c->htiles.clear();
    c->hchunks.clear();
    c->rb.lds_push = push_lds_bytes(c->maxN, c->precision);
    int64_t poff = 0, red0 = 0;
    int32_t goff = 0;
    for (int i = 0; i < n; ++i) {
        const pic_config& k = c->cfg[i];
        InstDesc& d = c->hinst[i];
        d.N = k.N;
        d.log2N = ilog2i(k.N);
        d.nmask = k.N - 1;
        d.R = k.shape;
        d.P = k.P;
        d.poff = poff;
        d.goff = goff;
        d.q = fxp_frac_bits(k.P, k.w);
        d.max_sweeps = k.max_sweeps;
        d.w = k.w;
        d.dt = k.dt;
        d.W = k.W;
        d.rtol = k.rtol;
        d.wq = k.w * 0x1p-30;
        d.vstep = k.dt * 4294967296.0;
        d.inv_vstep = 1.0 / d.vstep;
        d.escale2 = (k.dt * k.dt) * 4.0;
        d.red0 = red0;
        red0 += k.N;
        d.chunk0 = (int32_t)c->hchunks.size();
        const int64_t nch = (k.P + kTreeChunk - 1) / kTreeChunk;
        d.nchunks = (int32_t)nch;
        for (int64_t ch = 0; ch < nch; ++ch) {
            ChunkDesc C{};
            C.inst = i;
            C.start = ch * kTreeChunk;
            C.count = std::min<int64_t>(kTreeChunk, k.P - ch * kTreeChunk);
            c->hchunks.push_back(C);
        }
        poff += k.P;
        goff += k.N;
    }
    RunBuffers& b = c->rb;
    b.ninst = n;
    b.nchunks = (int32_t)c->hchunks.size();
    b.maxN = c->maxN;
    b.R = c->R;
    b.scheme = c->scheme;
    b.precision = c->precision;
    b.max_sweeps = c->max_sweeps;
    b.block_solve = 256;
    int rc;
    if ((rc = dalloc(c, &b.insts, n))) return rc;
    if ((rc = dalloc(c, &b.st, n))) return rc;
    if ((rc = dalloc(c, &b.chunks, b.nchunks))) return rc;
    int64_t tileP;
    int block;
    default_tiling(c, &tileP, &block);
    if ((rc = layout_tiles(c, tileP, block))) return rc;
    HIPCHK(c, hipMemcpy(b.chunks, c->hchunks.data(), sizeof(ChunkDesc) * b.nchunks, hipMemcpyHostToDevice));
    std::vector<InstState> st(n);
    for (auto& s : st) {
        std::memset(&s, 0, sizeof s);
        s.last_sweep = -1;
    }
    HIPCHK(c, hipMemcpy(b.st, st.data(), sizeof(InstState) * n, hipMemcpyHostToDevice));
    const size_t P = (size_t)c->Ptot, G = (size_t)goff;
    if (c->precision == PIC_PREC_FP64) {
        for (int j = 0; j < 2; ++j) {
            if ((rc = dalloc(c, &b.X[j], P))) return rc;
            if ((rc = dalloc(c, &b.V[j], P))) return rc;
            if (c->scheme != PIC_SCHEME_LEAPFROG && (rc = dalloc(c, &b.vit[j], P))) return rc;
        }
        if ((rc = dalloc(c, reinterpret_cast<unsigned long long**>(&b.racc), 3 * (size_t)red0))) return rc;
        HIPCHK(c, hipMemset(b.racc, 0, 3 * 8 * (size_t)red0));
        b.racc_cells = 3 * red0;
        b.prime_next = c->scheme == PIC_SCHEME_MIDPOINT ? 1 : 0;
    } else {
        for (int j = 0; j < 2; ++j) {
            if ((rc = dalloc(c, &b.Xq[j], P))) return rc;
            if ((rc = dalloc(c, &b.Vh[j], P))) return rc;
            if ((rc = dalloc(c, &b.Vl[j], P))) return rc;
            if ((rc = dalloc(c, &b.vf[j], P))) return rc;
        }
        if ((rc = dalloc(c, reinterpret_cast<unsigned long long**>(&b.racc), (size_t)red0))) return rc;
        HIPCHK(c, hipMemset(b.racc, 0, 8 * (size_t)red0));
        b.racc_cells = red0;
        b.prime_next = c->scheme == PIC_SCHEME_MIDPOINT ? 1 : 0;
    }
    if (c->scheme == PIC_SCHEME_SIMPSON) {
        // FP64: three limbs per cell; mixed: one u64 of quanta
        const size_t words = (c->precision == PIC_PREC_FP64 ? 3 : 1) * (size_t)red0;
        if ((rc = dalloc(c, &b.Es1, G))) return rc;
        if ((rc = dalloc(c, &b.Es2[0], G))) return rc;
        if ((rc = dalloc(c, &b.Es2[1], G))) return rc;
        if (c->precision == PIC_PREC_FP64) {
            if ((rc = dalloc(c, &b.Vg, P))) return rc;
        } else {
            if ((rc = dalloc(c, &b.G1s, P))) return rc;
            if ((rc = dalloc(c, &b.E32s1, G))) return rc;
            if ((rc = dalloc(c, &b.E32s2, G))) return rc;
            HIPCHK(c, hipMemset(b.E32s1, 0, sizeof(float) * G));
            HIPCHK(c, hipMemset(b.E32s2, 0, sizeof(float) * G));
        }
        if ((rc = dalloc(c, reinterpret_cast<unsigned long long**>(&b.racc2), words))) return rc;
        if ((rc = dalloc(c, reinterpret_cast<unsigned long long**>(&b.racc3), words))) return rc;
        HIPCHK(c, hipMemset(b.racc2, 0, 8 * words));
        HIPCHK(c, hipMemset(b.racc3, 0, 8 * words));
        HIPCHK(c, hipMemset(b.Es1, 0, sizeof(double) * G));
        HIPCHK(c, hipMemset(b.Es2[0], 0, sizeof(double) * G));
        HIPCHK(c, hipMemset(b.Es2[1], 0, sizeof(double) * G));
    }
    if ((rc = dalloc(c, &b.rho, G))) return rc;
    if ((rc = dalloc(c, &b.E[0], G))) return rc;
    if ((rc = dalloc(c, &b.E[1], G))) return rc;
    if ((rc = dalloc(c, &b.E32, G))) return rc;
    if ((rc = dalloc(c, &b.halt, 1))) return rc;
    HIPCHK(c, hipMemset(b.halt, 0, sizeof(int32_t)));
    HIPCHK(c, hipMemset(b.E[0], 0, sizeof(double) * G));
    HIPCHK(c, hipMemset(b.E[1], 0, sizeof(double) * G));
    HIPCHK(c, hipMemset(b.E32, 0, sizeof(float) * G));
    b.tw = c->tw;
    b.diag = nullptr;
    b.sweeps = nullptr;
    b.diag_cap = 0;
    return PIC_OK;
}

int pic_create_batch(pic_ctx** out, const pic_config* cfgs, int32_t n) {
    PIC_TRACE("pic_create_batch");
    if (!out || !cfgs || n < 1) return PIC_EINVAL;
    *out = nullptr;
    pic_ctx* c = new pic_ctx();
    c->cfg.assign(cfgs, cfgs + n);
    c->ninst = n;
    for (int i = 0; i < n; ++i) {
        const char* why = validate(cfgs[i]);
        if (!why && (cfgs[i].shape != cfgs[0].shape || cfgs[i].scheme != cfgs[0].scheme ||
                     cfgs[i].precision != cfgs[0].precision || cfgs[i].max_sweeps != cfgs[0].max_sweeps ||
                     cfgs[i].device != cfgs[0].device))
            why = "all instances of a batch share shape, scheme, precision, max_sweeps and device";
        if (why) {
            g_create_err = std::string("instance ") + std::to_string(i) + ": " + why;
            delete c;
            return PIC_EINVAL;
        }
        c->maxN = std::max(c->maxN, cfgs[i].N);
        c->Ptot += cfgs[i].P;
    }
    c->R = cfgs[0].shape;
    c->scheme = cfgs[0].scheme;
    c->precision = cfgs[0].precision;
    c->max_sweeps = cfgs[0].scheme != PIC_SCHEME_LEAPFROG ? cfgs[0].max_sweeps : 1;
    c->device = cfgs[0].device;
    int rc = PIC_OK;
    auto fail = [&](int code) {
        g_create_err = c->err.empty() ? pic_strerror(code) : c->err;
        pic_destroy(c);
        return code;
    };
    if (hipSetDevice(c->device) != hipSuccess) return fail(PIC_EHIP);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, c->device) != hipSuccess) return fail(PIC_EHIP);
    c->numCU = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    if (configure_kernels() != 0) {
        set_err(c, "could not raise the dynamic LDS limit of the kernels");
        return fail(PIC_EHIP);
    }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return fail(PIC_EHIP);
    const std::vector<double> twp = pass_twiddle_table();
    if ((rc = dalloc(c, &c->tw, twp.size()))) return fail(rc);
    if (hipMemcpy(c->tw, twp.data(), sizeof(double) * twp.size(), hipMemcpyHostToDevice) != hipSuccess)
        return fail(PIC_EHIP);
    if ((rc = build_tables(c))) return fail(rc);
    if (c->ninst == 1 && small_eligible(c->hinst[0], c->scheme, c->precision, &c->small_K)) {
        c->small_ok = true;
        if ((rc = dalloc(c, &c->small_gacc, 9 * (size_t)c->hinst[0].N))) return fail(rc);
        if ((rc = dalloc(c, &c->small_bar, 4))) return fail(rc);
    }
    *out = c;
    return PIC_OK;
}

int pic_create(pic_ctx** out, const pic_config* cfg) {
    PIC_TRACE("pic_create");
    return pic_create_batch(out, cfg, 1);
}

int pic_destroy(pic_ctx* c) {
    PIC_TRACE("pic_destroy");
    if (!c) return PIC_OK;
    if (c->async_thread.joinable()) c->async_thread.join();
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    drop_graphs(c);
    for (void* p : c->allocs) (void)hipFree(p);
    for (DevBuf* b : {&c->op_a, &c->op_b, &c->op_c, &c->op_d, &c->op_e, &c->op_f, &c->op_part,
                      &c->op_dpart, &c->op_flag, &c->tiles_buf, &c->dpart_buf, &c->snap_buf})
        if (b->p) (void)hipFree(b->p);
    if (c->rb.diag) (void)hipFree(c->rb.diag);
    if (c->rb.sweeps) (void)hipFree(c->rb.sweeps);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return PIC_OK;
}

static int read_states(pic_ctx* c, std::vector<InstState>& st) {
    st.resize(c->ninst);
    HIPCHK(c, hipMemcpyAsync(st.data(), c->rb.st, sizeof(InstState) * c->ninst, hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PIC_OK;
}

static int write_states(pic_ctx* c, const std::vector<InstState>& st) {
    HIPCHK(c, hipMemcpyAsync(c->rb.st, st.data(), sizeof(InstState) * c->ninst, hipMemcpyHostToDevice,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PIC_OK;
}

int pic_upload(pic_ctx* c, int32_t inst, const double* x, const double* v, const double* E) {
    PIC_TRACE("pic_upload");
    if (!c) return PIC_EINVAL;
    if (c->busy.load()) return in_flight(c);
    if (inst < 0 || inst >= c->ninst || !x || !v) {
        set_err(c, "pic_upload: bad instance or null array");
        return PIC_EINVAL;
    }
    HIPCHK(c, hipSetDevice(c->device));
    const InstDesc& d = c->hinst[inst];
    const int par = (int)(c->steps_done & 1);
    const size_t P = (size_t)d.P;
    if (c->precision == PIC_PREC_FP64) {
        HIPCHK(c, hipMemcpyAsync(c->rb.X[par] + d.poff, x, P * 8, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->rb.V[par] + d.poff, v, P * 8, hipMemcpyHostToDevice, c->stream));
    } else {
        std::vector<uint32_t> xq(P);
        std::vector<float> vh(P), vl(P);
        for (size_t j = 0; j < P; ++j) {
            double f = x[j] - std::floor(x[j]);
            double s = std::nearbyint(f * 4294967296.0);
            xq[j] = (uint32_t)(uint64_t)(int64_t)s;  // 2^32 wraps to 0
            const double vs = v[j] * d.vstep;  // steps of the fixed-point position
            vh[j] = (float)vs;
            vl[j] = (float)(vs - (double)vh[j]);
        }
        HIPCHK(c, hipMemcpyAsync(c->rb.Xq[par] + d.poff, xq.data(), P * 4, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->rb.Vh[par] + d.poff, vh.data(), P * 4, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->rb.Vl[par] + d.poff, vl.data(), P * 4, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    std::vector<double> e((size_t)d.N, 0.0), zero((size_t)d.N, 0.0);
    if (E) std::memcpy(e.data(), E, sizeof(double) * d.N);
    if (c->scheme == PIC_SCHEME_SIMPSON) {
        // E is the script's E[end,:]; E[1,:] and E[2,:] start at zero
        for (double* g : {c->rb.Es1, c->rb.Es2[0], c->rb.Es2[1]})
            HIPCHK(c, hipMemcpyAsync(g + d.goff, zero.data(), sizeof(double) * d.N, hipMemcpyHostToDevice,
                                     c->stream));
    }
    std::vector<float> e32((size_t)d.N);
    for (int i = 0; i < d.N; ++i) e32[i] = (float)(e[i] * d.escale2);  // the mixed gather's units
    HIPCHK(c, hipMemcpyAsync(c->rb.E[0] + d.goff, e.data(), sizeof(double) * d.N, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->rb.E32 + d.goff, e32.data(), sizeof(float) * d.N, hipMemcpyHostToDevice, c->stream));
    std::vector<InstState> st;
    int rc = read_states(c, st);
    if (rc) return rc;
    st[inst].e_cur = 0;
    st[inst].err = 0;
    c->primed = false;  // the fused first deposit no longer matches the state
    return write_states(c, st);
}

int pic_download(pic_ctx* c, int32_t inst, double* x, double* v, double* E) {
    PIC_TRACE("pic_download");
    if (!c) return PIC_EINVAL;
    if (c->busy.load()) return in_flight(c);
    if (inst < 0 || inst >= c->ninst) {
        set_err(c, "pic_download: bad instance");
        return PIC_EINVAL;
    }
    HIPCHK(c, hipSetDevice(c->device));
    const InstDesc& d = c->hinst[inst];
    const int par = (int)(c->steps_done & 1);
    const size_t P = (size_t)d.P;
    if (c->precision == PIC_PREC_FP64) {
        if (x) HIPCHK(c, hipMemcpyAsync(x, c->rb.X[par] + d.poff, P * 8, hipMemcpyDeviceToHost, c->stream));
        if (v) HIPCHK(c, hipMemcpyAsync(v, c->rb.V[par] + d.poff, P * 8, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    } else if (x || v) {
        std::vector<uint32_t> xq(P);
        std::vector<float> vh(P), vl(P);
        HIPCHK(c, hipMemcpyAsync(xq.data(), c->rb.Xq[par] + d.poff, P * 4, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipMemcpyAsync(vh.data(), c->rb.Vh[par] + d.poff, P * 4, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipMemcpyAsync(vl.data(), c->rb.Vl[par] + d.poff, P * 4, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        for (size_t j = 0; j < P; ++j) {
            if (x) x[j] = (double)xq[j] * 0x1p-32;
            if (v) v[j] = ((double)vh[j] + (double)vl[j]) * d.inv_vstep;
        }
    }
    if (E) {
        std::vector<InstState> st;
        int rc = read_states(c, st);
        if (rc) return rc;
        const double* src = c->rb.E[st[inst].e_cur] + d.goff;
        HIPCHK(c, hipMemcpy(E, src, sizeof(double) * d.N, hipMemcpyDeviceToHost));
    }
    return PIC_OK;
}

// ---- run ------------------------------------------------------------------

// One step with `slots` sweep slots (midpoint; slots <= max_sweeps).  With
// fewer slots than max_sweeps the last slot's solve halts the graph when an
// instance needs more sweeps (RunBuffers::halt).
static void enqueue_step(pic_ctx* c, hipStream_t s, int slots) {
    const RunBuffers& b = c->rb;
    if (c->scheme == PIC_SCHEME_MIDPOINT) {
        if (!b.prime_next) launch_mid_push0(b, s);
        for (int k = 0; k < slots; ++k) {
            const bool last = k == slots - 1;
            launch_solve(b, k, last && slots < c->max_sweeps, s);
            launch_mid_push(b, k, last, s);
        }
    } else if (c->scheme == PIC_SCHEME_SIMPSON) {
        launch_simp_first(b, s);
        for (int k = 0; k < slots; ++k) {
            const bool last = k == slots - 1;
            launch_simp_solve(b, k, last && slots < c->max_sweeps, s);
            launch_simp_push(b, k, last, s);
        }
    } else {
        launch_lf_first(b, s);
        launch_solve(b, 0, false, s);
        launch_lf_second(b, s);
    }
    launch_diag(b, s);
}

// The rest of a step that halted after sweep slot k0 - 1.
static void enqueue_tail(pic_ctx* c, hipStream_t s, int k0) {
    const RunBuffers& b = c->rb;
    for (int k = k0; k < c->max_sweeps; ++k) {
        if (c->scheme == PIC_SCHEME_SIMPSON) {
            launch_simp_solve(b, k, false, s);
            launch_simp_push(b, k, false, s);
        } else {
            launch_solve(b, k, false, s);
            launch_mid_push(b, k, false, s);
        }
    }
    launch_diag(b, s);
}

static int capture(pic_ctx* c, int64_t steps, int slots, hipGraphExec_t* out) {
    hipGraph_t g = nullptr;
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    for (int64_t i = 0; i < steps; ++i) enqueue_step(c, c->stream, slots);
    HIPCHK(c, hipStreamEndCapture(c->stream, &g));
    hipError_t e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        set_err(c, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
        return PIC_EHIP;
    }
    return PIC_OK;
}

static int ensure_diag(pic_ctx* c, int64_t nsteps) {
    RunBuffers& b = c->rb;
    if (b.diag_cap >= nsteps && b.diag) return PIC_OK;
    drop_graphs(c);
    if (b.diag) (void)hipFree(b.diag);
    if (b.sweeps) (void)hipFree(b.sweeps);
    b.diag = nullptr;
    b.sweeps = nullptr;
    const int64_t cap = std::max<int64_t>(nsteps, 64);
    HIPCHK(c, hipMalloc(&b.diag, sizeof(double) * 4 * cap * c->ninst));
    HIPCHK(c, hipMalloc(&b.sweeps, sizeof(int32_t) * cap * c->ninst));
    b.diag_cap = cap;
    return PIC_OK;
}

// Copies the leading dst_row bytes of every instance's row (src_row bytes
// apart on the device) into consecutive host rows: one transfer of the whole
// buffer (strided device-to-host copies are slow), then a host-side gather.
static int copy_rows(pic_ctx* c, void* dst, const void* src, size_t dst_row, size_t src_row) {
    const size_t ni = (size_t)c->ninst;
    if (dst_row == src_row) {
        HIPCHK(c, hipMemcpyAsync(dst, src, ni * src_row, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return PIC_OK;
    }
    std::vector<char> h(ni * src_row);
    HIPCHK(c, hipMemcpyAsync(h.data(), src, h.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < ni; ++i) std::memcpy((char*)dst + i * dst_row, h.data() + i * src_row, dst_row);
    return PIC_OK;
}

static int begin_run(pic_ctx* c, int64_t nsteps) {
    int rc = ensure_diag(c, nsteps);
    if (rc) return rc;
    std::vector<InstState> st;
    if ((rc = read_states(c, st))) return rc;
    for (auto& s : st) s.run_base = s.step;
    return write_states(c, st);
}

static int finish_run(pic_ctx* c, int64_t nsteps, double* diag, int32_t* sweeps) {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipGetLastError());
    c->steps_done += nsteps;
    std::vector<InstState> st;
    int rc = read_states(c, st);
    if (rc) return rc;
    const size_t cap = (size_t)c->rb.diag_cap, n = (size_t)nsteps;
    if (diag && (rc = copy_rows(c, diag, c->rb.diag, n * 4 * sizeof(double), cap * 4 * sizeof(double)))) return rc;
    if (sweeps && (rc = copy_rows(c, sweeps, c->rb.sweeps, n * sizeof(int32_t), cap * sizeof(int32_t)))) return rc;
    for (int i = 0; i < c->ninst; ++i)
        if (st[i].err & 1) {
            set_err(c, "instance %d: non-finite charge or field", i);
            return PIC_ENONFINITE;
        }
    return PIC_OK;
}

// With the first deposit fused into the final sweep, a run that does not
// follow a completed step (new context, or after an upload) starts with the
// first deposit of its own.
static int prime(pic_ctx* c) {
    if (!c->rb.prime_next || c->primed) return PIC_OK;
    const size_t bytes = sizeof(unsigned long long) * (size_t)c->rb.racc_cells;
    HIPCHK(c, hipMemsetAsync(c->rb.racc, 0, bytes, c->stream));
    launch_mid_push0(c->rb, c->stream);
    HIPCHK(c, hipGetLastError());
    c->primed = true;
    return PIC_OK;
}

// Sweep slots per step for the next graphs: every slot the loop may need
// until a step has been observed, then the most sweeps seen so far.
static int choose_slots(const pic_ctx* c) {
    if (c->scheme == PIC_SCHEME_LEAPFROG || c->sweeps_seen == 0) return c->max_sweeps;
    return std::min(c->max_sweeps, std::max(1, c->sweeps_seen));
}

static int note_sweeps(pic_ctx* c, int64_t row0, int64_t row1) {
    if (c->scheme == PIC_SCHEME_LEAPFROG || row1 <= row0) return PIC_OK;
    const size_t n = (size_t)row1, cap = (size_t)c->rb.diag_cap;
    std::vector<int32_t> h(n * (size_t)c->ninst);
    int rc = copy_rows(c, h.data(), c->rb.sweeps, n * sizeof(int32_t), cap * sizeof(int32_t));
    if (rc) return rc;
    for (int i = 0; i < c->ninst; ++i)
        for (size_t r = (size_t)row0; r < n; ++r) c->sweeps_seen = std::max(c->sweeps_seen, (int)h[i * n + r]);
    return PIC_OK;
}

static bool use_small(const pic_ctx* c) { return c->small_ok && c->run_path != PIC_RUN_GRAPH; }

// The whole call in one persistent launch (kernels_small.hip).
static int run_small(pic_ctx* c, int64_t nsteps, double* diag, int32_t* sweeps) {
    if (nsteps > INT32_MAX - 1) {
        set_err(c, "pic_run: nsteps too large for one persistent run");
        return PIC_EINVAL;
    }
    int rc = begin_run(c, nsteps);
    if (rc) return rc;
    const RunBuffers& b = c->rb;
    SmallRun a{};
    a.I = c->hinst[0];
    a.st = b.st;
    a.X0 = b.X[0];
    a.X1 = b.X[1];
    a.V0 = b.V[0];
    a.V1 = b.V[1];
    a.E0 = b.E[0];
    a.E1 = b.E[1];
    a.tw = reinterpret_cast<const double2*>(c->tw);
    a.diag = b.diag;
    a.sweeps = b.sweeps;
    a.nsteps = nsteps;
    a.gacc = c->small_gacc;
    a.bar = c->small_bar;
    a.K = (int32_t)c->small_K;
    a.spin_limit = 1u << 22;
    HIPCHK(c, hipMemsetAsync(c->small_gacc, 0, 9 * sizeof(unsigned long long) * (size_t)a.I.N, c->stream));
    HIPCHK(c, hipMemsetAsync(c->small_bar, 0, 4 * sizeof(uint32_t), c->stream));
    HIPCHK(c, launch_small_run(a, c->scheme, c->stream));
    uint32_t bar[4] = {0, 0, 0, 0};
    HIPCHK(c, hipMemcpyAsync(bar, c->small_bar, sizeof bar, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    // the run moved the state past any charge a graph step primed
    c->primed = false;
    if (bar[2]) {
        set_err(c, "pic_run: a grid barrier of the persistent run timed out (run state undefined)");
        return PIC_EHIP;
    }
    if ((rc = note_sweeps(c, 0, nsteps))) return rc;
    return finish_run(c, nsteps, diag, sweeps);
}

int pic_set_run_path(pic_ctx* c, int32_t path, int32_t workgroups) {
    PIC_TRACE("pic_set_run_path");
    if (!c) return PIC_EINVAL;
    if (c->busy.load()) return in_flight(c);
    if (path != PIC_RUN_AUTO && path != PIC_RUN_GRAPH && path != PIC_RUN_PERSISTENT) {
        set_err(c, "pic_set_run_path: unknown path %d", (int)path);
        return PIC_EINVAL;
    }
    int64_t K = 0;
    const bool fits = c->ninst == 1 && small_eligible(c->hinst[0], c->scheme, c->precision, &K);
    if (path == PIC_RUN_PERSISTENT && !fits) {
        set_err(c, "pic_set_run_path: the persistent run takes one FP64 instance with N <= 256 and NGP "
                   "leapfrog (P <= 8192) or a Gaussian shape (P <= 16384)");
        return PIC_EINVAL;
    }
    if (workgroups != 0 && (!fits || !small_workgroups_ok(c->hinst[0], workgroups))) {
        set_err(c, "pic_set_run_path: %d workgroups do not fit this instance", (int)workgroups);
        return PIC_EINVAL;
    }
    c->run_path = path;
    if (fits) c->small_K = workgroups ? workgroups : K;
    return PIC_OK;
}

int pic_get_run_path(const pic_ctx* c, int32_t* path, int32_t* workgroups) {
    PIC_TRACE("pic_get_run_path");
    if (!c) return PIC_EINVAL;
    const bool small = use_small(c);
    if (path) *path = small ? PIC_RUN_PERSISTENT : PIC_RUN_GRAPH;
    if (workgroups) *workgroups = small ? (int32_t)c->small_K : 0;
    return PIC_OK;
}

static int run_impl(pic_ctx* c, int64_t nsteps, double* diag, int32_t* sweeps);

int pic_run(pic_ctx* c, int64_t nsteps, double* diag, int32_t* sweeps) {
    PIC_TRACE("pic_run");
    if (!c) return PIC_EINVAL;
    if (c->busy.load()) return in_flight(c);
    return run_impl(c, nsteps, diag, sweeps);
}

int pic_run_async(pic_ctx* c, int64_t nsteps, double* diag, int32_t* sweeps) {
    PIC_TRACE("pic_run_async");
    if (!c) return PIC_EINVAL;
    if (c->busy.load()) return in_flight(c);
    if (c->async_thread.joinable()) c->async_thread.join();
    c->busy.store(true);
    try {
        c->async_thread = std::thread([c, nsteps, diag, sweeps] {
            c->async_rc = run_impl(c, nsteps, diag, sweeps);
            c->busy.store(false);
        });
    } catch (...) {
        c->busy.store(false);
        set_err(c, "pic_run_async: could not start the run's thread");
        return PIC_ESTATE;
    }
    return PIC_OK;
}

int pic_sync(pic_ctx* c) {
    PIC_TRACE("pic_sync");
    if (!c) return PIC_EINVAL;
    if (!c->async_thread.joinable()) return PIC_OK;
    c->async_thread.join();
    c->busy.store(false);
    const int rc = c->async_rc;
    c->async_rc = PIC_OK;
    return rc;
}

static int run_impl(pic_ctx* c, int64_t nsteps, double* diag, int32_t* sweeps) {
    if (nsteps < 0) {
        set_err(c, "pic_run: nsteps < 0");
        return PIC_EINVAL;
    }
    if (nsteps == 0) return PIC_OK;
    HIPCHK(c, hipSetDevice(c->device));
    if (use_small(c)) return run_small(c, nsteps, diag, sweeps);
    int rc = begin_run(c, nsteps);
    if (rc) return rc;
    if ((rc = prime(c))) return rc;
    const int64_t G = kGraphSteps;
    int64_t done = 0;
    while (done < nsteps) {
        const int slots = choose_slots(c);
        // until a step has been observed, run one with every slot in place
        const int64_t chunk = c->sweeps_seen == 0 ? 1 : nsteps - done;
        pic_ctx::StepGraphs& g = c->graphs[slots];
        if (!g.one && (rc = capture(c, 1, slots, &g.one))) return rc;
        if (chunk >= G && !g.many && (rc = capture(c, G, slots, &g.many))) return rc;
        int64_t left = chunk;
        while (g.many && left >= G) {
            HIPCHK(c, hipGraphLaunch(g.many, c->stream));
            left -= G;
        }
        while (left > 0) {
            HIPCHK(c, hipGraphLaunch(g.one, c->stream));
            --left;
        }
        int32_t halt = 0;
        HIPCHK(c, hipMemcpyAsync(&halt, c->rb.halt, sizeof halt, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        HIPCHK(c, hipGetLastError());
        std::vector<InstState> st;
        if ((rc = read_states(c, st))) return rc;
        int64_t completed = (int64_t)st[0].step - st[0].run_base;
        if (halt) {
            // a step needed more sweeps than the graph holds: finish it here
            const int32_t zero = 0;
            HIPCHK(c, hipMemcpyAsync(c->rb.halt, &zero, sizeof zero, hipMemcpyHostToDevice, c->stream));
            enqueue_tail(c, c->stream, slots);
            HIPCHK(c, hipStreamSynchronize(c->stream));
            HIPCHK(c, hipGetLastError());
            completed += 1;
        }
        if ((rc = note_sweeps(c, done, completed))) return rc;
        if (completed <= done || completed > nsteps) {
            set_err(c, "pic_run: step bookkeeping lost (%lld of %lld)", (long long)completed, (long long)nsteps);
            return PIC_ESTATE;
        }
        done = completed;
    }
    return finish_run(c, nsteps, diag, sweeps);
}

int pic_run_profiled(pic_ctx* c, int64_t nsteps, double* diag, int32_t* sweeps, pic_kernel_stat* stats,
                     int32_t* n_stats) {
    PIC_TRACE("pic_run_profiled");
    if (!c) return PIC_EINVAL;
    if (c->busy.load()) return in_flight(c);
    if (nsteps <= 0) {
        set_err(c, "pic_run_profiled: nsteps must be > 0");
        return PIC_EINVAL;
    }
    HIPCHK(c, hipSetDevice(c->device));
    int rc = begin_run(c, nsteps);
    if (rc) return rc;
    if ((rc = prime(c))) return rc;
    enum { K_PUSH0, K_PUSH, K_SOLVE, K_LF1, K_LF2, K_DIAG, K_N };
    const char* names[K_N] = {"push_first", "push_sweep", "solve", "leapfrog_first", "leapfrog_second",
                              "diagnostics"};
    double total[K_N] = {0};
    int64_t count[K_N] = {0};
    hipEvent_t e0, e1;
    HIPCHK(c, hipEventCreate(&e0));
    HIPCHK(c, hipEventCreate(&e1));
    auto timed = [&](int kind, auto&& launch) -> int {
        HIPCHK(c, hipEventRecord(e0, c->stream));
        launch();
        HIPCHK(c, hipEventRecord(e1, c->stream));
        HIPCHK(c, hipEventSynchronize(e1));
        float ms = 0.f;
        HIPCHK(c, hipEventElapsedTime(&ms, e0, e1));
        total[kind] += ms;
        count[kind] += 1;
        return PIC_OK;
    };
    const RunBuffers& b = c->rb;
    std::vector<InstState> st;
    for (int64_t t = 0; t < nsteps && rc == PIC_OK; ++t) {
        if (c->scheme == PIC_SCHEME_MIDPOINT) {
            if (!b.prime_next && (rc = timed(K_PUSH0, [&] { launch_mid_push0(b, c->stream); }))) break;
            for (int k = 0; k < c->max_sweeps; ++k) {
                if ((rc = timed(K_SOLVE, [&] { launch_solve(b, k, false, c->stream); }))) break;
                if ((rc = timed(K_PUSH, [&] { launch_mid_push(b, k, false, c->stream); }))) break;
                if ((rc = read_states(c, st))) break;
                bool all_done = true;
                for (auto& s : st) all_done = all_done && s.last_sweep >= 0 && s.last_sweep <= k;
                if (all_done) break;
            }
        } else if (c->scheme == PIC_SCHEME_SIMPSON) {
            if ((rc = timed(K_PUSH0, [&] { launch_simp_first(b, c->stream); }))) break;
            for (int k = 0; k < c->max_sweeps; ++k) {
                if ((rc = timed(K_SOLVE, [&] { launch_simp_solve(b, k, false, c->stream); }))) break;
                if ((rc = timed(K_PUSH, [&] { launch_simp_push(b, k, false, c->stream); }))) break;
                if ((rc = read_states(c, st))) break;
                bool all_done = true;
                for (auto& s : st) all_done = all_done && s.last_sweep >= 0 && s.last_sweep <= k;
                if (all_done) break;
            }
        } else {
            if ((rc = timed(K_LF1, [&] { launch_lf_first(b, c->stream); }))) break;
            if ((rc = timed(K_SOLVE, [&] { launch_solve(b, 0, false, c->stream); }))) break;
            if ((rc = timed(K_LF2, [&] { launch_lf_second(b, c->stream); }))) break;
        }
        if (rc) break;
        rc = timed(K_DIAG, [&] { launch_diag(b, c->stream); });
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    if (stats && n_stats) {
        int m = 0;
        for (int k = 0; k < K_N && m < *n_stats; ++k) {
            if (!count[k]) continue;
            std::memset(&stats[m], 0, sizeof(pic_kernel_stat));
            std::snprintf(stats[m].name, sizeof stats[m].name, "%s", names[k]);
            stats[m].launches = count[k];
            stats[m].total_ms = total[k];
            ++m;
        }
        *n_stats = m;
    }
    return finish_run(c, nsteps, diag, sweeps);
}

// ---- tiling and the autotuner ----------------------------------------------

// Device buffers a run step changes: particle state (both parities), field,
// charge accumulator, loop state.  The autotuner restores them after each trial.
static std::vector<std::pair<void*, size_t>> run_state(pic_ctx* c) {
    RunBuffers& b = c->rb;
    const size_t P = (size_t)c->Ptot, G = (size_t)(c->hinst.back().goff + c->hinst.back().N);
    std::vector<std::pair<void*, size_t>> v;
    for (int j = 0; j < 2; ++j) {
        if (c->precision == PIC_PREC_FP64) {
            v.push_back({b.X[j], P * 8});
            v.push_back({b.V[j], P * 8});
            if (b.vit[j]) v.push_back({b.vit[j], P * 8});
        } else {
            v.push_back({b.Xq[j], P * 4});
            v.push_back({b.Vh[j], P * 4});
            v.push_back({b.Vl[j], P * 4});
            v.push_back({b.vf[j], P * 4});
        }
        v.push_back({b.E[j], G * 8});
    }
    v.push_back({b.E32, G * 4});
    const size_t cells = (size_t)(c->hinst.back().red0 + c->hinst.back().N);
    v.push_back({b.racc, 8 * cells * (c->precision == 0 ? 3 : 1)});
    if (c->scheme == PIC_SCHEME_SIMPSON) {
        v.push_back({b.Es1, G * 8});
        v.push_back({b.Es2[0], G * 8});
        v.push_back({b.Es2[1], G * 8});
        if (c->precision == PIC_PREC_FP64) {
            v.push_back({b.Vg, P * 8});
        } else {
            v.push_back({b.G1s, P * 4});
            v.push_back({b.E32s1, G * 4});
            v.push_back({b.E32s2, G * 4});
        }
        v.push_back({b.racc2, 8 * cells * (c->precision == 0 ? 3 : 1)});
        v.push_back({b.racc3, 8 * cells * (c->precision == 0 ? 3 : 1)});
    }
    v.push_back({b.st, sizeof(InstState) * (size_t)c->ninst});
    v.push_back({b.halt, sizeof(int32_t)});
    return v;
}

static int copy_run_state(pic_ctx* c, bool save) {
    size_t off = 0;
    for (auto& e : run_state(c)) {
        char* snap = (char*)c->snap_buf.p + off;
        if (save) HIPCHK(c, hipMemcpyAsync(snap, e.first, e.second, hipMemcpyDeviceToDevice, c->stream));
        else HIPCHK(c, hipMemcpyAsync(e.first, snap, e.second, hipMemcpyDeviceToDevice, c->stream));
        off += (e.second + 255) & ~(size_t)255;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PIC_OK;
}

int pic_set_tiling(pic_ctx* c, int64_t tile_particles, int32_t block) {
    PIC_TRACE("pic_set_tiling");
    if (!c) return PIC_EINVAL;
    if (c->busy.load()) return in_flight(c);
    if (tile_particles < 1 || block < 64 || block > 1024 || (block & (block - 1)) != 0) {
        set_err(c, "pic_set_tiling: tile_particles must be >= 1 and block a power of two in [64, 1024]");
        return PIC_EINVAL;
    }
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return layout_tiles(c, tile_particles, block);
}

int pic_get_tiling(const pic_ctx* c, int64_t* tile_particles, int32_t* block, int32_t* n_tiles) {
    PIC_TRACE("pic_get_tiling");
    if (!c) return PIC_EINVAL;
    if (tile_particles) *tile_particles = c->tileP;
    if (block) *block = c->rb.block_push;
    if (n_tiles) *n_tiles = c->rb.ntiles;
    return PIC_OK;
}

// Times trial steps (eager launches, every sweep slot in place) of the
// current state under the current tiling; the state is restored by the caller.
static int time_trial(pic_ctx* c, int steps, double* ms_per_step) {
    hipEvent_t e0, e1;
    HIPCHK(c, hipEventCreate(&e0));
    HIPCHK(c, hipEventCreate(&e1));
    const bool primed = c->primed;
    int rc = prime(c);
    if (rc == PIC_OK) {
        enqueue_step(c, c->stream, c->max_sweeps);  // warm-up
        (void)hipEventRecord(e0, c->stream);
        for (int i = 0; i < steps; ++i) enqueue_step(c, c->stream, c->max_sweeps);
        (void)hipEventRecord(e1, c->stream);
        hipError_t e = hipEventSynchronize(e1);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e != hipSuccess) {
            set_err(c, "pic_autotune: trial failed: %s", hipGetErrorString(e));
            rc = PIC_EHIP;
        }
        *ms_per_step = ms / steps;
    }
    c->primed = primed;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

int pic_autotune(pic_ctx* c, int32_t trial_steps, pic_tuning* out) {
    PIC_TRACE("pic_autotune");
    if (!c) return PIC_EINVAL;
    if (c->busy.load()) return in_flight(c);
    if (trial_steps < 1 || trial_steps > 64) {
        set_err(c, "pic_autotune: trial_steps must be in [1, 64]");
        return PIC_EINVAL;
    }
    HIPCHK(c, hipSetDevice(c->device));
    int rc = begin_run(c, trial_steps + 1);
    if (rc) return rc;
    size_t bytes = 0;
    for (auto& e : run_state(c)) bytes += (e.second + 255) & ~(size_t)255;
    if ((rc = grow(c, c->snap_buf, bytes))) return rc;
    if ((rc = copy_run_state(c, true))) return rc;
    const int64_t tile0 = c->tileP;
    const int block0 = c->rb.block_push;
    // candidates: the particles per workgroup for 1..16 tiles per CU, and
    // blocks of 256, 512 and 1024 threads
    struct Cand { int64_t tileP; int block; double ms; };
    std::vector<Cand> cands;
    const int mult[] = {1, 2, 3, 4, 6, 8, 12, 16};
    for (int m : mult)
        for (int block : {256, 512, 1024}) {
            int64_t t = (c->Ptot + (int64_t)m * c->numCU - 1) / ((int64_t)m * c->numCU);
            t = std::min<int64_t>(std::max<int64_t>(64, ((t + 63) / 64) * 64), (int64_t)1 << 20);
            int bl = block;
            while (bl > 64 && bl / 2 >= t) bl /= 2;
            bool dup = false;
            for (auto& q : cands) dup |= q.tileP == t && q.block == bl;
            if (!dup) cands.push_back({t, bl, 0.0});
        }
    double best = 0.0, ms0 = 0.0;
    int ibest = -1;
    for (size_t i = 0; i <= cands.size() && rc == PIC_OK; ++i) {
        // i == cands.size(): the tiling the context had on entry
        const int64_t t = i < cands.size() ? cands[i].tileP : tile0;
        const int bl = i < cands.size() ? cands[i].block : block0;
        double ms = 0.0;
        if ((rc = layout_tiles(c, t, bl))) break;
        if ((rc = time_trial(c, trial_steps, &ms))) break;
        rc = copy_run_state(c, false);
        if (i < cands.size()) cands[i].ms = ms;
        else ms0 = ms;
        if (i < cands.size() && (ibest < 0 || ms < best)) {
            best = ms;
            ibest = (int)i;
        }
    }
    if (rc) {
        (void)copy_run_state(c, false);
        (void)layout_tiles(c, tile0, block0);
        return rc;
    }
    // keep the entry tiling unless a candidate beats it
    const bool keep = ms0 <= best;
    if ((rc = layout_tiles(c, keep ? tile0 : cands[ibest].tileP, keep ? block0 : cands[ibest].block))) return rc;
    if (out) {
        out->tile_particles = c->tileP;
        out->block = c->rb.block_push;
        out->n_tiles = c->rb.ntiles;
        out->candidates = (int32_t)cands.size();
        out->ms_per_step = keep ? ms0 : best;
        out->entry_ms_per_step = ms0;
    }
    return PIC_OK;
}

// ---- operators ------------------------------------------------------------

static OpBuffers op_view(pic_ctx* c) {
    OpBuffers o{};
    o.N = c->cfg[0].N;
    o.log2N = ilog2i(o.N);
    o.tw = c->tw;
    o.part = c->op_part.p;
    o.dpart = (double*)c->op_dpart.p;
    o.flag = (int32_t*)c->op_flag.p;
    return o;
}

static int h2d(pic_ctx* c, DevBuf& b, const double* src, int64_t n) {
    int rc = grow(c, b, sizeof(double) * (size_t)n);
    if (rc) return rc;
    if (src) HIPCHK(c, hipMemcpyAsync(b.p, src, sizeof(double) * (size_t)n, hipMemcpyHostToDevice, c->stream));
    return PIC_OK;
}

static int d2h(pic_ctx* c, double* dst, const DevBuf& b, int64_t n) {
    HIPCHK(c, hipMemcpyAsync(dst, b.p, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipGetLastError());
    return PIC_OK;
}

static int check_shape(pic_ctx* c, int32_t shape, bool allow_ngp, bool allow_area = false) {
    if (shape == PIC_SHAPE_GAUSS6 || shape == PIC_SHAPE_GAUSS7 || (allow_ngp && shape == PIC_SHAPE_NGP) ||
        (allow_area && shape == PIC_SHAPE_AREA))
        return PIC_OK;
    set_err(c, "unsupported shape %d", shape);
    return PIC_EINVAL;
}

int pic_deposit(pic_ctx* c, const double* x, const double* y, int64_t P, double w, int32_t shape,
                double* rho) {
    PIC_TRACE("pic_deposit");
    if (!c) return PIC_EINVAL;
    if (c->busy.load()) return in_flight(c);
    if (!x || !rho || P < 1 || !(w >= 0.0) || !std::isfinite(w)) {
        set_err(c, "pic_deposit: null array, P < 1 or bad w");
        return PIC_EINVAL;
    }
    int rc = check_shape(c, shape, true, true);
    if (rc) return rc;
    HIPCHK(c, hipSetDevice(c->device));
    const int N = c->cfg[0].N;
    if ((rc = h2d(c, c->op_a, x, P))) return rc;
    if (y && (rc = h2d(c, c->op_b, y, P))) return rc;
    if ((rc = h2d(c, c->op_c, nullptr, N))) return rc;
    if ((rc = grow(c, c->op_part, (size_t)op_tiles(P) * N * 16))) return rc;
    OpBuffers o = op_view(c);
    launch_op_deposit(o, (const double*)c->op_a.p, y ? (const double*)c->op_b.p : nullptr, P, w, shape,
                      fxp_frac_bits(P, w), (double*)c->op_c.p, c->stream);
    return d2h(c, rho, c->op_c, N);
}

int pic_fieldsolve(pic_ctx* c, const double* rho, double* E) {
    PIC_TRACE("pic_fieldsolve");
    if (!c) return PIC_EINVAL;
    if (c->busy.load()) return in_flight(c);
    if (!rho || !E) {
        set_err(c, "pic_fieldsolve: null array");
        return PIC_EINVAL;
    }
    HIPCHK(c, hipSetDevice(c->device));
    const int N = c->cfg[0].N;
    int rc;
    if ((rc = h2d(c, c->op_a, rho, N))) return rc;
    if ((rc = h2d(c, c->op_c, nullptr, N))) return rc;
    OpBuffers o = op_view(c);
    launch_op_fieldsolve(o, (const double*)c->op_a.p, (double*)c->op_c.p, c->stream);
    return d2h(c, E, c->op_c, N);
}

int pic_drift_half(pic_ctx* c, double* x, const double* v, int64_t P, double dt) {
    PIC_TRACE("pic_drift_half");
    if (!c) return PIC_EINVAL;
    if (c->busy.load()) return in_flight(c);
    if (!x || !v || P < 1) {
        set_err(c, "pic_drift_half: null array or P < 1");
        return PIC_EINVAL;
    }
    HIPCHK(c, hipSetDevice(c->device));
    int rc;
    if ((rc = h2d(c, c->op_a, x, P))) return rc;
    if ((rc = h2d(c, c->op_b, v, P))) return rc;
    launch_op_drift_half((double*)c->op_a.p, (const double*)c->op_b.p, P, dt, c->stream);
    return d2h(c, x, c->op_a, P);
}

int pic_push_leapfrog(pic_ctx* c, double* x, double* v, const double* E, int64_t P, double dt,
                      int32_t shape) {
    PIC_TRACE("pic_push_leapfrog");
    if (!c) return PIC_EINVAL;
    if (c->busy.load()) return in_flight(c);
    if (!x || !v || !E || P < 1) {
        set_err(c, "pic_push_leapfrog: null array or P < 1");
        return PIC_EINVAL;
    }
    int rc = check_shape(c, shape, true);
    if (rc) return rc;
    HIPCHK(c, hipSetDevice(c->device));
    const int N = c->cfg[0].N;
    if ((rc = h2d(c, c->op_a, x, P))) return rc;
    if ((rc = h2d(c, c->op_b, v, P))) return rc;
    if ((rc = h2d(c, c->op_c, E, N))) return rc;
    OpBuffers o = op_view(c);
    launch_op_push_leapfrog(o, (double*)c->op_a.p, (double*)c->op_b.p, (const double*)c->op_c.p, P, dt, shape,
                            c->stream);
    if ((rc = d2h(c, x, c->op_a, P))) return rc;
    return d2h(c, v, c->op_b, P);
}

int pic_midpoint_position(pic_ctx* c, const double* X, const double* V, const double* v, double* x,
                          int64_t P, double dt) {
    PIC_TRACE("pic_midpoint_position");
    if (!c) return PIC_EINVAL;
    if (c->busy.load()) return in_flight(c);
    if (!X || !V || !v || !x || P < 1) {
        set_err(c, "pic_midpoint_position: null array or P < 1");
        return PIC_EINVAL;
    }
    HIPCHK(c, hipSetDevice(c->device));
    int rc;
    if ((rc = h2d(c, c->op_a, X, P))) return rc;
    if ((rc = h2d(c, c->op_b, V, P))) return rc;
    if ((rc = h2d(c, c->op_d, v, P))) return rc;
    if ((rc = h2d(c, c->op_e, nullptr, P))) return rc;
    launch_op_midpos((const double*)c->op_a.p, (const double*)c->op_b.p, (const double*)c->op_d.p,
                     (double*)c->op_e.p, P, dt, c->stream);
    return d2h(c, x, c->op_e, P);
}

int pic_push_midpoint(pic_ctx* c, const double* X, const double* V, const double* x, double* v,
                      const double* E, int64_t P, double dt, int32_t shape) {
    PIC_TRACE("pic_push_midpoint");
    if (!c) return PIC_EINVAL;
    if (c->busy.load()) return in_flight(c);
    if (!X || !V || !x || !v || !E || P < 1) {
        set_err(c, "pic_push_midpoint: null array or P < 1");
        return PIC_EINVAL;
    }
    int rc = check_shape(c, shape, false);
    if (rc) return rc;
    HIPCHK(c, hipSetDevice(c->device));
    const int N = c->cfg[0].N;
    if ((rc = h2d(c, c->op_a, X, P))) return rc;
    if ((rc = h2d(c, c->op_b, V, P))) return rc;
    if ((rc = h2d(c, c->op_d, x, P))) return rc;
    if ((rc = h2d(c, c->op_e, nullptr, P))) return rc;
    if ((rc = h2d(c, c->op_c, E, N))) return rc;
    OpBuffers o = op_view(c);
    launch_op_push_midpoint(o, (const double*)c->op_a.p, (const double*)c->op_b.p, (const double*)c->op_d.p,
                            (double*)c->op_e.p, (const double*)c->op_c.p, P, dt, shape, c->stream);
    return d2h(c, v, c->op_e, P);
}

int pic_simpson_kick(pic_ctx* c, const double* a, const double* b, const double* v_in, double* v_out,
                     const double* E, int64_t P, double coef, int32_t shape) {
    PIC_TRACE("pic_simpson_kick");
    if (!c) return PIC_EINVAL;
    if (c->busy.load()) return in_flight(c);
    if (!a || !b || !v_in || !v_out || !E || P < 1) {
        set_err(c, "pic_simpson_kick: null array or P < 1");
        return PIC_EINVAL;
    }
    int rc = check_shape(c, shape, false, true);
    if (rc) return rc;
    HIPCHK(c, hipSetDevice(c->device));
    const int N = c->cfg[0].N;
    if ((rc = h2d(c, c->op_a, a, P))) return rc;
    if ((rc = h2d(c, c->op_b, b, P))) return rc;
    if ((rc = h2d(c, c->op_d, v_in, P))) return rc;
    if ((rc = h2d(c, c->op_e, nullptr, P))) return rc;
    if ((rc = h2d(c, c->op_c, E, N))) return rc;
    OpBuffers o = op_view(c);
    launch_op_stage_kick(o, (const double*)c->op_a.p, (const double*)c->op_b.p, (const double*)c->op_d.p,
                         (double*)c->op_e.p, (const double*)c->op_c.p, P, coef, shape, c->stream);
    return d2h(c, v_out, c->op_e, P);
}

int pic_wrap(pic_ctx* c, double* x, int64_t P) {
    PIC_TRACE("pic_wrap");
    if (!c) return PIC_EINVAL;
    if (c->busy.load()) return in_flight(c);
    if (!x || P < 1) {
        set_err(c, "pic_wrap: null array or P < 1");
        return PIC_EINVAL;
    }
    HIPCHK(c, hipSetDevice(c->device));
    int rc;
    if ((rc = h2d(c, c->op_a, x, P))) return rc;
    launch_op_wrap((double*)c->op_a.p, P, c->stream);
    return d2h(c, x, c->op_a, P);
}

int pic_isapprox(pic_ctx* c, const double* F, const double* E, int32_t n, double rtol, double atol,
                 int32_t* result) {
    PIC_TRACE("pic_isapprox");
    if (!c) return PIC_EINVAL;
    if (c->busy.load()) return in_flight(c);
    if (!F || !E || !result || n < 1 || n > kTreeChunk) {
        set_err(c, "pic_isapprox: null array or n outside [1, 4096]");
        return PIC_EINVAL;
    }
    HIPCHK(c, hipSetDevice(c->device));
    int rc;
    if ((rc = h2d(c, c->op_a, F, n))) return rc;
    if ((rc = h2d(c, c->op_b, E, n))) return rc;
    if ((rc = grow(c, c->op_flag, 64))) return rc;
    launch_op_isapprox((const double*)c->op_a.p, (const double*)c->op_b.p, n, rtol, atol,
                       (int32_t*)c->op_flag.p, c->stream);
    HIPCHK(c, hipMemcpyAsync(result, c->op_flag.p, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PIC_OK;
}

int pic_diagnostics(pic_ctx* c, const double* E, const double* v, int64_t P, double W, double* out4) {
    PIC_TRACE("pic_diagnostics");
    if (!c) return PIC_EINVAL;
    if (c->busy.load()) return in_flight(c);
    if (!E || !v || !out4 || P < 1) {
        set_err(c, "pic_diagnostics: null array or P < 1");
        return PIC_EINVAL;
    }
    HIPCHK(c, hipSetDevice(c->device));
    const int N = c->cfg[0].N;
    int rc;
    if ((rc = h2d(c, c->op_a, E, N))) return rc;
    if ((rc = h2d(c, c->op_b, v, P))) return rc;
    if ((rc = h2d(c, c->op_c, nullptr, 4))) return rc;
    if ((rc = grow(c, c->op_dpart, sizeof(double) * 2 * (size_t)((P + kTreeChunk - 1) / kTreeChunk)))) return rc;
    OpBuffers o = op_view(c);
    launch_op_diagnostics(o, (const double*)c->op_a.p, (const double*)c->op_b.p, P, W, (double*)c->op_c.p,
                          c->stream);
    return d2h(c, out4, c->op_c, 4);
}

}  // extern "C"
