// qk_capi.hip -- the C ABI (include/qk.h): context, device buffers, argument
// validation, host precomputation of scheme constants, and the pricing entry
// points.  Host code only; the kernels live in qk_kernels.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <mutex>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "qk_capi_internal.hpp"
#include "qk_launch.hpp"

using namespace qk;

// ---------------------------------------------------------------------------
// Errors.
// ---------------------------------------------------------------------------
extern "C" const char *qk_strerror(qk_status s) {
    switch (s) {
    case QK_OK: return "ok";
    case QK_EINVAL: return "invalid argument";
    case QK_ENOMEM: return "out of memory budget";
    case QK_EDEVICE: return "device error";
    case QK_EAGAIN: return "queue full, retry";
    case QK_ENOTFOUND: return "unknown ticket";
    case QK_ENOTREADY: return "not ready";
    }
    return "unknown status";
}

extern "C" const char *qk_last_error(const qk_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

namespace qk {

qk_status fail(qk_ctx *ctx, qk_status s, const std::string &msg) {
    if (ctx) ctx->err = msg;
    return s;
}

qk_status hip_fail(qk_ctx *ctx, hipError_t e, const char *what) {
    return fail(ctx, QK_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

#define QK_HIP(call)                                  \
    do {                                              \
        hipError_t e_ = (call);                       \
        if (e_ != hipSuccess) return hip_fail(ctx, e_, #call); \
    } while (0)

qk_status DevBuf::reserve(qk_ctx *ctx, size_t bytes) {
    if (p != nullptr && bytes <= cap) return QK_OK;   // never OK without memory, even for 0 bytes
    if (p) {
        hipError_t e = hipFree(p);
        p = nullptr;
        cap = 0;
        if (e != hipSuccess) return hip_fail(ctx, e, "hipFree");
    }
    const size_t want = std::max<size_t>(bytes, 256);
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
        p = nullptr;
        return fail(ctx, QK_ENOMEM, "hipMalloc of " + std::to_string(want) + " bytes failed: " + hipGetErrorString(e));
    }
    cap = want;
    return QK_OK;
}

void DevBuf::release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
}

// ---------------------------------------------------------------------------
// Validation.
// ---------------------------------------------------------------------------
qk_status validate_contract(qk_ctx *ctx, const qk_contract *c) {
    if (!c) return fail(ctx, QK_EINVAL, "null contract");
    if (c->model < QK_MODEL_HESTON || c->model > QK_MODEL_BATES) return fail(ctx, QK_EINVAL, "contract.model out of range");
    if (c->scheme < QK_SCHEME_EULER || c->scheme > QK_SCHEME_QE) return fail(ctx, QK_EINVAL, "contract.scheme out of range");
    if (c->precision != QK_FP32 && c->precision != QK_FP64) return fail(ctx, QK_EINVAL, "contract.precision out of range");
    if (c->n_steps < 1) return fail(ctx, QK_EINVAL, "contract.n_steps must be >= 1");
    // Heston QE draws words 3t..3t+2 of a 32-bit block counter.
    if (c->n_steps > (1 << 30)) return fail(ctx, QK_EINVAL, "contract.n_steps must be <= 2^30");
    if (c->n_paths < 1) return fail(ctx, QK_EINVAL, "contract.n_paths must be >= 1");
    if (!(c->T > 0.0) || !std::isfinite(c->T)) return fail(ctx, QK_EINVAL, "contract.T must be positive");
    if (!(c->S0 > 0.0) || !std::isfinite(c->S0)) return fail(ctx, QK_EINVAL, "contract.S0 must be positive");
    if (!(c->v0 >= 0.0)) return fail(ctx, QK_EINVAL, "contract.v0 must be >= 0");
    if (!(c->kappa >= 0.0) || !(c->theta >= 0.0) || !(c->sigma >= 0.0))
        return fail(ctx, QK_EINVAL, "contract kappa, theta, sigma must be >= 0");
    if (!(c->rho >= -1.0 && c->rho <= 1.0)) return fail(ctx, QK_EINVAL, "contract.rho must be in [-1, 1]");
    if (c->model == QK_MODEL_BATES && (!(c->lambda_j >= 0.0) || !(c->sigma_j >= 0.0) || !std::isfinite(c->mu_j)))
        return fail(ctx, QK_EINVAL, "Bates lambda_j and sigma_j must be >= 0");
    if (c->model == QK_MODEL_BATES && c->lambda_j * (c->T / c->n_steps) > 8.0)
        return fail(ctx, QK_EINVAL, "Bates lambda_j * dt must be <= 8");
    if (!(c->price_scale > 0.0)) return fail(ctx, QK_EINVAL, "contract.price_scale must be positive");
    if (c->n_paths > ((int64_t)1 << 40)) return fail(ctx, QK_EINVAL, "contract.n_paths must be <= 2^40");
    // Path-step counts, path-store bytes (x 8) and lane products (x 64) stay
    // far inside 64 bits below this bound.
    if (c->n_paths > (MAX_PATH_STEPS / c->n_steps))
        return fail(ctx, QK_EINVAL, "contract.n_paths * n_steps must be <= 2^56");
    return QK_OK;
}

qk_status validate_payoff(qk_ctx *ctx, const qk_payoff *p) {
    if (!p) return fail(ctx, QK_EINVAL, "null payoff");
    if (p->type < QK_EUROPEAN || p->type > QK_BINARY_ASSET)
        return fail(ctx, QK_EINVAL, "payoff.type out of range (American goes through qk_price_american)");
    if (p->right != QK_CALL && p->right != QK_PUT) return fail(ctx, QK_EINVAL, "payoff.right out of range");
    if (!(p->K > 0.0) || !std::isfinite(p->K)) return fail(ctx, QK_EINVAL, "payoff.K must be positive");
    return QK_OK;
}

bool path_dependent(int type) { return !(type == QK_EUROPEAN || type == QK_BINARY_CASH || type == QK_BINARY_ASSET); }

// The uniform maps of the kernels, evaluated on the host (the FP32 one in
// float arithmetic: the product by 2^-32 is exact, the sum rounds to nearest
// even as on the device).
float uniform_fp32(uint32_t w) { return (float)w * 0x1.0p-32f + 0x1.0p-33f; }
double uniform_fp64(uint32_t w) { return ((double)w + 0.5) * 0x1.0p-32; }

// Largest word w* with U(w*) below the Poisson threshold e^{-lambda dt}: both
// uniform maps are non-decreasing in the word, so U(w) >= threshold exactly
// when w > w*; 0xFFFFFFFF when no word reaches the threshold (lambda = 0).
template <typename Real>
uint32_t jump_threshold_word(Real threshold) {
    auto U = [](uint32_t w) -> double {
        if constexpr (sizeof(Real) == 4) return (double)uniform_fp32(w);
        else return uniform_fp64(w);
    };
    if (U(0u) >= (double)threshold) return 0u;   // unreachable: lambda dt <= 8 keeps the threshold above U(0)
    uint64_t lo = 0, hi = 0x100000000ull;        // U(lo) < threshold; U(hi) >= threshold, hi = 2^32 standing for none
    while (hi - lo > 1) {
        const uint64_t mid = (lo + hi) / 2;
        if (U((uint32_t)mid) >= (double)threshold) hi = mid; else lo = mid;
    }
    return (uint32_t)lo;
}

// ---------------------------------------------------------------------------
// Scheme constants in FP64 (SDEScheme constructors, MSchemes.hpp:104-122,
// :199-216, :296-352; Bates compensator :535-550).
// ---------------------------------------------------------------------------
template <typename Real>
LaneC<Real> make_lane(const qk_contract &c) {
    const double dt = c.T / (double)c.n_steps;
    const double eps = c.precision == QK_FP32 ? 1e-6 : 1e-12;
    double q = c.q, eml = 0.0;
    if (c.model == QK_MODEL_BATES) {
        const double kJ = std::exp(c.mu_j + 0.5 * c.sigma_j * c.sigma_j) - 1.0;
        q += c.lambda_j * kJ;
        // lambda = 0 never jumps (the reference's uniforms are below 1; the
        // FP32 map reaches 1, which would pass a threshold of exactly 1).
        eml = c.lambda_j > 0.0 ? std::exp(-c.lambda_j * dt) : 2.0;
    }
    const double k = c.kappa, th = c.theta, sg = c.sigma;
    double rho = c.rho;
    LaneC<Real> L{};
    L.S0 = (Real)c.S0;
    L.v0 = (Real)c.v0;
    L.rqdt = (Real)((c.r - q) * dt);
    L.half_dt = (Real)(0.5 * dt);
    L.sqrt_dt = (Real)std::sqrt(dt);
    L.rho = (Real)rho;
    L.rho_c = (Real)std::sqrt(1.0 - rho * rho);
    L.omkdt = (Real)(1.0 - k * dt);
    L.ktdt = (Real)(k * th * dt);
    L.ssdt = (Real)(sg * std::sqrt(dt));
    L.qs2dt = (Real)(0.25 * sg * sg * dt);
    L.iden = (Real)(1.0 / (1.0 + k * dt));
    const double ekdt = std::exp(-k * dt);
    double tmf, s2f1, s2f2;
    if (k < eps) {
        tmf = 0.0;
        s2f1 = sg * sg * dt;
        s2f2 = 0.0;
    } else {
        const double ome = 1.0 - ekdt;
        tmf = th * ome;
        s2f1 = (sg * sg * ekdt / k) * ome;
        s2f2 = (th * sg * sg / (2.0 * k)) * (ome * ome);
    }
    double rds;
    if (sg < eps) {
        rds = 0.0;
        rho = 0.0;
    } else {
        rds = rho / sg;
    }
    const double omr2 = 1.0 - rho * rho;
    const double K12 = k * rds - 0.5;
    const double K2 = dt * 0.5 * K12 + rds, K3 = dt * 0.5 * omr2, K4 = dt * 0.5 * omr2;
    L.ekdt = (Real)ekdt;
    L.tmf = (Real)tmf;
    L.s2f1 = (Real)s2f1;
    L.s2f2 = (Real)s2f2;
    L.K2 = (Real)K2;
    L.K3 = (Real)K3;
    L.K4 = (Real)K4;
    L.A = (Real)(K2 + 0.5 * K4);
    L.cv = (Real)(-0.5 * K3);
    L.eps = (Real)eps;
    L.eml_dt = (Real)eml;
    L.mu_j = (Real)c.mu_j;
    L.sigma_j = (Real)c.sigma_j;
    L.jump_w = c.model == QK_MODEL_BATES ? jump_threshold_word<Real>(L.eml_dt) : 0xFFFFFFFFu;
    return L;
}

template LaneC<float> make_lane<float>(const qk_contract &);
template LaneC<double> make_lane<double>(const qk_contract &);

// ---------------------------------------------------------------------------
// Footprints (device bytes of one call), the budget test of admission.
// ---------------------------------------------------------------------------
// Saturating size arithmetic: a footprint that does not fit in size_t is
// SIZE_MAX, which no budget admits.
static size_t sat_mul(size_t a, size_t b) {
    size_t r;
    return __builtin_mul_overflow(a, b, &r) ? SIZE_MAX : r;
}
static size_t sat_add(size_t a, size_t b) {
    size_t r;
    return __builtin_add_overflow(a, b, &r) ? SIZE_MAX : r;
}

size_t footprint(const qk_contract &c, int n_lanes, int n_pay, bool american) {
    if (c.n_paths < 1 || c.n_steps < 1) return SIZE_MAX;
    const size_t rs = c.precision == QK_FP32 ? 4 : 8;
    const size_t n_tiles = (size_t)((c.n_paths + TILE - 1) / TILE);
    const size_t groups = (size_t)std::max(n_lanes, 1) * std::max(n_pay, 1);
    size_t b = sat_mul(sat_mul(groups, n_tiles), 8 + 8 + 4);          // tile moments
    b = sat_add(b, (size_t)(n_lanes + 4) * sizeof(LaneC<double>) + MAX_PAYOFFS * sizeof(PayD<double>));
    b = sat_add(b, groups * (sizeof(qk_stats) + 16) + 4096);
    b = sat_add(b, sat_mul(groups, sat_mul((size_t)merge_chunks((int64_t)n_tiles), MERGE_MOM_BYTES)));
    if (american) {
        b = sat_add(b, sat_mul(sat_mul((size_t)c.n_paths, (size_t)c.n_steps), rs));   // path store [date][path]
        b = sat_add(b, sat_mul((size_t)c.n_paths, rs));                                // cash flows
        int64_t chunk = 0;
        // Right-hand-side partials: as many as the larger of the plain call's
        // and the request server's block counts; the Gram pre-pass partials
        // and factor records; the coefficients.
        b = sat_add(b, sat_mul((size_t)c.n_steps, LSM_MAX_BASIS * 8));
        const int nb = std::max(pump_blocks(c.n_paths, LSM_SHARED_GROUP, PUMP_MAX_BLOCKS, &chunk),
                                server_pump_blocks(c.n_paths, &chunk));
        b = sat_add(b, (size_t)2 * nb * LSM_STRIDE * 8);
        b = sat_add(b, lsm_prepass_bytes(c.n_paths, c.n_steps));
    }
    return b;
}

qk_status check_budget(qk_ctx *ctx, size_t bytes) {
    if (bytes > ctx->budget)
        return fail(ctx, QK_ENOMEM, "call needs " + std::to_string(bytes) + " device bytes, budget is " +
                                        std::to_string(ctx->budget));
    return QK_OK;
}

}  // namespace qk

extern "C" size_t qk_footprint_bytes(const qk_contract *c, int32_t n_lanes, int32_t n_payoffs, int32_t american) {
    if (!c) return 0;
    return footprint(*c, std::max(n_lanes, 1), n_payoffs, american != 0);
}

// ---------------------------------------------------------------------------
// Context.
// ---------------------------------------------------------------------------
extern "C" qk_status qk_ctx_create(int device, size_t mem_budget_bytes, qk_ctx **out) {
    if (!out) return QK_EINVAL;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return QK_EDEVICE;
    if (device < 0 || device >= n) return QK_EINVAL;
    if (hipSetDevice(device) != hipSuccess) return QK_EDEVICE;
    qk_ctx *ctx = new qk_ctx();
    ctx->device = device;
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
        delete ctx;
        return QK_EDEVICE;
    }
    ctx->budget = mem_budget_bytes ? mem_budget_bytes : (size_t)(0.85 * (double)free_b);
    noalloc_init(device);   // device address of the allocation-free entry point's array
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return QK_EDEVICE;
    }
    for (auto &e : ctx->ev) {
        if (hipEventCreate(&e) != hipSuccess) {
            delete ctx;
            return QK_EDEVICE;
        }
    }
    if (ctx->counter.reserve(ctx, 2048) != QK_OK || hipMemset(ctx->counter.p, 0, 2048) != hipSuccess) {
        delete ctx;
        return QK_EDEVICE;
    }
    *out = ctx;
    return QK_OK;
}

qk_ctx::~qk_ctx() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (DevBuf *b : {&lanes, &pays, &psum, &pm2, &pitm, &disc, &scale, &stats, &greeks, &store, &cf, &coeffs,
                      &bpart, &counter, &sums, &mscratch, &payv, &hist, &tpose, &gpart, &factor})
        b->release();
    for (auto &e : ev)
        if (e) (void)hipEventDestroy(e);
    if (hstage) (void)hipHostFree(hstage);
    if (stream) (void)hipStreamDestroy(stream);
}

extern "C" qk_status qk_ctx_destroy(qk_ctx *ctx) {
    if (!ctx) return QK_EINVAL;
    (void)hipSetDevice(ctx->device);
    delete ctx;
    return QK_OK;
}

extern "C" qk_status qk_synchronize(qk_ctx *ctx) {
    if (!ctx) return QK_EINVAL;
    QK_HIP(hipStreamSynchronize(ctx->stream));
    return QK_OK;
}

extern "C" qk_status qk_last_timing(const qk_ctx *ctx, qk_timing *out) {
    if (!ctx || !out) return QK_EINVAL;
    *out = ctx->timing;
    return QK_OK;
}

// ---------------------------------------------------------------------------
// Lane pricing: the common core of qk_price / qk_price_payoffs /
// qk_price_lanes / qk_price_greeks.  Leaves per-(lane, payoff) qk_stats in
// ctx->stats on the device.
// ---------------------------------------------------------------------------
namespace qk {

static float elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.f;
    return ms;
}

// Pinned staging of at least `bytes` (kept for later calls).
static qk_status host_stage(qk_ctx *ctx, size_t bytes) {
    if (bytes <= ctx->hstage_cap) return QK_OK;
    if (ctx->hstage) (void)hipHostFree(ctx->hstage);
    ctx->hstage = nullptr;
    ctx->hstage_cap = 0;
    const size_t want = std::max<size_t>(bytes, 65536);
    hipError_t e = hipHostMalloc(&ctx->hstage, want, hipHostMallocDefault);
    if (e != hipSuccess) {
        ctx->hstage = nullptr;
        return fail(ctx, QK_ENOMEM, std::string("hipHostMalloc of the staging buffer failed: ") + hipGetErrorString(e));
    }
    ctx->hstage_cap = want;
    return QK_OK;
}

// Small device -> caller copies go through the pinned staging buffer: the
// runtime carries out a copy into or out of pageable memory in a way that
// waits for work on other streams (measured: a 0.5 ms American pricing
// waited 2.3 s behind a request server's batches on another context), a
// pinned one does not.  Queues the copy on ctx->stream, synchronises it and
// copies out; the staging buffer must already hold `bytes` (host_stage before
// any copy from it is queued: a regrowth frees the old buffer).
static qk_status copy_out(qk_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (bytes > ctx->hstage_cap) return fail(ctx, QK_EINVAL, "staging buffer too small");
    QK_HIP(hipMemcpyAsync(ctx->hstage, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    QK_HIP(hipStreamSynchronize(ctx->stream));
    std::memcpy(dst, ctx->hstage, bytes);
    return QK_OK;
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Byte layout of a lane call's constants block: lane constants, payoffs,
// per-lane discounts, per-payoff price scales (each part 256-byte aligned).
struct ConstLayout {
    size_t o_pay, o_disc, o_scale, bytes;
};
template <typename Real>
static ConstLayout const_layout(int n_lanes, int n_pay) {
    ConstLayout l;
    l.o_pay = align256((size_t)n_lanes * sizeof(LaneC<Real>));
    l.o_disc = l.o_pay + align256((size_t)n_pay * sizeof(PayD<Real>));
    l.o_scale = l.o_disc + align256((size_t)n_lanes * 8);
    l.bytes = l.o_scale + (size_t)n_pay * 8;
    return l;
}

// Writes the constants block of a lane call into host memory hb (at least
// const_layout(...).bytes long): scheme constants precomputed in FP64, the
// payoffs, e^{-rT} per lane and the price scale per payoff (1 for cash
// binaries, OptionPricer.hpp:323-332).
template <typename Real>
static void pack_consts(const qk_contract *lanes, int n_lanes, const qk_payoff *pays, int n_pay,
                        const ConstLayout &lay, char *hb) {
    LaneC<Real> *hl = (LaneC<Real> *)hb;
    PayD<Real> *hp = (PayD<Real> *)(hb + lay.o_pay);
    double *hdisc = (double *)(hb + lay.o_disc), *hscale = (double *)(hb + lay.o_scale);
    for (int i = 0; i < n_lanes; ++i) {
        hl[i] = make_lane<Real>(lanes[i]);
        hdisc[i] = std::exp(-lanes[i].r * lanes[i].T);
    }
    for (int j = 0; j < n_pay; ++j) {
        hp[j] = PayD<Real>{pays[j].type, pays[j].right, (Real)pays[j].K, (Real)pays[j].B};
        hscale[j] = pays[j].type == QK_BINARY_CASH ? 1.0 : lanes[0].price_scale;
    }
}

// Device buffers of one lane call.
struct LaneBufs {
    char *consts;          // the constants block (ConstLayout)
    double *psum, *pm2;    // tile moments [(lane * n_pay + j) * n_tiles + tile]
    uint32_t *pitm;
    void *mscratch;        // first-stage merge results
    qk_stats *stats;       // [n_lanes * n_pay]
};

// Launches the path kernel(s) and the merge of a lane call on ctx->stream,
// once the constants block is on the device (or its copy is queued before).
// path0 numbers the first path in the random stream (a multiple of 64); with
// chunks_only the merge stops after its first stage, leaving the chunk moments
// [lane * n_pay + payoff][chunk] in B.mscratch.
template <typename Real>
static qk_status launch_lanes(qk_ctx *ctx, const qk_contract *lanes, int n_lanes, const qk_payoff *pays, int n_pay,
                              const ConstLayout &lay, const LaneBufs &B, Real *pay_out, Real *store = nullptr,
                              int64_t path0 = 0, bool chunks_only = false) {
    const qk_contract &c0 = lanes[0];
    const bool pathdep = std::any_of(pays, pays + n_pay, [](const qk_payoff &p) { return path_dependent(p.type); });
    const int L = qk_lanes_per_thread(ctx, c0.precision, c0.scheme, n_lanes);
    const int64_t n_tiles = (c0.n_paths + TILE - 1) / TILE;
    hipStream_t s = ctx->stream;
    PathArgs<Real> a{};
    a.lanes = (const LaneC<Real> *)B.consts;
    a.pays = (const PayD<Real> *)(B.consts + lay.o_pay);
    a.n_lanes = n_lanes;
    a.n_pay = n_pay;
    a.n_steps = c0.n_steps;
    a.n_paths = c0.n_paths;
    a.n_tiles = n_tiles;
    a.path0 = path0;
    a.key0 = (uint32_t)c0.seed;
    a.key1 = (uint32_t)(c0.seed >> 32);
    a.itm_eps = (Real)(c0.precision == QK_FP32 ? 1e-6 : 1e-12);   // OptionPricer.hpp:316
    a.part_sum = B.psum;
    a.part_m2 = B.pm2;
    a.part_itm = B.pitm;
    a.store = store;
    a.pay_out = pay_out;

    QK_HIP(hipEventRecord(ctx->ev[0], s));
    if (store != nullptr) {
        // One lane: the path kernel also writes S after every step to the
        // [step][path] store (the path-dependent form, S multiplied step by
        // step, whatever the payoff).
        QK_HIP(launch_paths<Real>(a, c0.model, c0.scheme, true, true, 1, s));
    }
    // Whole groups of L lanes per thread; the lanes left over go to a launch
    // with L/2 (then 1) lanes per thread instead of being padded.
    for (int done = store != nullptr ? n_lanes : 0, Lk = L; done < n_lanes;) {
        while (Lk > 1 && n_lanes - done < Lk) Lk >>= 1;
        const int cnt = (n_lanes - done) / Lk * Lk;
        PathArgs<Real> ak = a;
        const int64_t off = (int64_t)done * n_pay * n_tiles;
        ak.lanes = a.lanes + done;
        ak.n_lanes = cnt;
        ak.part_sum = a.part_sum + off;
        ak.part_m2 = a.part_m2 + off;
        ak.part_itm = a.part_itm + off;
        ak.pay_out = done == 0 ? a.pay_out : nullptr;
        QK_HIP(launch_paths<Real>(ak, c0.model, c0.scheme, pathdep, false, Lk, s));
        done += cnt;
    }
    QK_HIP(hipEventRecord(ctx->ev[1], s));
    QK_HIP(launch_merge(a.part_sum, a.part_m2, a.part_itm, n_tiles, c0.n_paths, n_lanes, n_pay,
                        (const double *)(B.consts + lay.o_disc), (const double *)(B.consts + lay.o_scale), B.mscratch,
                        B.stats, s, chunks_only));
    QK_HIP(hipEventRecord(ctx->ev[2], s));
    return QK_OK;
}

// The allocating lane call: buffers from the ctx (grown as needed), constants
// through the pinned staging buffer.  Leaves per-(lane, payoff) qk_stats in
// ctx->stats on the device.
template <typename Real>
static qk_status run_lanes(qk_ctx *ctx, const qk_contract *lanes, int n_lanes, const qk_payoff *pays, int n_pay,
                           Real *pay_out = nullptr, Real *store = nullptr, int64_t path0 = 0,
                           bool chunks_only = false) {
    const ConstLayout lay = const_layout<Real>(n_lanes, n_pay);
    const int64_t n_tiles = (lanes[0].n_paths + TILE - 1) / TILE;
    const size_t groups = (size_t)n_lanes * n_pay;
    qk_status st;
    // Room for the results too, so that qk_price_lanes reuses the buffer as is.
    if ((st = host_stage(ctx, std::max(lay.bytes, sizeof(qk_stats) * groups))) != QK_OK) return st;
    if ((st = ctx->lanes.reserve(ctx, lay.bytes)) != QK_OK) return st;
    if ((st = ctx->psum.reserve(ctx, groups * n_tiles * 8)) != QK_OK) return st;
    if ((st = ctx->pm2.reserve(ctx, groups * n_tiles * 8)) != QK_OK) return st;
    if ((st = ctx->pitm.reserve(ctx, groups * n_tiles * 4)) != QK_OK) return st;
    if ((st = ctx->stats.reserve(ctx, groups * sizeof(qk_stats))) != QK_OK) return st;
    if ((st = ctx->mscratch.reserve(ctx, groups * merge_chunks(n_tiles) * MERGE_MOM_BYTES)) != QK_OK) return st;
    // ev[0] follows the last copy out of the staging buffer on the stream.
    QK_HIP(hipEventSynchronize(ctx->ev[0]));
    pack_consts<Real>(lanes, n_lanes, pays, n_pay, lay, (char *)ctx->hstage);
    QK_HIP(hipMemcpyAsync(ctx->lanes.p, ctx->hstage, lay.bytes, hipMemcpyHostToDevice, ctx->stream));
    const LaneBufs B{(char *)ctx->lanes.p, (double *)ctx->psum.p, (double *)ctx->pm2.p, (uint32_t *)ctx->pitm.p,
                     ctx->mscratch.p, (qk_stats *)ctx->stats.p};
    return launch_lanes<Real>(ctx, lanes, n_lanes, pays, n_pay, lay, B, pay_out, store, path0, chunks_only);
}

static qk_status finish_timing(qk_ctx *ctx, int64_t path_steps, bool has_lsm) {
    QK_HIP(hipEventSynchronize(ctx->ev[2]));
    ctx->timing = qk_timing{};
    ctx->timing.path_kernel_ms = elapsed(ctx->ev[0], ctx->ev[1]);
    if (has_lsm) {
        ctx->timing.lsm_backward_ms = elapsed(ctx->ev[1], ctx->ev[3]);
        ctx->timing.reduce_ms = elapsed(ctx->ev[3], ctx->ev[2]);
    } else {
        ctx->timing.reduce_ms = elapsed(ctx->ev[1], ctx->ev[2]);
    }
    ctx->timing.total_ms = elapsed(ctx->ev[0], ctx->ev[2]);
    ctx->timing.path_steps = path_steps;
    return QK_OK;
}

static qk_status validate_lanes(qk_ctx *ctx, const qk_contract *lanes, int32_t n_lanes, const qk_payoff *pays,
                                int32_t n_pay) {
    if (!lanes || n_lanes < 1 || n_lanes > MAX_LANES) return fail(ctx, QK_EINVAL, "n_lanes must be in [1, 64]");
    if (!pays || n_pay < 1 || n_pay > MAX_PAYOFFS) return fail(ctx, QK_EINVAL, "n_payoffs must be in [1, 16]");
    for (int i = 0; i < n_lanes; ++i) {
        qk_status st = validate_contract(ctx, &lanes[i]);
        if (st != QK_OK) return st;
        const qk_contract &a = lanes[0], &b = lanes[i];
        if (a.model != b.model || a.scheme != b.scheme || a.precision != b.precision || a.n_steps != b.n_steps ||
            a.n_paths != b.n_paths || a.seed != b.seed || a.price_scale != b.price_scale)
            return fail(ctx, QK_EINVAL,
                        "lanes must share model, scheme, precision, n_steps, n_paths, seed and price_scale");
    }
    for (int j = 0; j < n_pay; ++j) {
        qk_status st = validate_payoff(ctx, &pays[j]);
        if (st != QK_OK) return st;
    }
    return QK_OK;
}

}  // namespace qk

extern "C" qk_status qk_set_schedule(qk_ctx *ctx, const qk_schedule_entry *entries, int32_t n) {
    if (!ctx) return QK_EINVAL;
    if (n < 0 || n > 256 || (n > 0 && !entries)) return fail(ctx, QK_EINVAL, "schedule: 0 <= n <= 256 entries");
    for (int32_t i = 0; i < n; ++i) {
        const qk_schedule_entry &e = entries[i];
        if (e.lanes_per_thread != 1 && e.lanes_per_thread != 2 && e.lanes_per_thread != 4 &&
            !(e.lanes_per_thread == 8 && e.precision == QK_FP32))
            return fail(ctx, QK_EINVAL, "schedule: lanes_per_thread must be 1, 2 or 4 (or 8 in FP32)");
        if (e.n_lanes < 1 || e.n_lanes > MAX_LANES) return fail(ctx, QK_EINVAL, "schedule: n_lanes must be in [1, 64]");
    }
    ctx->schedule.assign(entries, entries + n);
    return QK_OK;
}

extern "C" int32_t qk_lanes_per_thread(const qk_ctx *ctx, int32_t precision, int32_t scheme, int32_t n_lanes) {
    if (ctx)
        for (const qk_schedule_entry &e : ctx->schedule)
            if (e.precision == precision && e.scheme == scheme && e.n_lanes == n_lanes) return e.lanes_per_thread;
    return n_lanes >= 4 ? 4 : 1;
}

extern "C" qk_status qk_price_lanes(qk_ctx *ctx, const qk_contract *lanes, int32_t n_lanes, const qk_payoff *pays,
                                    int32_t n_pay, qk_stats *out) {
    if (!ctx || !out) return fail(ctx, QK_EINVAL, "null argument");
    qk_status st = validate_lanes(ctx, lanes, n_lanes, pays, n_pay);
    if (st != QK_OK) return st;
    if ((st = check_budget(ctx, footprint(lanes[0], n_lanes, n_pay, false))) != QK_OK) return st;
    QK_HIP(hipSetDevice(ctx->device));
    st = lanes[0].precision == QK_FP32 ? run_lanes<float>(ctx, lanes, n_lanes, pays, n_pay)
                                       : run_lanes<double>(ctx, lanes, n_lanes, pays, n_pay);
    if (st != QK_OK) return st;
    // Results through the pinned staging (its constants' copy precedes this
    // one on the stream, so the buffer is free to reuse).
    const size_t n_out = sizeof(qk_stats) * n_lanes * n_pay;
    if ((st = host_stage(ctx, n_out)) != QK_OK) return st;
    QK_HIP(hipMemcpyAsync(ctx->hstage, ctx->stats.p, n_out, hipMemcpyDeviceToHost, ctx->stream));
    QK_HIP(hipStreamSynchronize(ctx->stream));
    std::memcpy(out, ctx->hstage, n_out);
    return finish_timing(ctx, lanes[0].n_paths * (int64_t)lanes[0].n_steps * n_lanes, false);
}

extern "C" qk_status qk_price_payoffs(qk_ctx *ctx, const qk_contract *c, const qk_payoff *pays, int32_t n_pay,
                                      qk_stats *out) {
    return qk_price_lanes(ctx, c, 1, pays, n_pay, out);
}

extern "C" qk_status qk_price(qk_ctx *ctx, const qk_contract *c, const qk_payoff *p, qk_stats *out) {
    return qk_price_lanes(ctx, c, 1, p, 1, out);
}

// ---------------------------------------------------------------------------
// Batches as path windows (run_forward_all_mc_batches, MCEngine.hpp:489-520):
// a window's first-stage merge chunks out, the second stage over every
// window's chunks on the device.
// ---------------------------------------------------------------------------
static_assert(sizeof(qk_moments) == MERGE_MOM_BYTES, "qk_moments is the merge chunk record");
static_assert(QK_WINDOW_CHUNK_PATHS == (int64_t)MERGE_CHUNK * TILE, "a merge chunk holds 65,536 paths");

extern "C" qk_status qk_price_window(qk_ctx *ctx, const qk_contract *c, int64_t path0, const qk_payoff *pays,
                                     int32_t n_pay, qk_moments *chunks) {
    if (!ctx || !chunks) return fail(ctx, QK_EINVAL, "null argument");
    qk_status st = validate_lanes(ctx, c, 1, pays, n_pay);
    if (st != QK_OK) return st;
    if (path0 < 0 || path0 % TILE != 0) return fail(ctx, QK_EINVAL, "path0 must be a non-negative multiple of 64");
    if (path0 > MAX_PATH_STEPS - c->n_paths) return fail(ctx, QK_EINVAL, "path0 + n_paths must be <= 2^56");
    if ((st = check_budget(ctx, footprint(*c, 1, n_pay, false))) != QK_OK) return st;
    QK_HIP(hipSetDevice(ctx->device));
    st = c->precision == QK_FP32 ? run_lanes<float>(ctx, c, 1, pays, n_pay, nullptr, nullptr, path0, true)
                                 : run_lanes<double>(ctx, c, 1, pays, n_pay, nullptr, nullptr, path0, true);
    if (st != QK_OK) return st;
    const int64_t n_tiles = (c->n_paths + TILE - 1) / TILE;
    const size_t bytes = (size_t)n_pay * (size_t)merge_chunks(n_tiles) * MERGE_MOM_BYTES;
    QK_HIP(hipStreamSynchronize(ctx->stream));           // nothing in flight reads the staging buffer
    if ((st = host_stage(ctx, bytes)) != QK_OK) return st;
    if ((st = copy_out(ctx, chunks, ctx->mscratch.p, bytes)) != QK_OK) return st;
    return finish_timing(ctx, c->n_paths * (int64_t)c->n_steps, false);
}

extern "C" qk_status qk_merge_windows(qk_ctx *ctx, const qk_contract *c, const qk_payoff *pays, int32_t n_pay,
                                      const qk_moments *chunks, int64_t n_chunks, qk_stats *out) {
    if (!ctx || !chunks || !out) return fail(ctx, QK_EINVAL, "null argument");
    qk_status st = validate_lanes(ctx, c, 1, pays, n_pay);
    if (st != QK_OK) return st;
    if (n_chunks < 1 || n_chunks > ((int64_t)1 << 40) / n_pay)
        return fail(ctx, QK_EINVAL, "n_chunks must be in [1, 2^40 / n_payoffs]");
    const size_t cbytes = (size_t)n_pay * (size_t)n_chunks * MERGE_MOM_BYTES;
    if ((st = check_budget(ctx, cbytes + 4096 + sizeof(qk_stats) * n_pay)) != QK_OK) return st;
    QK_HIP(hipSetDevice(ctx->device));
    // The discount and price scales as the lane call computes them.
    const ConstLayout lay = c->precision == QK_FP32 ? const_layout<float>(1, n_pay) : const_layout<double>(1, n_pay);
    // constants at 0, the chunks after them, both from the pinned staging buffer
    const size_t o_chunks = (lay.bytes + 255) / 256 * 256;
    QK_HIP(hipEventSynchronize(ctx->ev[0]));
    if ((st = host_stage(ctx, std::max(o_chunks + cbytes, sizeof(qk_stats) * n_pay))) != QK_OK) return st;
    if ((st = ctx->lanes.reserve(ctx, lay.bytes)) != QK_OK) return st;
    if ((st = ctx->mscratch.reserve(ctx, cbytes)) != QK_OK) return st;
    if ((st = ctx->stats.reserve(ctx, sizeof(qk_stats) * n_pay)) != QK_OK) return st;
    if (c->precision == QK_FP32) pack_consts<float>(c, 1, pays, n_pay, lay, (char *)ctx->hstage);
    else pack_consts<double>(c, 1, pays, n_pay, lay, (char *)ctx->hstage);
    std::memcpy((char *)ctx->hstage + o_chunks, chunks, cbytes);
    QK_HIP(hipMemcpyAsync(ctx->lanes.p, ctx->hstage, lay.bytes, hipMemcpyHostToDevice, ctx->stream));
    QK_HIP(hipMemcpyAsync(ctx->mscratch.p, (char *)ctx->hstage + o_chunks, cbytes, hipMemcpyHostToDevice,
                          ctx->stream));
    QK_HIP(hipEventRecord(ctx->ev[0], ctx->stream));
    const char *cb = (const char *)ctx->lanes.p;
    QK_HIP(launch_merge_final(ctx->mscratch.p, n_chunks, n_pay, n_pay, (const double *)(cb + lay.o_disc),
                              (const double *)(cb + lay.o_scale), (qk_stats *)ctx->stats.p, ctx->stream));
    QK_HIP(hipMemcpyAsync(ctx->hstage, ctx->stats.p, sizeof(qk_stats) * n_pay, hipMemcpyDeviceToHost, ctx->stream));
    QK_HIP(hipStreamSynchronize(ctx->stream));
    std::memcpy(out, ctx->hstage, sizeof(qk_stats) * n_pay);
    return QK_OK;
}

// ---------------------------------------------------------------------------
// Pricing that keeps its paths (the reference's StorePaths kernel,
// MSolver.hpp:117-119, taken when an output file is named, Config.hpp:622,
// :750-758), and the copy of stored paths out in the ".paths" payload layout.
// ---------------------------------------------------------------------------
extern "C" qk_status qk_price_store_paths(qk_ctx *ctx, const qk_contract *c, const qk_payoff *p, qk_stats *out) {
    if (!ctx || !out) return fail(ctx, QK_EINVAL, "null argument");
    qk_status st = validate_lanes(ctx, c, 1, p, 1);
    if (st != QK_OK) return st;
    if ((st = check_budget(ctx, footprint(*c, 1, 1, true))) != QK_OK) return st;
    QK_HIP(hipSetDevice(ctx->device));
    const size_t rs = c->precision == QK_FP32 ? 4 : 8;
    ctx->store_paths = 0;
    if ((st = ctx->store.reserve(ctx, (size_t)c->n_paths * c->n_steps * rs)) != QK_OK) return st;
    st = c->precision == QK_FP32 ? run_lanes<float>(ctx, c, 1, p, 1, nullptr, (float *)ctx->store.p)
                                 : run_lanes<double>(ctx, c, 1, p, 1, nullptr, (double *)ctx->store.p);
    if (st != QK_OK) return st;
    ctx->store_paths = c->n_paths;
    ctx->store_steps = c->n_steps;
    ctx->store_prec = c->precision;
    QK_HIP(hipMemcpyAsync(ctx->hstage, ctx->stats.p, sizeof(qk_stats), hipMemcpyDeviceToHost, ctx->stream));
    QK_HIP(hipStreamSynchronize(ctx->stream));
    std::memcpy(out, ctx->hstage, sizeof(qk_stats));
    return finish_timing(ctx, c->n_paths * (int64_t)c->n_steps, false);
}

extern "C" qk_status qk_copy_stored_paths(qk_ctx *ctx, int64_t path0, int64_t n, void *out) {
    if (!ctx || !out) return fail(ctx, QK_EINVAL, "null argument");
    if (ctx->store_paths <= 0) return fail(ctx, QK_EINVAL, "no stored path set on this ctx");
    if (path0 < 0 || n < 1 || path0 > ctx->store_paths - n)
        return fail(ctx, QK_EINVAL, "path window outside the stored path set");
    const size_t rs = ctx->store_prec == QK_FP32 ? 4 : 8;
    const int32_t T = ctx->store_steps;
    // Windows of at most 64 MiB through the transpose buffer.
    const int64_t win = std::max<int64_t>(1, (int64_t)(64u << 20) / ((int64_t)T * (int64_t)rs));
    qk_status st;
    QK_HIP(hipSetDevice(ctx->device));
    if ((st = ctx->tpose.reserve(ctx, (size_t)std::min(win, n) * T * rs)) != QK_OK) return st;
    for (int64_t d = 0; d < n; d += win) {
        const int64_t k = std::min(win, n - d);
        if (rs == 4)
            QK_HIP(launch_transpose_paths<float>((const float *)ctx->store.p, ctx->store_paths, T, path0 + d, k,
                                                 (float *)ctx->tpose.p, ctx->stream));
        else
            QK_HIP(launch_transpose_paths<double>((const double *)ctx->store.p, ctx->store_paths, T, path0 + d, k,
                                                  (double *)ctx->tpose.p, ctx->stream));
        QK_HIP(hipMemcpyAsync((char *)out + (size_t)d * T * rs, ctx->tpose.p, (size_t)k * T * rs,
                              hipMemcpyDeviceToHost, ctx->stream));
        QK_HIP(hipStreamSynchronize(ctx->stream));
    }
    return QK_OK;
}

// ---------------------------------------------------------------------------
// Allocation-free pricing: qk_price_payoffs over a fixed device array.
//
// `scratch` holds every device buffer of the call (constants block, tile
// moments, merge results, result structs).  Each device that loads the
// library holds its own copy until the process ends.  In host code the name
// `scratch` is only a host-side stand-in: kernels get the device address that
// hipGetSymbolAddress returned for the current device (noalloc_ptr), never the
// name.  Every call on a device shares that device's array, so calls that
// overlap on one device (from several threads or several ctxs) race.  Nothing
// in it needs clearing: the path kernels write every tile moment the merge
// reads, and the merge writes every result.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[QK_NOALLOC_SCRATCH_BYTES];

namespace qk {
namespace {
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex noalloc_mu;
void *noalloc_ptr[NOALLOC_MAX_DEVICES] = {};
hipError_t noalloc_err[NOALLOC_MAX_DEVICES] = {};
}  // namespace

// Looks up the device address of `scratch` on `device`, which must be the
// current device; once per device (qk_ctx_create calls it).
void noalloc_init(int device) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return;
    std::lock_guard<std::mutex> lk(noalloc_mu);
    if (noalloc_ptr[device] != nullptr) return;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e == hipSuccess && p != nullptr) noalloc_ptr[device] = p;
    else noalloc_err[device] = e == hipSuccess ? hipErrorInvalidSymbol : e;
}

static char *noalloc_base(qk_ctx *ctx, qk_status *st) {
    const int d = ctx->device;
    if (d < 0 || d >= NOALLOC_MAX_DEVICES) {
        *st = fail(ctx, QK_EDEVICE, "no-allocation scratch: device index beyond 63");
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(noalloc_mu);
    if (noalloc_ptr[d] == nullptr) {
        *st = hip_fail(ctx, noalloc_err[d], "hipGetSymbolAddress(scratch)");
        return nullptr;
    }
    *st = QK_OK;
    return (char *)noalloc_ptr[d];
}
}  // namespace qk

extern "C" qk_status qk_price_payoffs_noalloc(qk_ctx *ctx, const qk_contract *c, const qk_payoff *pays,
                                              int32_t n_pay, qk_stats *out) {
    if (!ctx || !out) return fail(ctx, QK_EINVAL, "null argument");
    qk_status st = validate_lanes(ctx, c, 1, pays, n_pay);
    if (st != QK_OK) return st;
    const bool f32 = c->precision == QK_FP32;
    const ConstLayout lay = f32 ? const_layout<float>(1, n_pay) : const_layout<double>(1, n_pay);
    // Carve the array: constants, results, tile moments, merge results.
    const size_t n_tiles = (size_t)((c->n_paths + TILE - 1) / TILE);
    const size_t groups = (size_t)n_pay;
    const size_t o_stats = align256(lay.bytes);
    const size_t o_sum = o_stats + align256(groups * sizeof(qk_stats));
    const size_t mom = groups * n_tiles;                 // <= 16 * 2^34: no overflow
    const size_t o_m2 = o_sum + align256(mom * 8);
    const size_t o_itm = o_m2 + align256(mom * 8);
    const size_t o_merge = o_itm + align256(mom * 4);
    const size_t need = o_merge + groups * (size_t)merge_chunks((int64_t)n_tiles) * MERGE_MOM_BYTES;
    if (need > QK_NOALLOC_SCRATCH_BYTES)
        return fail(ctx, QK_ENOMEM, "qk_price_payoffs_noalloc needs " + std::to_string(need) +
                                        " bytes of scratch, the fixed array holds " +
                                        std::to_string((size_t)QK_NOALLOC_SCRATCH_BYTES) +
                                        " (about 20 bytes per 64 paths per payoff)");
    QK_HIP(hipSetDevice(ctx->device));
    char *base = noalloc_base(ctx, &st);
    if (st != QK_OK) return st;
    // The constants block is built on the stack: this call allocates nothing,
    // on the device or the host.
    alignas(256) char hb[NOALLOC_CONST_BYTES];
    if (lay.bytes > sizeof(hb)) return fail(ctx, QK_EINVAL, "constants block too large");
    if (f32) pack_consts<float>(c, 1, pays, n_pay, lay, hb);
    else pack_consts<double>(c, 1, pays, n_pay, lay, hb);
    QK_HIP(hipMemcpyAsync(base, hb, lay.bytes, hipMemcpyHostToDevice, ctx->stream));
    const LaneBufs B{base, (double *)(base + o_sum), (double *)(base + o_m2), (uint32_t *)(base + o_itm),
                     base + o_merge, (qk_stats *)(base + o_stats)};
    st = f32 ? launch_lanes<float>(ctx, c, 1, pays, n_pay, lay, B, nullptr)
             : launch_lanes<double>(ctx, c, 1, pays, n_pay, lay, B, nullptr);
    if (st != QK_OK) {
        (void)hipStreamSynchronize(ctx->stream);   // hb must outlive any queued copy
        return st;
    }
    QK_HIP(hipMemcpyAsync(out, B.stats, groups * sizeof(qk_stats), hipMemcpyDeviceToHost, ctx->stream));
    QK_HIP(hipStreamSynchronize(ctx->stream));
    return finish_timing(ctx, c->n_paths * (int64_t)c->n_steps, false);
}

// ---------------------------------------------------------------------------
// Greeks: the reference's 12-pricing stencil (MCEngine.hpp:576-735) as lanes
// of one launch under common random numbers, differenced in FP64 on the device.
// ---------------------------------------------------------------------------
extern "C" qk_status qk_price_greeks(qk_ctx *ctx, const qk_contract *c, const qk_payoff *pays, int32_t n_pay,
                                     const qk_bumps *b, qk_greeks *out) {
    if (!ctx || !out || !b) return fail(ctx, QK_EINVAL, "null argument");
    qk_status st = validate_contract(ctx, c);
    if (st != QK_OK) return st;
    const double vol0 = std::sqrt(c->v0);
    if ((b->delta_gamma || b->vanna) && !(b->rel_S > 0.0 && b->rel_S <= 0.05))
        return fail(ctx, QK_EINVAL, "rel_S must be in (0, 0.05]");
    // abs_vol = sqrt(v0) is allowed: the down-bump is v0 = 0 (Config.hpp:518).
    if ((b->vega_vomma || b->vanna) && !(b->abs_vol > 0.0 && b->abs_vol <= vol0))
        return fail(ctx, QK_EINVAL, "abs_vol must be in (0, sqrt(v0)]");
    if (b->rho && !(b->abs_r > 0.0 && b->abs_r <= 0.05)) return fail(ctx, QK_EINVAL, "abs_r must be in (0, 0.05]");
    if (b->theta && !(b->abs_T > 0.0 && b->abs_T < c->T)) return fail(ctx, QK_EINVAL, "abs_T must be in (0, T)");
    std::vector<qk_contract> L;
    GreekMap m{};
    m.base = 0;
    m.s_up = m.s_dn = m.v_up = m.v_dn = m.r_up = m.r_dn = m.t_dn = m.t_up = m.uu = m.du = m.ud = m.dd = -1;
    L.push_back(*c);
    const double h_S = c->S0 * b->rel_S;
    m.h_S_real = c->price_scale * c->S0 * b->rel_S;
    m.h_vol = b->abs_vol;
    m.h_r = b->abs_r;
    m.h_T = b->abs_T;
    const double vu = (vol0 + b->abs_vol) * (vol0 + b->abs_vol), vd = (vol0 - b->abs_vol) * (vol0 - b->abs_vol);
    auto add = [&](auto f) {
        qk_contract x = *c;
        f(x);
        L.push_back(x);
        return (int32_t)L.size() - 1;
    };
    if (b->delta_gamma) {
        m.s_up = add([&](qk_contract &x) { x.S0 += h_S; });
        m.s_dn = add([&](qk_contract &x) { x.S0 -= h_S; });
    }
    if (b->vega_vomma) {
        m.v_up = add([&](qk_contract &x) { x.v0 = vu; });
        m.v_dn = add([&](qk_contract &x) { x.v0 = vd; });
    }
    if (b->rho) {
        m.r_up = add([&](qk_contract &x) { x.r += b->abs_r; });
        m.r_dn = add([&](qk_contract &x) { x.r -= b->abs_r; });
    }
    if (b->theta) {
        m.t_dn = add([&](qk_contract &x) { x.T -= b->abs_T; });
        if (b->theta_central) m.t_up = add([&](qk_contract &x) { x.T += b->abs_T; });
    }
    if (b->vanna) {
        m.uu = add([&](qk_contract &x) { x.S0 += h_S; x.v0 = vu; });
        m.du = add([&](qk_contract &x) { x.S0 -= h_S; x.v0 = vu; });
        m.ud = add([&](qk_contract &x) { x.S0 += h_S; x.v0 = vd; });
        m.dd = add([&](qk_contract &x) { x.S0 -= h_S; x.v0 = vd; });
    }
    m.n_lanes = (int32_t)L.size();
    st = validate_lanes(ctx, L.data(), (int32_t)L.size(), pays, n_pay);
    if (st != QK_OK) return st;
    if ((st = check_budget(ctx, footprint(*c, (int)L.size(), n_pay, false))) != QK_OK) return st;
    QK_HIP(hipSetDevice(ctx->device));
    st = c->precision == QK_FP32 ? run_lanes<float>(ctx, L.data(), (int)L.size(), pays, n_pay)
                                 : run_lanes<double>(ctx, L.data(), (int)L.size(), pays, n_pay);
    if (st != QK_OK) return st;
    if ((st = ctx->greeks.reserve(ctx, sizeof(qk_greeks) * n_pay)) != QK_OK) return st;
    QK_HIP(launch_greeks((const qk_stats *)ctx->stats.p, n_pay, m, (qk_greeks *)ctx->greeks.p, ctx->stream));
    QK_HIP(hipEventRecord(ctx->ev[2], ctx->stream));
    if ((st = copy_out(ctx, out, ctx->greeks.p, sizeof(qk_greeks) * n_pay)) != QK_OK) return st;
    return finish_timing(ctx, c->n_paths * (int64_t)c->n_steps * (int64_t)L.size(), false);
}

// ---------------------------------------------------------------------------
// Path store + Longstaff-Schwartz.
// ---------------------------------------------------------------------------
namespace qk {

// Staging-buffer slots of the path-store calls: generate_store's lane
// constants at 0, run_american's merge discount and scale after them.
constexpr size_t STAGE_LANE_BYTES = 1024, STAGE_AUX = STAGE_LANE_BYTES;

template <typename Real>
static qk_status generate_store(qk_ctx *ctx, const qk_contract &c, int64_t path0 = 0) {
    const LaneC<Real> hl = make_lane<Real>(c);
    qk_status st;
    if ((st = ctx->lanes.reserve(ctx, sizeof(LaneC<Real>))) != QK_OK) return st;
    ctx->store_paths = 0;   // the store holds no complete path set until the launch is queued
    if ((st = ctx->store.reserve(ctx, (size_t)c.n_paths * c.n_steps * sizeof(Real))) != QK_OK) return st;
    // staging buffer: the lane at 0 (callers keep 0 .. STAGE_LANE_BYTES free of copies in flight)
    static_assert(sizeof(hl) <= STAGE_LANE_BYTES, "lane constants exceed their staging slot");
    if ((st = host_stage(ctx, STAGE_LANE_BYTES)) != QK_OK) return st;
    std::memcpy(ctx->hstage, &hl, sizeof(hl));
    QK_HIP(hipMemcpyAsync(ctx->lanes.p, ctx->hstage, sizeof(hl), hipMemcpyHostToDevice, ctx->stream));
    PathArgs<Real> a{};
    a.lanes = (const LaneC<Real> *)ctx->lanes.p;
    a.pays = nullptr;
    a.n_lanes = 1;
    a.n_pay = 0;
    a.n_steps = c.n_steps;
    a.n_paths = c.n_paths;
    a.n_tiles = (c.n_paths + TILE - 1) / TILE;
    a.path0 = path0;
    a.key0 = (uint32_t)c.seed;
    a.key1 = (uint32_t)(c.seed >> 32);
    a.store = (Real *)ctx->store.p;
    QK_HIP(hipEventRecord(ctx->ev[0], ctx->stream));
    QK_HIP(launch_paths<Real>(a, c.model, c.scheme, true, true, 1, ctx->stream));
    QK_HIP(hipEventRecord(ctx->ev[1], ctx->stream));
    ctx->store_paths = c.n_paths;
    ctx->store_steps = c.n_steps;
    ctx->store_prec = c.precision;
    return QK_OK;
}

template <typename Real>
static qk_status run_american(qk_ctx *ctx, const qk_contract &c, int right, double K, qk_stats *out,
                              double *coeffs_out, int n_basis = 4, bool weighted = false) {
    // The merge's discount (1: no e^{-rT}, OptionPricer.hpp:211-216) and price
    // scale, in one copy from the pinned staging buffer ahead of the paths.
    qk_status st;
    const int N_ = c.n_steps;
    // Staging for the whole call, reserved before any copy from it is queued:
    // lane constants and discount/scale going in, results and coefficients
    // coming back (after the kernels, so they reuse the slots).
    QK_HIP(hipStreamSynchronize(ctx->stream));
    if ((st = host_stage(ctx, std::max<size_t>(STAGE_AUX + 16, 256 + (size_t)N_ * n_basis * 8))) != QK_OK) return st;
    if ((st = ctx->disc.reserve(ctx, 16)) != QK_OK) return st;
    double *aux = (double *)((char *)ctx->hstage + STAGE_AUX);
    aux[0] = 1.0;
    aux[1] = c.price_scale;
    QK_HIP(hipMemcpyAsync(ctx->disc.p, aux, 16, hipMemcpyHostToDevice, ctx->stream));
    if ((st = generate_store<Real>(ctx, c)) != QK_OK) return st;
    const int64_t n = c.n_paths;
    const int N = c.n_steps;
    const int64_t n_tiles = (n + TILE - 1) / TILE;
    int64_t chunk = 0;
    const int nb = pump_blocks(n, lsm_group<Real>(), PUMP_MAX_BLOCKS, &chunk);
    if ((st = ctx->cf.reserve(ctx, (size_t)n * sizeof(Real))) != QK_OK) return st;
    if ((st = ctx->coeffs.reserve(ctx, (size_t)N * n_basis * 8)) != QK_OK) return st;
    if ((st = ctx->bpart.reserve(ctx, (size_t)2 * nb * LSM_STRIDE * 8)) != QK_OK) return st;
    int64_t gchunk = 0;
    const int gchunks = lsm_gram_chunks(n, &gchunk);
    const size_t gwin = (size_t)std::max(1, std::min(N - 1, lsm_gram_window(gchunks)));
    if ((st = ctx->gpart.reserve(ctx, gwin * gchunks * LsmGram<5>::STRIDE * 8)) != QK_OK) return st;
    if ((st = ctx->factor.reserve(ctx, (size_t)N * LSM_FACTOR_STRIDE * 8)) != QK_OK) return st;
    if ((st = ctx->psum.reserve(ctx, n_tiles * 8)) != QK_OK) return st;
    if ((st = ctx->pm2.reserve(ctx, n_tiles * 8)) != QK_OK) return st;
    if ((st = ctx->pitm.reserve(ctx, n_tiles * 4)) != QK_OK) return st;
    if ((st = ctx->stats.reserve(ctx, sizeof(qk_stats))) != QK_OK) return st;
    if ((st = ctx->mscratch.reserve(ctx, merge_chunks(n_tiles) * MERGE_MOM_BYTES)) != QK_OK) return st;
    hipStream_t s = ctx->stream;
    // Gram matrices of every date (they depend only on the stored spots),
    // factorised before the backward sweep.
    LsmGramArgs<Real> ga{};
    ga.store = (const Real *)ctx->store.p;
    ga.gpart = (double *)ctx->gpart.p;
    ga.factor = (double *)ctx->factor.p;
    ga.n_paths = n;
    ga.gchunk = gchunk;
    ga.n_steps = N;
    ga.right = right;
    ga.gchunks = gchunks;
    ga.K = (Real)K;
    QK_HIP(launch_lsm_gram<Real>(ga, n_basis, weighted, s));
    PumpArgs<Real> a{};
    PumpReq<Real> &q = a.r[0];
    q.store = (const Real *)ctx->store.p;
    q.cf = (Real *)ctx->cf.p;
    q.bpart = (double *)ctx->bpart.p;
    q.factor = (const double *)ctx->factor.p;
    q.coeffs = (double *)ctx->coeffs.p;
    q.psum = (double *)ctx->psum.p;
    q.pm2 = (double *)ctx->pm2.p;
    q.pitm = (uint32_t *)ctx->pitm.p;
    q.n_paths = n;
    q.chunk = chunk;
    q.n_steps = N;
    q.right = right;
    q.nb = nb;
    q.K = (Real)K;
    q.disc = (Real)std::exp(-c.r * (c.T / N));           // ForwardBackwardMC.hpp:209
    q.itm_eps = (Real)(c.precision == QK_FP32 ? 1e-6 : 1e-12);
    a.n_req = 1;
    a.block_start[0] = 0;
    a.block_start[1] = nb;
    // One launch per date, or on request all dates in one cooperative launch
    // (when the cash flows of a block's chunks fit its LDS and every block can
    // be resident).  Both give bit-identical results; the per-date launches
    // measure faster on MI355X, so they are the default.
    const int coop_blocks = nb;
    const size_t coop_lds = (size_t)chunk * sizeof(Real);
    bool coop = false;
    if (ctx->lsm_mode == QK_LSM_COOPERATIVE) {
        int max_lds = 0, n_cu = 0, coop_ok = 0;
        (void)hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device);
        (void)hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, ctx->device);
        (void)hipDeviceGetAttribute(&coop_ok, hipDeviceAttributeCooperativeLaunch, ctx->device);
        coop = coop_ok && coop_lds + 16384 <= (size_t)max_lds &&
               (int64_t)lsm_coop_occupancy<Real>(n_basis, weighted, coop_lds) * n_cu >= coop_blocks;
    }
    if (coop) {
        LsmCoopArgs<Real> ca{};
        ca.store = q.store;
        ca.cf = q.cf;
        ca.bpart = q.bpart;
        ca.factor = q.factor;
        ca.coeffs = q.coeffs;
        ca.psum = q.psum;
        ca.pm2 = q.pm2;
        ca.pitm = q.pitm;
        ca.sync = (unsigned int *)ctx->counter.p;
        ca.n_paths = n;
        ca.chunk = chunk;
        ca.n_steps = N;
        ca.right = right;
        ca.nb = nb;
        ca.K = q.K;
        ca.disc = q.disc;
        ca.itm_eps = q.itm_eps;
        QK_HIP(hipMemsetAsync(ctx->counter.p, 0, 2048, s));
        QK_HIP(launch_lsm_coop<Real>(ca, n_basis, weighted, coop_blocks, coop_lds, s));
    } else {
        for (int t = N - 1; t >= 0; --t) {               // ForwardBackwardMC.hpp:219-242
            q.t = t;
            QK_HIP(launch_pump<Real>(a, n_basis, weighted, s));
        }
    }
    QK_HIP(hipEventRecord(ctx->ev[3], s));
    QK_HIP(launch_merge((const double *)ctx->psum.p, (const double *)ctx->pm2.p, (const uint32_t *)ctx->pitm.p,
                        n_tiles, n, 1, 1, (const double *)ctx->disc.p, (const double *)ctx->disc.p + 1,
                        ctx->mscratch.p, (qk_stats *)ctx->stats.p, s));
    QK_HIP(hipEventRecord(ctx->ev[2], s));
    // results through the staging buffer: stats at 0, the error word at 128,
    // the coefficients at 256
    char *hs = (char *)ctx->hstage;
    QK_HIP(hipMemcpyAsync(hs, ctx->stats.p, sizeof(qk_stats), hipMemcpyDeviceToHost, s));
    if (coeffs_out)
        QK_HIP(hipMemcpyAsync(hs + 256, ctx->coeffs.p, (size_t)N * n_basis * 8, hipMemcpyDeviceToHost, s));
    unsigned int sync_err = 0;
    if (coop) QK_HIP(hipMemcpyAsync(hs + 128, (char *)ctx->counter.p + 4 * LSM_COOP_SYNC_ERR_WORD, 4,
                                    hipMemcpyDeviceToHost, s));
    QK_HIP(hipStreamSynchronize(s));
    std::memcpy(out, hs, sizeof(qk_stats));
    if (coeffs_out) std::memcpy(coeffs_out, hs + 256, (size_t)N * n_basis * 8);
    if (coop) std::memcpy(&sync_err, hs + 128, 4);
    if (sync_err) return fail(ctx, QK_EDEVICE, "Longstaff-Schwartz grid barrier timed out");
    // Immediate exercise check (LSMAmerican.hpp:409-418): if today's intrinsic
    // beats the mean discounted cash flow, every path exercises at t0.
    const Real S0 = (Real)c.S0, Kr = (Real)K;
    const Real today = right == QK_CALL ? (S0 > Kr ? S0 - Kr : Real(0)) : (Kr > S0 ? Kr - S0 : Real(0));
    if ((double)today > out->mean) {
        const double eps = c.precision == QK_FP32 ? 1e-6 : 1e-12;
        out->mean = (double)today;
        out->m2 = 0.0;
        out->price = (double)today * c.price_scale;
        out->se = 0.0;
        out->itm = (double)today > eps ? n : 0;
        out->prob_itm = (double)out->itm / (double)n;
    }
    return QK_OK;
}

}  // namespace qk

extern "C" qk_status qk_price_american(qk_ctx *ctx, const qk_contract *c, int32_t right, double K, qk_stats *out,
                                       double *coeffs_out) {
    if (!ctx || !out) return fail(ctx, QK_EINVAL, "null argument");
    qk_status st = validate_contract(ctx, c);
    if (st != QK_OK) return st;
    if (right != QK_CALL && right != QK_PUT) return fail(ctx, QK_EINVAL, "right out of range");
    if (!(K > 0.0)) return fail(ctx, QK_EINVAL, "K must be positive");
    if ((st = check_budget(ctx, footprint(*c, 1, 1, true))) != QK_OK) return st;
    QK_HIP(hipSetDevice(ctx->device));
    st = c->precision == QK_FP32 ? run_american<float>(ctx, *c, right, K, out, coeffs_out)
                                 : run_american<double>(ctx, *c, right, K, out, coeffs_out);
    if (st != QK_OK) return st;
    return finish_timing(ctx, c->n_paths * (int64_t)c->n_steps, true);
}

// ---------------------------------------------------------------------------
// Payoff distribution (OptionPricer.hpp:340-360): exact order statistics of
// the per-path values at the reference's ranks, selected on the device.
// ---------------------------------------------------------------------------
namespace qk {

constexpr int N_QUANTILES = 15;
// Rank fractions of p1 ... p99 in qk_distribution order (OptionPricer.hpp:345-357).
constexpr double QUANTILE_Q[N_QUANTILES - 2] = {0.01, 0.05, 0.10, 0.20, 0.30, 0.40, 0.50,
                                                 0.60, 0.70, 0.80, 0.90, 0.95, 0.99};

template <typename Real>
static Real key_to_value(uint64_t k) {
    if constexpr (sizeof(Real) == 4) {
        uint32_t b = (uint32_t)k;
        b = (b & 0x80000000u) ? (b & 0x7fffffffu) : ~b;
        float f;
        std::memcpy(&f, &b, 4);
        return f;
    } else {
        uint64_t b = (k & 0x8000000000000000ull) ? (k & 0x7fffffffffffffffull) : ~k;
        double d;
        std::memcpy(&d, &b, 8);
        return d;
    }
}

// Values of the given ranks (0-based, ascending order) of vals[0..n).
template <typename Real>
static qk_status select_ranks(qk_ctx *ctx, const Real *vals, int64_t n, const int64_t *ranks, int n_ranks,
                              double *out) {
    constexpr int BITS = 8 * (int)sizeof(Real);
    qk_status st;
    if ((st = ctx->hist.reserve(ctx, SELECT_MAX_GROUPS * 256 * 4)) != QK_OK) return st;
    QK_HIP(hipStreamSynchronize(ctx->stream));
    if ((st = host_stage(ctx, SELECT_MAX_GROUPS * 256 * 4)) != QK_OK) return st;
    std::vector<uint64_t> prefix(n_ranks, 0);
    std::vector<int64_t> rest(ranks, ranks + n_ranks);
    std::vector<uint32_t> h(SELECT_MAX_GROUPS * 256);
    int n_cu = 256;
    (void)hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, ctx->device);
    const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, (int64_t)n_cu * 8));
    hipStream_t s = ctx->stream;
    for (int shift = BITS - 8; shift >= 0; shift -= 8) {
        SelectArgs a{};
        std::vector<int> group(n_ranks);
        for (int r = 0; r < n_ranks; ++r) {
            int g = 0;
            while (g < a.n_groups && a.prefix[g] != prefix[r]) ++g;
            if (g == a.n_groups) a.prefix[a.n_groups++] = prefix[r];
            group[r] = g;
        }
        QK_HIP(hipMemsetAsync(ctx->hist.p, 0, (size_t)a.n_groups * 256 * 4, s));
        a.shift = shift;
        QK_HIP(launch_select_hist<Real>(vals, n, a, (uint32_t *)ctx->hist.p, blocks, s));
        if ((st = copy_out(ctx, h.data(), ctx->hist.p, (size_t)a.n_groups * 256 * 4)) != QK_OK) return st;
        for (int r = 0; r < n_ranks; ++r) {
            const uint32_t *hg = h.data() + group[r] * 256;
            int d = 0;
            while (d < 255 && rest[r] >= (int64_t)hg[d]) rest[r] -= hg[d++];
            prefix[r] = (prefix[r] << 8) | (uint64_t)d;
        }
    }
    for (int r = 0; r < n_ranks; ++r) out[r] = (double)key_to_value<Real>(prefix[r]);
    return QK_OK;
}

template <typename Real>
static qk_status run_distribution(qk_ctx *ctx, const qk_contract &c, const qk_payoff &p, qk_stats *out,
                                  qk_distribution *dist) {
    const int64_t n = c.n_paths;
    const bool american = p.type == QK_AMERICAN;
    qk_status st;
    const Real *vals;
    double factor;
    if (american) {
        if ((st = run_american<Real>(ctx, c, p.right, p.K, out, nullptr)) != QK_OK) return st;
        if ((st = finish_timing(ctx, n * (int64_t)c.n_steps, true)) != QK_OK) return st;
        vals = (const Real *)ctx->cf.p;
        factor = c.price_scale;                               // no e^{-rT}: OptionPricer.hpp:211-216
    } else {
        if ((st = ctx->payv.reserve(ctx, (size_t)n * sizeof(Real))) != QK_OK) return st;
        if ((st = run_lanes<Real>(ctx, &c, 1, &p, 1, (Real *)ctx->payv.p)) != QK_OK) return st;
        if ((st = copy_out(ctx, out, ctx->stats.p, sizeof(qk_stats))) != QK_OK) return st;
        if ((st = finish_timing(ctx, n * (int64_t)c.n_steps, false)) != QK_OK) return st;
        vals = (const Real *)ctx->payv.p;
        factor = std::exp(-c.r * c.T) * (p.type == QK_BINARY_CASH ? 1.0 : c.price_scale);
    }
    const auto t0 = std::chrono::steady_clock::now();
    int64_t ranks[N_QUANTILES];
    const double N = (double)n;
    ranks[0] = 0;
    for (int i = 0; i < N_QUANTILES - 2; ++i) ranks[1 + i] = (int64_t)(N * QUANTILE_Q[i]);
    ranks[N_QUANTILES - 1] = n - 1;
    double v[N_QUANTILES];
    // Immediate exercise at t0 replaces every cash flow (LSMAmerican.hpp:409-418).
    if (american && out->se == 0.0 && out->m2 == 0.0 && n > 1) {
        for (double &x : v) x = out->mean;
    } else if ((st = select_ranks<Real>(ctx, vals, n, ranks, N_QUANTILES, v)) != QK_OK) {
        return st;
    }
    double *d = &dist->p0;
    for (int i = 0; i < N_QUANTILES; ++i) d[i] = v[i] * factor;
    dist->select_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return QK_OK;
}

}  // namespace qk

extern "C" qk_status qk_price_distribution(qk_ctx *ctx, const qk_contract *c, const qk_payoff *p, qk_stats *out,
                                           qk_distribution *dist) {
    if (!ctx || !out || !dist || !p) return fail(ctx, QK_EINVAL, "null argument");
    qk_status st = validate_contract(ctx, c);
    if (st != QK_OK) return st;
    const bool american = p->type == QK_AMERICAN;
    if (american) {
        if (p->right != QK_CALL && p->right != QK_PUT) return fail(ctx, QK_EINVAL, "payoff.right out of range");
        if (!(p->K > 0.0)) return fail(ctx, QK_EINVAL, "payoff.K must be positive");
    } else if ((st = validate_payoff(ctx, p)) != QK_OK) {
        return st;
    }
    const size_t extra = american ? 0 : (size_t)c->n_paths * (c->precision == QK_FP32 ? 4 : 8);
    if ((st = check_budget(ctx, footprint(*c, 1, 1, american) + extra)) != QK_OK) return st;
    QK_HIP(hipSetDevice(ctx->device));
    return c->precision == QK_FP32 ? run_distribution<float>(ctx, *c, *p, out, dist)
                                   : run_distribution<double>(ctx, *c, *p, out, dist);
}

extern "C" qk_status qk_set_lsm_schedule(qk_ctx *ctx, int32_t mode) {
    if (!ctx) return QK_EINVAL;
    if (mode < QK_LSM_AUTO || mode > QK_LSM_COOPERATIVE) return fail(ctx, QK_EINVAL, "lsm schedule out of range");
    ctx->lsm_mode = mode;
    return QK_OK;
}

extern "C" qk_status qk_price_american_ex(qk_ctx *ctx, const qk_contract *c, int32_t right, double K,
                                          const qk_lsm_options *opt, qk_stats *out, double *coeffs_out) {
    if (!ctx || !out || !opt) return fail(ctx, QK_EINVAL, "null argument");
    if (opt->basis < QK_LAGUERRE_P02 || opt->basis > QK_LAGUERRE_P04)
        return fail(ctx, QK_EINVAL, "lsm basis must be LaguerreP02, P03 or P04");
    qk_status st = validate_contract(ctx, c);
    if (st != QK_OK) return st;
    if (right != QK_CALL && right != QK_PUT) return fail(ctx, QK_EINVAL, "right out of range");
    if (!(K > 0.0)) return fail(ctx, QK_EINVAL, "K must be positive");
    if ((st = check_budget(ctx, footprint(*c, 1, 1, true))) != QK_OK) return st;
    QK_HIP(hipSetDevice(ctx->device));
    const int nb = 3 + opt->basis;
    const bool w = opt->weighted != 0;
    st = c->precision == QK_FP32 ? run_american<float>(ctx, *c, right, K, out, coeffs_out, nb, w)
                                 : run_american<double>(ctx, *c, right, K, out, coeffs_out, nb, w);
    if (st != QK_OK) return st;
    return finish_timing(ctx, c->n_paths * (int64_t)c->n_steps, true);
}

extern "C" qk_status qk_simulate_paths(qk_ctx *ctx, const qk_contract *c, int64_t path0, double *S_out) {
    if (!ctx || !S_out) return fail(ctx, QK_EINVAL, "null argument");
    if (path0 < 0) return fail(ctx, QK_EINVAL, "path0 must be >= 0");
    qk_status st = validate_contract(ctx, c);
    if (st != QK_OK) return st;
    if ((st = check_budget(ctx, footprint(*c, 1, 1, true))) != QK_OK) return st;
    QK_HIP(hipSetDevice(ctx->device));
    const size_t n = (size_t)c->n_paths * c->n_steps;
    if (c->precision == QK_FP64) {
        if ((st = generate_store<double>(ctx, *c, path0)) != QK_OK) return st;
        QK_HIP(hipMemcpyAsync(S_out, ctx->store.p, n * 8, hipMemcpyDeviceToHost, ctx->stream));
        QK_HIP(hipStreamSynchronize(ctx->stream));
    } else {
        if ((st = generate_store<float>(ctx, *c, path0)) != QK_OK) return st;
        std::vector<float> tmp(n);
        QK_HIP(hipMemcpyAsync(tmp.data(), ctx->store.p, n * 4, hipMemcpyDeviceToHost, ctx->stream));
        QK_HIP(hipStreamSynchronize(ctx->stream));
        for (size_t i = 0; i < n; ++i) S_out[i] = tmp[i];
    }
    QK_HIP(hipEventRecord(ctx->ev[2], ctx->stream));
    return finish_timing(ctx, c->n_paths * (int64_t)c->n_steps, false);
}
