// qk_capi_internal.hpp -- the qk_ctx definition and helpers shared by the C ABI
// translation units (pricing entry points and the request server).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "qk_types.hpp"

namespace qk {

// A device buffer that only grows; freed with its context.
struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    qk_status reserve(qk_ctx *ctx, size_t bytes);
    void release();
};

qk_status fail(qk_ctx *ctx, qk_status s, const std::string &msg);
qk_status hip_fail(qk_ctx *ctx, hipError_t e, const char *what);
qk_status validate_contract(qk_ctx *ctx, const qk_contract *c);
qk_status validate_payoff(qk_ctx *ctx, const qk_payoff *p);
bool path_dependent(int type);
template <typename Real>
LaneC<Real> make_lane(const qk_contract &c);
size_t footprint(const qk_contract &c, int n_lanes, int n_pay, bool american);
qk_status check_budget(qk_ctx *ctx, size_t bytes);
void noalloc_init(int device);
// Host bytes of the constants block of qk_price_payoffs_noalloc (one lane,
// at most MAX_PAYOFFS payoffs), built on the stack.
constexpr size_t NOALLOC_CONST_BYTES = 2048;

}  // namespace qk

struct qk_ctx {
    int device = 0;
    size_t budget = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {};
    std::string err;
    qk_timing timing{};
    qk::DevBuf lanes, pays, psum, pm2, pitm, disc, scale, stats, greeks, store, cf, coeffs, bpart, counter, sums, mscratch;
    qk::DevBuf gpart, factor; // Longstaff-Schwartz Gram pre-pass partials and per-date factor records
    qk::DevBuf payv, hist;   // per-path payoffs and selection histograms of qk_price_distribution
    qk::DevBuf tpose;        // [path][step] window of the path store (qk_copy_stored_paths)
    // Shape of the complete path set in `store` (store_paths = 0: none), left
    // by qk_price_store_paths, qk_price_american* and qk_simulate_paths.
    int64_t store_paths = 0;
    int32_t store_steps = 0, store_prec = QK_FP64;
    std::vector<qk_schedule_entry> schedule;   // lanes per thread by (precision, scheme, n_lanes)
    int32_t lsm_mode = QK_LSM_AUTO;             // backward-pass schedule of qk_price_american*
    // Pinned host staging of a lane call's constants (one host-to-device copy)
    // and results; every call that uses it synchronises before it returns.
    void *hstage = nullptr;
    size_t hstage_cap = 0;
    ~qk_ctx();
};
