/*
 * qk.h -- C ABI of the MI355X-native QuantKos Monte Carlo engine (libquantkos.so).
 *
 * Plain C types only: no C++ or torch types cross this boundary, no exception
 * escapes it.  Every call returns a qk_status; qk_last_error() gives the text.
 * The caller owns every input and output struct; a qk_ctx owns all device
 * memory (path stores, partial sums, request queues) and one HIP stream.
 * Calls on one ctx are not thread-safe, except the qk_server_* family, whose
 * submit/poll/wait are internally synchronised.
 *
 * The reference (QuantKos) has no C ABI: its operator API is a set of C++
 * template policies.  Each entry point below names the reference interface it
 * replaces; the C++ drop-in types in quantkos_amd/host/ map onto these calls.
 */
#ifndef QK_H
#define QK_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    QK_OK = 0,
    QK_EINVAL = -1,     /* invalid argument (message in qk_last_error)            */
    QK_ENOMEM = -2,     /* request exceeds the ctx memory budget or device memory  */
    QK_EDEVICE = -3,    /* HIP runtime error                                       */
    QK_EAGAIN = -4,     /* server queue full; retry later                          */
    QK_ENOTFOUND = -5,  /* unknown ticket                                          */
    QK_ENOTREADY = -6   /* ticket not finished yet (qk_server_poll)                */
} qk_status;

/* Enumerations in the order of the reference's ConfigFileEnums.hpp:24-58. */
enum { QK_MODEL_HESTON = 0, QK_MODEL_BATES = 1 };
enum { QK_SCHEME_EULER = 0, QK_SCHEME_MILSTEIN = 1, QK_SCHEME_QE = 2 };
enum {
    QK_EUROPEAN = 0, QK_ASIAN = 1, QK_BARRIER_UP_OUT = 2, QK_BARRIER_DOWN_OUT = 3,
    QK_BARRIER_UP_IN = 4, QK_BARRIER_DOWN_IN = 5, QK_LOOKBACK_FLOATING = 6,
    QK_LOOKBACK_FIXED = 7, QK_BINARY_CASH = 8, QK_BINARY_ASSET = 9, QK_AMERICAN = 10
};
enum { QK_CALL = 0, QK_PUT = 1 };
enum { QK_FP32 = 0, QK_FP64 = 1 };

typedef struct qk_ctx qk_ctx;

/* One model + grid + path set.  Monetary inputs are taken as given: the host
 * layer applies the reference's Normalize_prices rule (S0 -> 1, K,B /= S0;
 * Config.hpp:710-723) and passes the scale in price_scale, by which every
 * price and SE is multiplied (except cash-or-nothing binaries, scale 1;
 * OptionPricer.hpp:323-332).  dt = T / n_steps, 1 <= n_steps <= 2^30,
 * 1 <= n_paths <= 2^40. */
typedef struct {
    int32_t model, scheme, precision, n_steps;
    int64_t n_paths;
    uint64_t seed;
    double S0, v0, r, q, T;
    double kappa, theta, sigma, rho;
    double lambda_j, mu_j, sigma_j;
    double price_scale;
} qk_contract;

/* Payoff spec (PayoffTracker<Type,Right>, Payoff.hpp:84-243). */
typedef struct { int32_t type, right; double K, B; } qk_payoff;

/* Pricing result (MCOpPrices, OptionPricer.hpp:50-70) plus the raw moments. */
typedef struct {
    double price, se, prob_itm, mean, m2;
    int64_t n, itm;
} qk_stats;

/* Greek stencil switches and bump sizes (MCConfig, Config.hpp:431-447). */
typedef struct {
    int32_t delta_gamma, vega_vomma, vanna, rho, theta;
    int32_t theta_central;   /* 0: reference one-sided (T-h); 1: central (T-h, T+h) */
    double rel_S, abs_vol, abs_r, abs_T;
} qk_bumps;

/* MCGreeks (MCEngine.hpp:93-101) plus the base price and the lane count. */
typedef struct {
    double delta, gamma, vega, vomma, vanna, rho, theta;
    double base_price, base_se;
    int32_t n_lanes;
} qk_greeks;

/* Device-time breakdown of the last call on a ctx (HIP events on its stream). */
typedef struct {
    double path_kernel_ms;   /* fused path-generation kernel(s)             */
    double reduce_ms;        /* tile-partial merge + finalisation           */
    double lsm_backward_ms;  /* Longstaff-Schwartz backward sweep (American) */
    double total_ms;         /* first to last event of the call             */
    int64_t path_steps;      /* path-steps simulated by the call            */
} qk_timing;

/* ---- context ---- */
/* Replaces Kokkos::initialize + the memory managers' budget (PathsMCBatchMem.hpp:49-421):
 * binds a device and a byte budget that every call's device footprint must fit. */
qk_status qk_ctx_create(int device, size_t mem_budget_bytes, qk_ctx **out);
qk_status qk_ctx_destroy(qk_ctx *ctx);
const char *qk_strerror(qk_status s);
const char *qk_last_error(const qk_ctx *ctx);
qk_status qk_synchronize(qk_ctx *ctx);
qk_status qk_last_timing(const qk_ctx *ctx, qk_timing *out);
/* Bytes of device memory a call with this contract / payload would need. */
size_t qk_footprint_bytes(const qk_contract *c, int32_t n_lanes, int32_t n_payoffs, int32_t american);

/* ---- launch schedule (the autotuner's knob for the lane kernels) ----
 * How many lanes of a qk_price_lanes / qk_price_greeks launch one thread
 * carries, sharing each step's random numbers: 1, 2 or 4, or 8 in FP32.  Entries are keyed
 * by (precision, scheme, n_lanes); a call without a matching entry uses 4 when
 * n_lanes >= 4, else 1.  The schedule changes speed only: prices agree to
 * rounding whatever it is.  Replaces the launch-policy choice the reference
 * leaves to Kokkos' RangePolicy (MSolver.hpp:195-232); quantkos_amd/autotune.py
 * measures the candidates and writes the table. */
typedef struct { int32_t precision, scheme, n_lanes, lanes_per_thread; } qk_schedule_entry;
/* Replaces the ctx's table with entries[0..n) (n = 0 clears it; at most 256). */
qk_status qk_set_schedule(qk_ctx *ctx, const qk_schedule_entry *entries, int32_t n);
/* The lanes per thread a call with this key would use. */
int32_t qk_lanes_per_thread(const qk_ctx *ctx, int32_t precision, int32_t scheme, int32_t n_lanes);

/* Schedule of the Longstaff-Schwartz backward pass of qk_price_american*:
 * QK_LSM_PER_DATE launches one kernel per exercise date; QK_LSM_COOPERATIVE
 * runs every date in one cooperative launch (one device-wide synchronisation
 * per date, cash flows held in LDS) when the request's cash flows fit the
 * blocks' LDS and falls back to per-date launches otherwise; QK_LSM_AUTO is
 * the faster of the two as measured on MI355X (per-date).  The schedules give
 * bit-identical results.  No counterpart in the reference, whose backward
 * pass is a host loop over dates (ForwardBackwardMC.hpp:219-242). */
enum { QK_LSM_AUTO = 0, QK_LSM_PER_DATE = 1, QK_LSM_COOPERATIVE = 2 };
qk_status qk_set_lsm_schedule(qk_ctx *ctx, int32_t mode);

/* ---- pricing ---- */
/* Plain single-contract launch.  Replaces MSolver<M,S,T,R>::execute_batch
 * (MSolver.hpp:195-232) + OptionPricer::accumulate_batch_payoffs /
 * evaluate_option_price (OptionPricer.hpp:123-249) for one OptType. */
qk_status qk_price(qk_ctx *ctx, const qk_contract *c, const qk_payoff *payoff, qk_stats *out);

/* Several payoffs from one path set in one launch (the reference needs one run
 * per OptType, Distpatcher.hpp:140-178).  out[n_payoffs]. */
qk_status qk_price_payoffs(qk_ctx *ctx, const qk_contract *c, const qk_payoff *payoffs,
                           int32_t n_payoffs, qk_stats *out);

/* qk_price_payoffs without any allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops that forbid hipMalloc).  Every device buffer of the
 * call comes from a fixed array of QK_NOALLOC_SCRATCH_BYTES per device,
 * reserved when the library loads; about 20 bytes per 64 paths per payoff, so
 * up to ~16M paths with 3 payoffs.  Larger calls fail with QK_ENOMEM and
 * allocate nothing.  The array is shared by every call on a device: calls
 * that overlap on one device race.  Results equal qk_price_payoffs'. */
#define QK_NOALLOC_SCRATCH_BYTES ((size_t)16 << 20)
qk_status qk_price_payoffs_noalloc(qk_ctx *ctx, const qk_contract *c, const qk_payoff *payoffs,
                                   int32_t n_payoffs, qk_stats *out);

/* ---- batches: path windows priced separately, merged on the device ----
 * The reference prices a path set in batches of MC.batch_size paths
 * (PathsMCBatchMem.hpp:293-338, run_forward_all_mc_batches in
 * MCEngine.hpp:489-520: MSolver::execute_batch per batch, then
 * OptionPricer::accumulate_batch_payoffs, OptionPricer.hpp:123-193).  Here a
 * batch is a window of the path set: qk_price_window returns the window's
 * merge chunks (the fixed-order merge's first stage, one per 65,536 paths of
 * the window) and qk_merge_windows runs the second stage over the chunks of
 * every window.  When every window starts at a multiple of
 * QK_WINDOW_CHUNK_PATHS, the merged result equals qk_price_payoffs over the
 * whole path set bit for bit; other 64-aligned windows merge exactly, in a
 * different summation order. */
#define QK_WINDOW_CHUNK_PATHS ((int64_t)65536)
/* Merge state of a run of paths: count, mean and sum of squared deviations of
 * the undiscounted payoff (before the price scale), in-the-money count. */
typedef struct { double n, mean, m2; uint64_t itm; } qk_moments;

/* Prices paths [path0, path0 + c->n_paths) of the path set of (c->seed, the
 * contract) -- the paths a single launch would number path0 .. -- and writes
 * chunks[j * nc + k] for payoff j and chunk k, nc = ceil(c->n_paths / 65536).
 * path0 must be a multiple of 64. */
qk_status qk_price_window(qk_ctx *ctx, const qk_contract *c, int64_t path0, const qk_payoff *payoffs,
                          int32_t n_payoffs, qk_moments *chunks);

/* Merges chunks[j * n_chunks + k] (the windows' chunks in path order for each
 * payoff j) and finalises price, SE and P(ITM) with the discount e^{-rT} and
 * the price scale of c: out[n_payoffs]. */
qk_status qk_merge_windows(qk_ctx *ctx, const qk_contract *c, const qk_payoff *payoffs, int32_t n_payoffs,
                           const qk_moments *chunks, int64_t n_chunks, qk_stats *out);

/* Lanes = contracts that share model, scheme, precision, n_steps, n_paths and
 * seed, and therefore the random stream (common random numbers); they may
 * differ in S0, v0, r, q, T and the Heston/jump parameters.  One launch
 * prices every lane x payoff.  out[n_lanes * n_payoffs], lane-major. */
qk_status qk_price_lanes(qk_ctx *ctx, const qk_contract *lanes, int32_t n_lanes,
                         const qk_payoff *payoffs, int32_t n_payoffs, qk_stats *out);

/* Finite-difference Greeks with the reference's stencils (MCEngine.hpp:576-735),
 * every bump a lane of one launch; differences taken in FP64 on the device.
 * out[n_payoffs]. */
qk_status qk_price_greeks(qk_ctx *ctx, const qk_contract *c, const qk_payoff *payoffs, int32_t n_payoffs,
                          const qk_bumps *bumps, qk_greeks *out);

/* American option by Longstaff-Schwartz with the unweighted Laguerre basis of
 * degree 3 on in-the-money paths, exercise at every grid date.  Replaces
 * ForwardBackwardMCRunner (ForwardBackwardMC.hpp:68-256) + LSMEngine
 * (LSMAmerican.hpp:217-614); the path store stays in HBM.  coeffs_out, if not
 * NULL, receives the regression coefficients [n_steps][4] (date t row = the
 * regression at date index t; row n_steps-1 is zero). */
qk_status qk_price_american(qk_ctx *ctx, const qk_contract *c, int32_t right, double K,
                            qk_stats *out, double *coeffs_out);

/* Regression basis of Longstaff-Schwartz.  The reference fixes LaguerreP03 at
 * compile time (LSMAmerican.hpp:217-233, LSRegressionBasis in
 * ConfigFileEnums.hpp:107-112, same order); weighted multiplies every basis
 * function by exp(-X/2), the weighted Laguerre functions of Longstaff and
 * Schwartz (2001), which the reference does not offer. */
enum { QK_LAGUERRE_P02 = 0, QK_LAGUERRE_P03 = 1, QK_LAGUERRE_P04 = 2 };
typedef struct { int32_t basis, weighted; } qk_lsm_options;

/* qk_price_american with another regression basis; coeffs_out, if not NULL,
 * receives [n_steps][n_basis] (n_basis = 3, 4, 5 for P02, P03, P04). */
qk_status qk_price_american_ex(qk_ctx *ctx, const qk_contract *c, int32_t right, double K,
                               const qk_lsm_options *opt, qk_stats *out, double *coeffs_out);

/* Discounted payoff distribution (OptionPricer::analyze_risk_neutral_discounted_
 * payoff_distribution, OptionPricer.hpp:340-360, enabled by the YAML key
 * MC.analyze_risk_neutral_payoff_distribution): the per-path payoffs sorted
 * ascending, taken at 0-based ranks floor(n*q) for q = 0.01, 0.05, 0.10, ...,
 * 0.90, 0.95, 0.99 (FP64 product, truncated), plus the minimum and maximum,
 * each times e^{-rT} (none for American) and the price scale.  The values are
 * exact order statistics, selected on the device. */
typedef struct {
    double p0, p1, p5, p10, p20, p30, p40, median, p60, p70, p80, p90, p95, p99, p100;
    double select_s;   /* wall time of the selection (the reference's sorting time) */
} qk_distribution;

/* qk_price (or, for payoff->type QK_AMERICAN, qk_price_american) plus the
 * payoff distribution of the same paths. */
qk_status qk_price_distribution(qk_ctx *ctx, const qk_contract *c, const qk_payoff *payoff, qk_stats *out,
                                qk_distribution *dist);

/* Simulated spot paths [path0, path0 + c->n_paths) of the path set, copied to
 * host memory laid out [step][path] (S after step t+1 at row t), as the
 * reference's StorePaths kernel writes them (MSolver.hpp:117-119).  Path p
 * draws the same random numbers whatever path0 and n_paths are. */
qk_status qk_simulate_paths(qk_ctx *ctx, const qk_contract *c, int64_t path0, double *S_out);

/* qk_price that also keeps its paths: the path kernel writes S after every
 * step into the ctx's path store [step][path] in the run's precision, as the
 * reference's StorePaths kernel does when an output file is named
 * (MSolver.hpp:117-119, Config.hpp:622, :750-758).  The price equals
 * qk_price's to rounding (the stored form multiplies S step by step). */
qk_status qk_price_store_paths(qk_ctx *ctx, const qk_contract *c, const qk_payoff *payoff, qk_stats *out);

/* Copies paths [path0, path0 + n) of the path set the last qk_price_store_paths,
 * qk_price_american* or qk_simulate_paths call left on the ctx (path 0 = its
 * first path) into host memory laid out [path][step] -- the ".paths" payload
 * (IOBinary.hpp:139-177) -- with elements of the run's precision (float for
 * QK_FP32, double for QK_FP64): n * n_steps elements.  The transpose runs on
 * the device (BinWriter's host loop, IOBinary.hpp:155-170). */
qk_status qk_copy_stored_paths(qk_ctx *ctx, int64_t path0, int64_t n, void *out);

/* ---- semi-analytic European prices (host only; no device work) ----
 * The reference's closed-form oracles (include/core/analytical:
 * HestonExact.hpp:174-195, BatesExact.hpp:130-145, BlackScholesExact.hpp:46-65)
 * as a product call: the European call or put of the contract's model
 * (Heston, or Bates when model is QK_MODEL_BATES) at strike K, on the same
 * normalised inputs as the Monte Carlo calls, times price_scale.  Heston and
 * Bates by characteristic-function inversion (Albrecher et al.'s form),
 * 64-point Gauss-Legendre panels; quad NULL: panels of 50 until the
 * integrand has decayed (at most phi = 8000); {100, 100} is the reference's
 * Heston rule, {8000, 50} its Bates rule.  sigma = 0: Black-Scholes at the
 * integrated variance (v0 T when kappa = 0), with Merton's jump series for
 * Bates. */
typedef struct { double phi_max, panel; } qk_cf_quadrature;
qk_status qk_analytic_european(const qk_contract *c, int32_t right, double K, const qk_cf_quadrature *quad,
                               double *price);

/* ---- continuous-batching request server (a stream of pricing requests) ----
 * The reference prices one configuration per process run (app/main.cpp:51-110);
 * these entry points replace a loop of MCDispatcher::launch_montecarlo calls
 * (Distpatcher.hpp:78-83) by a resident server.  Requests are admitted while
 * their device footprint (path store for American options, tile moments)
 * fits the server's arena; admitted requests are packed into batch launches
 * that run while later requests arrive, up to max_inflight batches at once.
 * A request's result does not depend on the requests it was batched with. */
typedef struct {
    size_t mem_budget_bytes;     /* device arena for admitted requests (0: 60% of free memory) */
    int32_t max_inflight;        /* batch launches in flight, one stream each (0: 2)           */
    int32_t max_lsm_streams;     /* American requests in the backward sweep at once (0 or >96: 96, four chains of 24) */
    int64_t item_path_steps;     /* path-steps per work item (0: 131072)                      */
    int32_t blocks_per_cu;       /* resident 256-thread blocks per CU of a batch launch (0: 4) */
    int32_t batch_items_per_wave; /* a batch takes requests until its work reaches this many work
                                     items per resident wavefront (0: 32); a larger request still
                                     forms a batch of its own */
    int32_t overlap_min_backlog; /* a batch is launched beside one already in flight only while
                                    the pending requests hold at least this many batches' work
                                    (the work cap above); 0: whenever a stream is free */
    int64_t express_max_path_steps; /* express lane: while its stream is free, pending requests
                                       of at most this many path-steps (an American request's path
                                       store generation) are batched ahead of arrival order (at
                                       most 4x this in all) and launched beside the running batch
                                       (0: 2^27; < 0: no lane) */
} qk_server_config;

typedef struct {
    qk_stats stats;
    double latency_s;            /* submit -> result ready on the host   */
    double queue_s;              /* submit -> admitted into a batch      */
    qk_status status;
} qk_result;

typedef struct {
    int64_t n_submitted, n_completed, n_batches, path_steps;
    int64_t arena_bytes, arena_peak_bytes;
    double busy_s;               /* wall time with at least one batch or sweep in flight */
} qk_server_stats;

typedef struct qk_server qk_server;

qk_status qk_server_create(qk_ctx *ctx, const qk_server_config *cfg, qk_server **out);
qk_status qk_server_destroy(qk_server *srv);
/* payoff->type QK_AMERICAN routes the request through Longstaff-Schwartz. */
qk_status qk_server_submit(qk_server *srv, const qk_contract *c, const qk_payoff *payoff, uint64_t *ticket);
/* qk_server_submit with the regression basis of an American request (lsm NULL:
 * LaguerreP03 unweighted, the reference's basis; ignored for other payoffs). */
qk_status qk_server_submit_ex(qk_server *srv, const qk_contract *c, const qk_payoff *payoff,
                              const qk_lsm_options *lsm, uint64_t *ticket);
qk_status qk_server_poll(qk_server *srv, uint64_t ticket, qk_result *out);    /* QK_ENOTREADY if pending */
qk_status qk_server_wait(qk_server *srv, uint64_t ticket, qk_result *out);
qk_status qk_server_stats_get(qk_server *srv, qk_server_stats *out);
/* Replays a request stream: request i is submitted arrival_s[i] seconds after
 * the call starts (non-decreasing), then every result is awaited. */
qk_status qk_server_replay(qk_server *srv, const qk_contract *cs, const qk_payoff *ps, const double *arrival_s,
                           int32_t n, qk_result *out);

/* ---- host layer: QuantKos YAML configuration files ---- */
typedef struct {
    qk_contract contract;     /* after the Normalize_prices scaling                 */
    qk_payoff payoff;         /* type QK_AMERICAN for American options              */
    qk_bumps bumps;           /* the Greeks requested (all switches 0 if none)      */
    int32_t compute_greeks;
    int32_t write_paths;      /* Output.Name_Paths_Out_File is not empty            */
    qk_lsm_options lsm;       /* MC.LSM_Basis, MC.LSM_Weighted_Basis (American)     */
    int32_t analyze_distribution;   /* MC.analyze_risk_neutral_payoff_distribution  */
} qk_run_spec;

/* Config::Parser::parse (IO/InputParser.hpp:56, src/IO/ConfigParser.cpp:36-294)
 * followed by UInputs::apply_price_scaling (Config.hpp:710-723): reads,
 * validates and scales a configuration file.  No device work.  On failure the
 * reference's error text is copied to err (if not NULL). */
qk_status qk_config_load(const char *yaml_path, qk_run_spec *out, char *err, size_t err_len);

typedef struct {
    qk_stats stats;            /* price, se, prob_itm, n, itm                       */
    qk_greeks greeks;          /* zero unless Greeks were requested                 */
    double t_total_s, t_pricing_s, t_greeks_s;   /* the reference's MCTimer lines  */
    qk_distribution dist;      /* MCOpPrices p0 ... p100 (-999 unless MC.analyze_risk_neutral_payoff_distribution) */
    double analytic_price;     /* European: qk_analytic_european of the same contract; NaN otherwise */
    double analytic_err_se;    /* (price - analytic_price) / se; NaN when not defined */
} qk_run_result;

/* MCDispatcher(Parser::parse(file)).launch_montecarlo() (app/main.cpp:78-93,
 * Distpatcher.hpp:78-83) on the given device; prints the reference's report
 * unless quiet. */
qk_status qk_run_config(const char *yaml_path, int device, int32_t quiet, qk_run_result *out, char *err,
                        size_t err_len);

#ifdef __cplusplus
}
#endif
#endif
