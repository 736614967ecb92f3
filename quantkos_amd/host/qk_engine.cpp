// qk_engine.cpp -- the engine layer over the C ABI (see qk_engine.hpp).
#include "qk_engine.hpp"
#include "qk_solver.hpp"

#include <cstdlib>
#include <iomanip>
#include <iostream>
#include <memory>
#include <mutex>
#include <stdexcept>

#include "qk_paths_io.hpp"

namespace quantkos::Engine {

namespace {
std::mutex g_mu;
int g_device = -1;
qk_ctx *g_ctx = nullptr;

struct CtxGuard {
    ~CtxGuard() {
        if (g_ctx) qk_ctx_destroy(g_ctx);
    }
} g_guard;

void check(qk_status st, const char *what) {
    if (st != QK_OK) {
        throw std::runtime_error(std::string(what) + " failed: " + qk_strerror(st) + ": " +
                                 (g_ctx ? qk_last_error(g_ctx) : ""));
    }
}
}  // namespace

void set_device(int device) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_ctx && g_device != device) {
        qk_ctx_destroy(g_ctx);
        g_ctx = nullptr;
    }
    g_device = device;
}

qk_ctx *device_context() {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ctx) {
        if (g_device < 0) {
            const char *e = std::getenv("QK_DEVICE");
            g_device = e ? std::atoi(e) : 0;
        }
        const qk_status st = qk_ctx_create(g_device, 0, &g_ctx);
        if (st != QK_OK) {
            g_ctx = nullptr;
            throw std::runtime_error(std::string("MI355X device ") + std::to_string(g_device) +
                                     " unavailable: " + qk_strerror(st));
        }
    }
    return g_ctx;
}

qk_contract to_contract(const KC::UInputs &c) {
    qk_contract k{};
    k.model = c.model.id_model == KI::MathModel::Bates ? QK_MODEL_BATES : QK_MODEL_HESTON;
    k.scheme = (int32_t)c.scheme.id_scheme;
    k.precision = c.mc.precision == KI::Precision::FP32 ? QK_FP32 : QK_FP64;
    k.n_steps = c.time.N_time_steps;
    k.n_paths = c.mc.N_Paths;
    k.seed = c.mc.rng_seed;
    k.S0 = c.market.S0;
    k.v0 = c.market.v0;
    k.r = c.market.r;
    k.q = c.market.q;
    k.T = c.time.t_end;
    const auto &h = c.model.active();
    k.kappa = h.k;
    k.theta = h.theta;
    k.sigma = h.sigma;
    k.rho = h.rho;
    k.lambda_j = c.model.bates.lambda_J;
    k.mu_j = c.model.bates.mu_J;
    k.sigma_j = c.model.bates.sigma_J;
    k.price_scale = c.get_price_scaling_factor();
    return k;
}

qk_payoff to_payoff(const KC::UInputs &c) {
    return qk_payoff{(int32_t)c.options.opt_type, c.options.opt_right == KI::OptRight::Call ? QK_CALL : QK_PUT,
                     c.options.StrikePrice, c.options.BarrierPrice};
}

qk_bumps to_bumps(const KC::UInputs &c) {
    qk_bumps b{};
    b.delta_gamma = c.mc.compute_delta_et_gamma;
    b.vega_vomma = c.mc.compute_vega_et_vomma;
    b.vanna = c.mc.compute_vanna;
    b.rho = c.mc.compute_rho;
    b.theta = c.mc.compute_theta && c.time.t_end > c.mc.time_absolute_bump_size;
    b.theta_central = 0;
    b.rel_S = c.mc.spot_price_relative_bump_size;
    b.abs_vol = c.mc.volatility_absolute_bump_size;
    b.abs_r = c.mc.risk_free_rate_absolute_bump_size;
    b.abs_T = c.mc.time_absolute_bump_size;
    return b;
}

// --------------------------------------------------------------------------
// Printing in the reference's formats (OptionPricer.hpp:251-290,
// MCEngine.hpp:76-86, :104-117).
// --------------------------------------------------------------------------
void MCTimer::print_summary() const {
    const char *ind = "  ";
    std::cout << "\n\n" << ind << "========================================\n"
              << ind << "        MC ELAPSED TIME    \n"
              << ind << "========================================\n"
              << ind << " Total time spent for the computations    : " << s_elapsed_total << "s\n"
              << ind << " Time spent evaluating the option price   : " << s_elapsed_for_pricing << "s\n"
              << ind << " Time spent evaluating the greeks         : " << s_elapsed_for_greeks << "s\n"
              << ind << "========================================\n";
}

void MCGreeks::print_summary() const {
    const char *ind = "  ";
    std::cout << "\n\n" << ind << "========================================\n"
              << ind << "        MONTE CARLO GREEKS    \n"
              << ind << "========================================\n"
              << ind << " Delta <-> Sensitivity to underlying spot price (S)    : " << delta << "\n"
              << ind << " Gamma <-> Delta curvature                             : " << gamma << "\n"
              << ind << " Vega <-> Sensitivity to volatility (v0)               : " << vega << "\n"
              << ind << " Vomma <-> Vega curvature                              : " << vomma << "\n"
              << ind << " Vanna <-> Cross sensitivity (spot price - volatility) : " << vanna << "\n"
              << ind << " Rho <-> Sensitivity to interest rates (r)             : " << rho << "\n"
              << ind << " Theta <-> Sensitivity to time decay (T)               : " << theta << "\n"
              << ind << "========================================\n";
}

// evaluate_option_price (OptionPricer.hpp:204-249): the device has already
// applied the discount factor (none for American) and the price scale.
void OptionPricer::evaluate_option_price() {
    if (config.mc.N_Paths <= 0) throw std::runtime_error("Cannot compute metrics: Invalid number of paths.");
    merge_batches();
    metrics.option_price = stats.price;
    metrics.standard_error = stats.se;
    metrics.prob_itm = stats.prob_itm;
    if (config.mc.analyze_risk_neutral_payoff_distribution) {
        if (!has_dist) throw std::runtime_error("Distribution analysis requested, but payoffs array is empty.");
        metrics.p0 = dist.p0;
        metrics.p1 = dist.p1;
        metrics.p5 = dist.p5;
        metrics.p10 = dist.p10;
        metrics.p20 = dist.p20;
        metrics.p30 = dist.p30;
        metrics.p40 = dist.p40;
        metrics.median = dist.median;
        metrics.p60 = dist.p60;
        metrics.p70 = dist.p70;
        metrics.p80 = dist.p80;
        metrics.p90 = dist.p90;
        metrics.p95 = dist.p95;
        metrics.p99 = dist.p99;
        metrics.p100 = dist.p100;
    }
    metrics.analytic_price = analytic;
    metrics.analytic_err_se = std::isfinite(analytic) && stats.se > 0.0 ? (stats.price - analytic) / stats.se
                                                                         : std::numeric_limits<double>::quiet_NaN();
    metrics.is_computed = true;
}

void OptionPricer::print_info_option_price() const {
    if (!metrics.is_computed)
        throw std::runtime_error("Attempted to print report before computing metrics. Call the method that computes "
                                 "the payoffs first.");
    std::ios old(nullptr);
    old.copyfmt(std::cout);
    std::cout << "  ========================================================\n"
              << "                     SIMULATION RESULTS\n"
              << "  ========================================================\n"
              << std::fixed << std::setprecision(6)
              << "   Estimated Option Price   :  " << metrics.option_price << "\n"
              << "   Statistical Std Error    :  " << metrics.standard_error << "\n"
              << "   Probability of ITM       :  " << metrics.prob_itm * 100.0 << " %\n";
    if (config.mc.analyze_risk_neutral_payoff_distribution) {
        std::cout << "  --------------------------------------------------------\n"
                  << "   [Risk-Neutral Discounted Payout Distribution]\n"
                  << "   Min  (Lower)             :  " << metrics.p0 << "\n"
                  << "   1st  Percentile          :  " << metrics.p1 << "\n"
                  << "   5th  Percentile          :  " << metrics.p5 << "\n"
                  << "   10th Percentile          :  " << metrics.p10 << "\n"
                  << "   20th Percentile          :  " << metrics.p20 << "\n"
                  << "   30th Percentile          :  " << metrics.p30 << "\n"
                  << "   40th Percentile          :  " << metrics.p40 << "\n"
                  << "   50th Percentile (Median) :  " << metrics.median << "\n"
                  << "   60th Percentile          :  " << metrics.p60 << "\n"
                  << "   70th Percentile          :  " << metrics.p70 << "\n"
                  << "   80th Percentile          :  " << metrics.p80 << "\n"
                  << "   90th Percentile          :  " << metrics.p90 << "\n"
                  << "   95th Percentile          :  " << metrics.p95 << "\n"
                  << "   99th Percentile          :  " << metrics.p99 << "\n"
                  << "   Max  (Upper)             :  " << metrics.p100 << "\n"
                  << "  --------------------------------------------------------\n"
                  << "   Time Spent to Sorting the Payoffs   :  " << dist.select_s << "s \n";
    }
    if (std::isfinite(metrics.analytic_price)) {
        std::cout << "  --------------------------------------------------------\n"
                  << "   Semi-analytic Price (CF) :  " << metrics.analytic_price << "\n"
                  << "   MC - Semi-analytic       :  " << metrics.option_price - metrics.analytic_price << "  ("
                  << std::setprecision(2) << metrics.analytic_err_se << " SE)\n"
                  << std::setprecision(6);
    }
    std::cout << "  ========================================================\n" << std::endl;
    std::cout.copyfmt(old);
}

const OptionPricer::MCOpPrices &OptionPricer::get_option_price_data() const {
    if (!metrics.is_computed) throw std::runtime_error("Metrics have not been computed yet.");
    return metrics;
}

// --------------------------------------------------------------------------
// Runners.
// --------------------------------------------------------------------------
OptionPricer::MCOpPrices price_forward(const KC::UInputs &scaled, bool print, bool *stored) {
    qk_ctx *ctx = device_context();
    const qk_contract c = to_contract(scaled);
    const qk_payoff p = to_payoff(scaled);
    qk_stats s{};
    OptionPricer pr(scaled);
    const bool keep_paths = !scaled.output.filename_paths_out.empty();
    if (stored) *stored = false;
    if (scaled.mc.analyze_risk_neutral_payoff_distribution) {
        qk_distribution d{};
        check(qk_price_distribution(ctx, &c, &p, &s, &d), "qk_price_distribution");
        pr.accumulate_device_distribution(d);
    } else if (keep_paths) {
        // the reference's StorePaths kernel when an output file is named (MSolver.hpp:117-119)
        check(qk_price_store_paths(ctx, &c, &p, &s), "qk_price_store_paths");
        if (stored) *stored = true;
    } else if (scaled.mc.batch_size > 0 && scaled.mc.batch_size < scaled.mc.N_Paths) {
        // MC.batch_size: the reference's batch loop (ForwardMC.hpp:113-135), each
        // batch a window of the path set; the same price as one launch
        s = price_in_batches(scaled);
    } else {
        check(qk_price(ctx, &c, &p, &s), "qk_price");
    }
    if (p.type == QK_EUROPEAN) {
        double a = 0.0;
        if (qk_analytic_european(&c, p.right, p.K, nullptr, &a) == QK_OK) pr.set_analytic_price(a);
    }
    pr.accumulate_device_stats(s);
    pr.evaluate_option_price();
    if (print) pr.print_info_option_price();
    return pr.get_option_price_data();
}

OptionPricer::MCOpPrices price_american(const KC::UInputs &scaled, bool print) {
    qk_ctx *ctx = device_context();
    const qk_contract c = to_contract(scaled);
    const int right = scaled.options.opt_right == KI::OptRight::Call ? QK_CALL : QK_PUT;
    qk_stats s{};
    OptionPricer pr(scaled);
    if (scaled.mc.analyze_risk_neutral_payoff_distribution) {
        const qk_payoff p{QK_AMERICAN, right, scaled.options.StrikePrice, 0.0};
        qk_distribution d{};
        check(qk_price_distribution(ctx, &c, &p, &s, &d), "qk_price_distribution");
        pr.accumulate_device_distribution(d);
    } else {
        const qk_lsm_options opt{(int32_t)scaled.mc.lsm_basis, scaled.mc.lsm_weighted_basis ? 1 : 0};
        check(qk_price_american_ex(ctx, &c, right, scaled.options.StrikePrice, &opt, &s, nullptr),
              "qk_price_american_ex");
    }
    pr.accumulate_device_stats(s);
    pr.evaluate_option_price();
    if (print) pr.print_info_option_price();
    return pr.get_option_price_data();
}

// ForwardMCRunner (ForwardMC.hpp:56-135) with the Greeks as lanes of one launch.
MCResults run_forward(KC::UInputs config) {
    config.apply_price_scaling();
    MCTimer timer;
    timer.start_total();
    timer.start_pricing();
    bool stored = false;
    const auto base = price_forward(config, true, &stored);
    if (!config.output.filename_paths_out.empty()) write_paths_file(config, stored);
    timer.stop_pricing();
    MCGreeks g{};
    int32_t lanes = 0;
    if (config.must_compute_greeks()) {
        timer.start_greeks();
        const qk_contract c = to_contract(config);
        const qk_payoff p = to_payoff(config);
        const qk_bumps b = to_bumps(config);
        qk_greeks out{};
        check(qk_price_greeks(device_context(), &c, &p, 1, &b, &out), "qk_price_greeks");
        g = MCGreeks{out.delta, out.gamma, out.vega, out.vomma, out.vanna, out.rho, out.theta};
        lanes = out.n_lanes;
        timer.stop_greeks();
        g.print_summary();
    }
    timer.stop_total();
    timer.print_summary();
    return MCResults{config, base, g, timer, lanes};
}

// ForwardBackwardMCRunner (ForwardBackwardMC.hpp:75-157): LSM pricing, Greeks
// by the generic sequential stencil.
MCResults run_forward_backward(KC::UInputs config) {
    config.apply_price_scaling();
    bool first = true;
    MCResults r = execute_mc_framework(config, [&](const KC::UInputs &c) {
        const bool print = first;
        first = false;
        auto m = price_american(c, print);
        // the base pricing's path store is the path set of the output file
        // (the Greek repricings name none, MCEngine.hpp:583)
        if (print && !c.output.filename_paths_out.empty()) write_paths_file(c, true);
        return m;
    });
    if (config.must_compute_greeks()) r.Greeks.print_summary();
    r.Timing.print_summary();
    return r;
}

MCResults MCDispatcher::launch_montecarlo() const {
    if (config.options.opt_type == KI::OptType::American) return run_forward_backward(config);
    return run_forward(config);
}

}  // namespace quantkos::Engine
