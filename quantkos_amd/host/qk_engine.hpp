// qk_engine.hpp -- QuantKos's engine interfaces, running on the MI355X kernels.
//
// Drop-in counterparts of the reference's orchestration types:
//   MCDispatcher            (engine/Distpatcher.hpp:52-230)
//   ForwardMCRunner         (engine/ForwardMC.hpp:48-200)
//   ForwardBackwardMCRunner (engine/ForwardBackwardMC.hpp:68-256)
//   execute_mc_framework    (engine/MCEngine.hpp:560-746)
//   OptionPricer, MCOpPrices (options/OptionPricer.hpp:50-305)
//   MCGreeks, MCTimer, MCResults (engine/MCEngine.hpp:33-135)
// Every price comes from libquantkos's C ABI (include/qk.h); this layer maps
// UInputs onto qk_contract / qk_payoff / qk_bumps and formats the results.
#pragma once

#include <chrono>
#include <cmath>
#include <limits>
#include <string>

#include "qk.h"
#include "qk_config.hpp"

namespace quantkos::Engine {

namespace KI = quantkos::Implemented;
namespace KC = quantkos::Config;

// The device context used by every runner of the process (device from the
// QK_DEVICE environment variable, default 0).  Throws if no device is usable.
qk_ctx *device_context();
void set_device(int device);

qk_contract to_contract(const KC::UInputs &scaled);
qk_payoff to_payoff(const KC::UInputs &scaled);
qk_bumps to_bumps(const KC::UInputs &cfg);

struct MCTimer {
    double s_elapsed_total = 0.0, s_elapsed_for_pricing = 0.0, s_elapsed_for_greeks = 0.0;
    void start_total() { t_total = clock::now(); }
    void stop_total() { s_elapsed_total = since(t_total); }
    void start_pricing() { t_pricing = clock::now(); }
    void stop_pricing() { s_elapsed_for_pricing = since(t_pricing); }
    void start_greeks() { t_greeks = clock::now(); }
    void stop_greeks() { s_elapsed_for_greeks = since(t_greeks); }
    void print_summary() const;

private:
    using clock = std::chrono::steady_clock;
    clock::time_point t_total{}, t_pricing{}, t_greeks{};
    static double since(clock::time_point t) { return std::chrono::duration<double>(clock::now() - t).count(); }
};

struct MCGreeks {
    double delta = 0.0, gamma = 0.0, vega = 0.0, vomma = 0.0, vanna = 0.0, rho = 0.0, theta = 0.0;
    void print_summary() const;
};

class PathsMCBatchMem;

class OptionPricer {
public:
    struct MCOpPrices {
        double option_price = 0.0, standard_error = 0.0, prob_itm = 0.0;
        double p0 = -999., p1 = -999., p5 = -999., p10 = -999., p20 = -999., p30 = -999., p40 = -999.,
               median = -999., p60 = -999., p70 = -999., p80 = -999., p90 = -999., p95 = -999., p99 = -999.,
               p100 = -999.;
        bool is_computed = false;
        // European options: the semi-analytic price of the model (Heston / Bates
        // characteristic function) and the Monte Carlo price's distance from it
        // in standard errors; NaN for other payoffs.
        double analytic_price = std::numeric_limits<double>::quiet_NaN();
        double analytic_err_se = std::numeric_limits<double>::quiet_NaN();
    };
    explicit OptionPricer(const KC::UInputs &config) : config(config) {}
    // Takes the device's merged statistics of the whole path set (the
    // reference merges batch by batch, OptionPricer.hpp:123-193).
    void accumulate_device_stats(const qk_stats &s) { stats = s; }
    // The device-selected distribution of the same paths (analyze_risk_neutral_
    // discounted_payoff_distribution, OptionPricer.hpp:340-360).
    void set_analytic_price(double p) { analytic = p; }
    void accumulate_device_distribution(const qk_distribution &d) {
        dist = d;
        has_dist = true;
    }
    // Batch by batch (OptionPricer.hpp:123-193): the batches' device merge
    // chunks accumulate in BatchMem; evaluate_option_price merges them on the
    // device (qk_merge_windows) once they cover N_Paths.
    void accumulate_batch_payoffs(const PathsMCBatchMem &BatchMem, int64_t curr_batch_size);
    void evaluate_option_price();
    void print_info_option_price() const;
    const MCOpPrices &get_option_price_data() const;
    const qk_stats &device_stats() const { return stats; }

private:
    void merge_batches();
    const KC::UInputs &config;
    const PathsMCBatchMem *batch_mem = nullptr;
    int64_t batch_paths = 0;
    qk_stats stats{};
    qk_distribution dist{};
    bool has_dist = false;
    double analytic = std::numeric_limits<double>::quiet_NaN();
    MCOpPrices metrics;
};
using MCOpPrices = OptionPricer::MCOpPrices;

struct MCResults {
    KC::UInputs MCConfig;
    OptionPricer::MCOpPrices OptionPrice;
    MCGreeks Greeks;
    MCTimer Timing;
    int32_t n_greek_lanes = 0;   // lanes of the fused Greek launch (0: sequential stencil)
};

// The reference's generic engine: base price, then each Greek bump as a full
// repricing on the same random stream (MCEngine.hpp:576-735).  The forward
// runner uses the fused one-launch Greeks instead; this path serves American
// options, whose pricing is a forward-backward pipeline.
template <typename PricingCallable>
MCResults execute_mc_framework(const KC::UInputs &config, PricingCallable &&pricing_function);

// One pricing of a scaled configuration through the device.
// stored (if not null) is set when the call left the path set in the ctx's
// path store (an output file is named and no distribution was analysed).
OptionPricer::MCOpPrices price_forward(const KC::UInputs &scaled, bool print, bool *stored = nullptr);
OptionPricer::MCOpPrices price_american(const KC::UInputs &scaled, bool print);
MCResults run_forward(KC::UInputs config);
MCResults run_forward_backward(KC::UInputs config);

template <KI::MathModel M, KI::NumScheme S, KI::OptType T, KI::OptRight R>
class ForwardMCRunner {
public:
    explicit ForwardMCRunner(KC::UInputs conf) : config(std::move(conf)) {}
    MCResults run() const { return run_forward(config); }

private:
    KC::UInputs config;
};

template <KI::MathModel M, KI::NumScheme S, KI::OptRight R>
class ForwardBackwardMCRunner {
public:
    explicit ForwardBackwardMCRunner(KC::UInputs conf) : config(std::move(conf)) {}
    MCResults run() const { return run_forward_backward(config); }

private:
    KC::UInputs config;
};

class MCDispatcher {
public:
    explicit MCDispatcher(const KC::UInputs &config) : config(config) {}
    MCResults launch_montecarlo() const;

private:
    KC::UInputs config;
};

// --------------------------------------------------------------------------
template <typename PricingCallable>
MCResults execute_mc_framework(const KC::UInputs &config, PricingCallable &&pricing_function) {
    MCTimer timer;
    timer.start_total();
    timer.start_pricing();
    const OptionPricer::MCOpPrices base = pricing_function(config);
    timer.stop_pricing();
    MCGreeks g{};
    if (config.must_compute_greeks()) {
        timer.start_greeks();
        KC::UInputs b = config;
        b.mc.analyze_risk_neutral_payoff_distribution = false;
        b.output.filename_paths_out = "";
        const double S0 = b.market.S0;
        const double h_S = S0 * b.mc.spot_price_relative_bump_size;
        const double h_S_real = b.mc.normalize_prices
                                    ? b.get_price_scaling_factor() * b.mc.spot_price_relative_bump_size
                                    : h_S;
        const double vol0 = std::sqrt(b.market.v0), h_vol = b.mc.volatility_absolute_bump_size;
        const double v_up = (vol0 + h_vol) * (vol0 + h_vol), v_dn = (vol0 - h_vol) * (vol0 - h_vol);
        auto P = [&](auto edit) {
            KC::UInputs c = b;
            edit(c);
            return pricing_function(c).option_price;
        };
        if (b.mc.compute_delta_et_gamma) {
            const double pu = P([&](KC::UInputs &c) { c.market.S0 = S0 + h_S; });
            const double pd = P([&](KC::UInputs &c) { c.market.S0 = S0 - h_S; });
            g.delta = (pu - pd) / (2.0 * h_S_real);
            g.gamma = (pu - 2.0 * base.option_price + pd) / (h_S_real * h_S_real);
        }
        if (b.mc.compute_vega_et_vomma) {
            const double pu = P([&](KC::UInputs &c) { c.market.v0 = v_up; });
            const double pd = P([&](KC::UInputs &c) { c.market.v0 = v_dn; });
            g.vega = (pu - pd) / (2.0 * h_vol);
            g.vomma = (pu - 2.0 * base.option_price + pd) / (h_vol * h_vol);
        }
        if (b.mc.compute_rho) {
            const double h = b.mc.risk_free_rate_absolute_bump_size, r = b.market.r;
            g.rho = (P([&](KC::UInputs &c) { c.market.r = r + h; }) - P([&](KC::UInputs &c) { c.market.r = r - h; })) /
                    (2.0 * h);
        }
        if (b.mc.compute_theta) {
            const double T = b.time.t_end, h = b.mc.time_absolute_bump_size;
            if (T > h) {
                const double pd = P([&](KC::UInputs &c) {
                    c.time.t_end = T - h;
                    c.time.dt = c.time.t_end / b.time.N_time_steps;
                });
                g.theta = (pd - base.option_price) / h;
            }
        }
        if (b.mc.compute_vanna) {
            const double puu = P([&](KC::UInputs &c) { c.market.S0 = S0 + h_S; c.market.v0 = v_up; });
            const double pdu = P([&](KC::UInputs &c) { c.market.S0 = S0 - h_S; c.market.v0 = v_up; });
            const double pud = P([&](KC::UInputs &c) { c.market.S0 = S0 + h_S; c.market.v0 = v_dn; });
            const double pdd = P([&](KC::UInputs &c) { c.market.S0 = S0 - h_S; c.market.v0 = v_dn; });
            g.vanna = (puu - pdu - pud + pdd) / (4.0 * h_S_real * h_vol);
        }
        timer.stop_greeks();
    }
    timer.stop_total();
    return MCResults{config, base, g, timer, 0};
}

}  // namespace quantkos::Engine
