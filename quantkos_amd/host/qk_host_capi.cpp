// qk_host_capi.cpp -- C entry points of the host layer (include/qk.h):
// configuration loading and whole-file runs.
#include <cmath>
#include <cstring>
#include <iostream>
#include <sstream>
#include <stdexcept>

#include "qk.h"
#include "qk_config.hpp"
#include "qk_engine.hpp"

namespace {

void copy_err(char *err, size_t len, const std::string &msg) {
    if (!err || len == 0) return;
    std::strncpy(err, msg.c_str(), len - 1);
    err[len - 1] = '\0';
}

// Silences std::cout for the lifetime of the object.
struct Mute {
    explicit Mute(bool on) : on(on) {
        if (on) old = std::cout.rdbuf(sink.rdbuf());
    }
    ~Mute() {
        if (on) std::cout.rdbuf(old);
    }
    bool on;
    std::ostringstream sink;
    std::streambuf *old = nullptr;
};

}  // namespace

extern "C" qk_status qk_config_load(const char *yaml_path, qk_run_spec *out, char *err, size_t err_len) {
    using namespace quantkos;
    if (!yaml_path || !out) {
        copy_err(err, err_len, "null argument");
        return QK_EINVAL;
    }
    try {
        Mute m(true);
        Config::UInputs conf = Config::Parser::parse(yaml_path, true);
        conf.apply_price_scaling();
        std::memset(out, 0, sizeof(*out));
        out->contract = Engine::to_contract(conf);
        out->payoff = Engine::to_payoff(conf);
        out->compute_greeks = conf.must_compute_greeks();
        if (out->compute_greeks) out->bumps = Engine::to_bumps(conf);
        out->write_paths = !conf.output.filename_paths_out.empty();
        out->lsm = qk_lsm_options{(int32_t)conf.mc.lsm_basis, conf.mc.lsm_weighted_basis ? 1 : 0};
        out->analyze_distribution = conf.mc.analyze_risk_neutral_payoff_distribution ? 1 : 0;
        return QK_OK;
    } catch (const std::exception &e) {
        copy_err(err, err_len, e.what());
        return QK_EINVAL;
    }
}

extern "C" qk_status qk_run_config(const char *yaml_path, int device, int32_t quiet, qk_run_result *out, char *err,
                                   size_t err_len) {
    using namespace quantkos;
    if (!yaml_path || !out) {
        copy_err(err, err_len, "null argument");
        return QK_EINVAL;
    }
    Config::UInputs conf;
    try {
        Mute m(quiet != 0);
        conf = Config::Parser::parse(yaml_path, quiet != 0);
    } catch (const std::exception &e) {
        copy_err(err, err_len, e.what());
        return QK_EINVAL;
    }
    try {
        Mute m(quiet != 0);
        Engine::set_device(device);
        const Engine::MCResults r = Engine::MCDispatcher(conf).launch_montecarlo();
        std::memset(out, 0, sizeof(*out));
        out->stats.price = r.OptionPrice.option_price;
        out->stats.se = r.OptionPrice.standard_error;
        out->stats.prob_itm = r.OptionPrice.prob_itm;
        out->stats.n = conf.mc.N_Paths;
        out->stats.itm = (int64_t)std::llround(r.OptionPrice.prob_itm * conf.mc.N_Paths);
        out->greeks = qk_greeks{r.Greeks.delta, r.Greeks.gamma, r.Greeks.vega, r.Greeks.vomma, r.Greeks.vanna,
                                r.Greeks.rho, r.Greeks.theta, r.OptionPrice.option_price,
                                r.OptionPrice.standard_error, r.n_greek_lanes};
        const auto &o = r.OptionPrice;
        out->dist = qk_distribution{o.p0,  o.p1,  o.p5,  o.p10, o.p20, o.p30, o.p40,  o.median,
                                    o.p60, o.p70, o.p80, o.p90, o.p95, o.p99, o.p100, 0.0};
        out->t_total_s = r.Timing.s_elapsed_total;
        out->t_pricing_s = r.Timing.s_elapsed_for_pricing;
        out->t_greeks_s = r.Timing.s_elapsed_for_greeks;
        out->analytic_price = r.OptionPrice.analytic_price;
        out->analytic_err_se = r.OptionPrice.analytic_err_se;
        return QK_OK;
    } catch (const std::exception &e) {
        copy_err(err, err_len, e.what());
        return QK_EDEVICE;
    }
}
