// qk_config.hpp -- QuantKos's configuration schema (UInputs) for the MI355X build.
//
// Same YAML keys, defaults, validation rules and error texts as the reference
// (include/core/config/Config.hpp:52-775, ConfigFileKeys.hpp:36-158,
// ConfigFileEnums.hpp:24-125, ConfigFileKeysEnumMaps.hpp:48-192), so existing
// configuration files load unchanged.  Differences: values are held in double
// and the arithmetic precision is a run-time choice (MC.Precision: FP32 | FP64,
// default FP64 as the reference's default build), where the reference fixes it
// at compile time (Typedefs.hpp:31-35).
#pragma once

#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <string_view>

namespace quantkos {

namespace Implemented {
enum class OptType {
    European, Asian, BarrierUpAndOut, BarrierDownAndOut, BarrierUpAndIn, BarrierDownAndIn,
    LookbackFloatingStrike, LookbackFixedStrike, BinaryCashOrNothing, BinaryAssetOrNothing, American
};
enum class OptRight { Call, Put };
enum class MathModel { Heston, Bates };
enum class NumScheme { Euler, ImplicitMilstein, AndersonQE };
enum class IOFormat { BIN, TXT };
enum class VerbosityLevel : int { None = 0, Low = 1, Medium = 2, High = 3 };
enum class Precision { FP32, FP64 };
// Longstaff-Schwartz regression basis (ConfigFileEnums.hpp:107-112).
enum class LSRegressionBasis { LaguerreP02, LaguerreP03, LaguerreP04 };

template <typename T> const std::map<std::string, T> &names();
template <typename T> std::string enum_to_string(T v) {
    for (const auto &[k, e] : names<T>())
        if (e == v) return k;
    return "Unknown";
}
// Throws "Config Error: Invalid value ..." listing the allowed names.
template <typename T> T string_to_enum(const std::string &input, const std::string &field);
}  // namespace Implemented

namespace Keys {
inline constexpr std::string_view Market = "Market", Options = "Options", Model = "Model",
                                  Numerics = "Numerics", Time = "Time", MC = "MC", Output = "Output";
}

namespace Config {

std::string config_err_msg(std::string_view block, std::string_view key, const std::string &msg);

struct MarketConfig {
    double S0 = 100.0, v0 = 1.0, r = 0.05, q = 0.0;
    void validate() const;
};

struct OptionsConfig {
    Implemented::OptType opt_type = Implemented::OptType::European;
    Implemented::OptRight opt_right = Implemented::OptRight::Call;
    double StrikePrice = 100.0, BarrierPrice = 100.0;
    void validate(double spot) const;
};

struct MathModelConfig {
    Implemented::MathModel id_model = Implemented::MathModel::Heston;
    struct Heston {
        double k = 2.0, theta = 0.04, sigma = 0.3, rho = -0.7;
    } heston;
    struct Bates : Heston {
        double lambda_J = 0.1, mu_J = -0.1, sigma_J = 0.15;
    } bates;
    void validate() const;
    const Heston &active() const { return id_model == Implemented::MathModel::Bates ? bates : heston; }
};

struct NumSchemeConfig {
    Implemented::NumScheme id_scheme = Implemented::NumScheme::Euler;
};

struct TimeConfig {
    double t_end = 1.0, inp_dt = 0.2, dt = 0.0;
    int N_time_steps = 0;
    void validate();
};

struct MCConfig {
    bool normalize_prices = true;
    int N_Paths = 1000;
    int batch_size = 0;
    uint64_t rng_seed = 184467440737095ULL;
    long long Max_VRAM_MB = 256, Max_CPU_RAM_MB = 4000;
    bool analyze_risk_neutral_payoff_distribution = false;
    bool compute_delta_et_gamma = false, compute_vega_et_vomma = false, compute_vanna = false,
         compute_rho = false, compute_theta = false;
    double spot_price_relative_bump_size = 0.001, volatility_absolute_bump_size = 0.01,
           risk_free_rate_absolute_bump_size = 0.0001, time_absolute_bump_size = 1.0 / 365.0;
    Implemented::Precision precision = Implemented::Precision::FP64;
    // The reference fixes the LSM basis at compile time (LSMAmerican.hpp:217);
    // here MC.LSM_Basis and MC.LSM_Weighted_Basis choose it per run.
    Implemented::LSRegressionBasis lsm_basis = Implemented::LSRegressionBasis::LaguerreP03;
    bool lsm_weighted_basis = false;
    void validate(const MarketConfig &m, double t_end) const;
    bool must_compute_greeks() const {
        return compute_delta_et_gamma || compute_vega_et_vomma || compute_vanna || compute_rho || compute_theta;
    }
};

struct OutputConfig {
    Implemented::VerbosityLevel verbosity = Implemented::VerbosityLevel::High;
    std::string out_dir = "outputs", filename_paths_out = "QuantKos.paths", filename_log = "QuantKos.log";
    Implemented::IOFormat format = Implemented::IOFormat::BIN;
    void validate() const;
};

struct UInputs {
    MarketConfig market;
    OptionsConfig options;
    MathModelConfig model;
    NumSchemeConfig scheme;
    TimeConfig time;
    MCConfig mc;
    OutputConfig output;

    void validate();
    void apply_price_scaling();
    void print_summary() const;
    bool requires_paths_allocated() const;
    double get_price_scaling_factor() const { return price_scale_factor; }
    bool must_compute_greeks() const { return mc.must_compute_greeks(); }

private:
    double price_scale_factor = 1.0;
};

// Config::Parser::parse (IO/InputParser.hpp:56, src/IO/ConfigParser.cpp:36-294).
struct Parser {
    static UInputs parse(const std::string &file, bool quiet = false);
    static UInputs parse_text(const std::string &yaml_text, bool quiet = false);
};

}  // namespace Config
}  // namespace quantkos
