// qk_config.cpp -- validation, price scaling and the YAML parser of UInputs.
// Rules and messages follow the reference's Config.hpp and ConfigParser.cpp.
#include "qk_config.hpp"

#include <chrono>
#include <cmath>
#include <filesystem>
#include <iostream>

#include "yaml_lite.hpp"

namespace quantkos {

namespace Implemented {
using enum OptType;

template <> const std::map<std::string, OptType> &names<OptType>() {
    static const std::map<std::string, OptType> m = {
        {"European", European}, {"Asian", Asian}, {"BarrierUpAndOut", BarrierUpAndOut},
        {"BarrierDownAndOut", BarrierDownAndOut}, {"BarrierUpAndIn", BarrierUpAndIn},
        {"BarrierDownAndIn", BarrierDownAndIn}, {"LookbackFloatingStrike", LookbackFloatingStrike},
        {"LookbackFixedStrike", LookbackFixedStrike}, {"BinaryCashOrNothing", BinaryCashOrNothing},
        {"BinaryAssetOrNothing", BinaryAssetOrNothing}, {"American", American}};
    return m;
}
template <> const std::map<std::string, OptRight> &names<OptRight>() {
    static const std::map<std::string, OptRight> m = {{"Call", OptRight::Call}, {"Put", OptRight::Put}};
    return m;
}
template <> const std::map<std::string, MathModel> &names<MathModel>() {
    static const std::map<std::string, MathModel> m = {{"Heston", MathModel::Heston}, {"Bates", MathModel::Bates}};
    return m;
}
template <> const std::map<std::string, NumScheme> &names<NumScheme>() {
    static const std::map<std::string, NumScheme> m = {
        {"Euler", NumScheme::Euler}, {"ImplicitMilstein", NumScheme::ImplicitMilstein},
        {"AndersonQE", NumScheme::AndersonQE}};
    return m;
}
template <> const std::map<std::string, IOFormat> &names<IOFormat>() {
    static const std::map<std::string, IOFormat> m = {{"BIN", IOFormat::BIN}, {"TXT", IOFormat::TXT}};
    return m;
}
template <> const std::map<std::string, VerbosityLevel> &names<VerbosityLevel>() {
    static const std::map<std::string, VerbosityLevel> m = {
        {"None", VerbosityLevel::None}, {"Low", VerbosityLevel::Low}, {"Medium", VerbosityLevel::Medium},
        {"High", VerbosityLevel::High}};
    return m;
}
template <> const std::map<std::string, Precision> &names<Precision>() {
    static const std::map<std::string, Precision> m = {{"FP32", Precision::FP32}, {"FP64", Precision::FP64}};
    return m;
}

template <> const std::map<std::string, LSRegressionBasis> &names<LSRegressionBasis>() {
    static const std::map<std::string, LSRegressionBasis> m = {{"LaguerreP02", LSRegressionBasis::LaguerreP02},
                                                               {"LaguerreP03", LSRegressionBasis::LaguerreP03},
                                                               {"LaguerreP04", LSRegressionBasis::LaguerreP04}};
    return m;
}

template <typename T> T string_to_enum(const std::string &input, const std::string &field) {
    const auto &m = names<T>();
    auto it = m.find(input);
    if (it != m.end()) return it->second;
    std::string err = "Config Error: Invalid value '" + input + "' for " + field + ".\nAllowed options: [ ";
    for (const auto &kv : m) err += kv.first + " ";
    throw std::runtime_error(err + "]");
}
template OptType string_to_enum<OptType>(const std::string &, const std::string &);
template OptRight string_to_enum<OptRight>(const std::string &, const std::string &);
template MathModel string_to_enum<MathModel>(const std::string &, const std::string &);
template NumScheme string_to_enum<NumScheme>(const std::string &, const std::string &);
template IOFormat string_to_enum<IOFormat>(const std::string &, const std::string &);
template VerbosityLevel string_to_enum<VerbosityLevel>(const std::string &, const std::string &);
template Precision string_to_enum<Precision>(const std::string &, const std::string &);
template LSRegressionBasis string_to_enum<LSRegressionBasis>(const std::string &, const std::string &);
}  // namespace Implemented

namespace Config {
namespace KI = Implemented;

std::string config_err_msg(std::string_view block, std::string_view key, const std::string &msg) {
    return "[" + std::string(block) + "." + std::string(key) + "] " + msg;
}

// Config.hpp:63-84
void MarketConfig::validate() const {
    if (S0 <= 0.0) throw std::invalid_argument(config_err_msg(Keys::Market, "Price", "must be positive."));
    if (v0 < 0.0) throw std::invalid_argument(config_err_msg(Keys::Market, "Variance", "must be positive."));
    if (q < 0.0)
        throw std::invalid_argument(config_err_msg(Keys::Market, "q", "Dividend yield (q) cannot be negative."));
    if (std::abs(r) > 0.5)
        throw std::invalid_argument(config_err_msg(
            Keys::Market, "r", "Risk-free rate (r) is unphysical (keep between -50% and +50%)."));
    if (q > 0.8)
        throw std::invalid_argument(
            config_err_msg(Keys::Market, "q", "Dividend yield (q) is unphysical (keep under 80%)."));
}

// Config.hpp:120-142
void OptionsConfig::validate(double spot) const {
    if (StrikePrice <= 0.0) throw std::invalid_argument(config_err_msg(Keys::Options, "StrikePrice", "must be positive."));
    if (StrikePrice < spot * 0.0001)
        throw std::invalid_argument("Config Error: Strike price (K) is too close to zero. Must be at least 1000 times "
                                    "smaller than the initial asset price (S0) to prevent numerical problems.");
    if (StrikePrice > spot * 10000.0)
        throw std::invalid_argument("Config Error: Strike price (K) is absurdly high (exceeds 10000x of S0). This will "
                                    "likely result in zero-variance path generation and statistical breakdown.");
}

// Config.hpp:198-313
void MathModelConfig::validate() const {
    const Heston &h = active();
    if (h.k < 0.0) throw std::invalid_argument("[Model Config] κ must be positive.");
    if (h.theta < 0.0) throw std::invalid_argument("[Model Config] θ must be positive.");
    if (h.sigma < 0.0) throw std::invalid_argument("[Model Config] σ must be positive.");
    if (h.rho < -1.0 || h.rho > 1.0) throw std::invalid_argument("[Model Config] ρ must fall within [-1.0, 1.0].");
    if (id_model == KI::MathModel::Bates) {
        if (bates.lambda_J < 0.0)
            throw std::invalid_argument("[Model Config] Bates: Jump intensity (λ) cannot be negative.");
        if (bates.sigma_J < 0.0)
            throw std::invalid_argument("[Model Config] Bates: Jump volatility (σJ) must be positive.");
    }
}

// Config.hpp:347-373: N = ceil(T / Inp_DT), dt = T / N.
void TimeConfig::validate() {
    if (t_end <= 0.0) throw std::invalid_argument(config_err_msg(Keys::Time, "T_End", "must be positive."));
    if (inp_dt <= 0.0) throw std::invalid_argument(config_err_msg(Keys::Time, "DT", "must be positive."));
    if (inp_dt > t_end) throw std::invalid_argument(config_err_msg(Keys::Time, "DT", "must be smaller that total time."));
    N_time_steps = static_cast<int>(std::ceil(t_end / inp_dt));
    dt = t_end / N_time_steps;
    if (std::abs(dt - inp_dt) > 1e-12) {
        std::cout << "[Warning] Time step dt (" << dt << "s) does not evenly divide End Time (" << t_end << ").\n"
                  << "          Automatically adjusted dt to " << dt << "across " << N_time_steps
                  << " uniform steps.\n" << std::endl;
    }
}

// Config.hpp:431-563
void MCConfig::validate(const MarketConfig &m, double t_end) const {
    if (N_Paths < 1) throw std::invalid_argument(config_err_msg(Keys::MC, "N_Paths", "Must be a positive integer."));
    if (batch_size < -1)
        throw std::invalid_argument(config_err_msg(
            Keys::MC, "Batch_Size",
            "Must be a positive integer if you want to impose it, -1 if you want it equal to the max number of paths, "
            "and 0 if you want to let the machine handle this."));
    if (batch_size > N_Paths)
        throw std::invalid_argument(config_err_msg(Keys::MC, "Batch_Size",
                                                   "The batch size must be smaller that the notal number of realizations."));
    if (rng_seed == 0) throw std::invalid_argument(config_err_msg(Keys::MC, "RNG_Seed", "The random number seed cannot be 0."));
    if (Max_VRAM_MB <= 0) throw std::invalid_argument(config_err_msg(Keys::MC, "Max_VRAM_MB", "must be a positive integer."));
    if (Max_CPU_RAM_MB <= 0)
        throw std::invalid_argument(config_err_msg(Keys::MC, "Max_CPU_RAM_MB", "must be a positive integer."));
    if (!must_compute_greeks()) return;
    const bool fp32 = precision == KI::Precision::FP32;
    if (compute_delta_et_gamma || compute_vanna) {
        if (spot_price_relative_bump_size <= 0.0 || spot_price_relative_bump_size > 0.05)
            throw std::invalid_argument(config_err_msg(Keys::MC, "spot_price_relative_bump_size",
                                                       "must be strictly positive and <= 5% (0.05)."));
        const double s_plus = fp32 ? (double)((float)m.S0 * (1.0f + (float)spot_price_relative_bump_size))
                                   : m.S0 * (1.0 + spot_price_relative_bump_size);
        if (s_plus == (fp32 ? (double)(float)m.S0 : m.S0))
            throw std::invalid_argument(config_err_msg(
                Keys::MC, "spot_price_relative_bump_size",
                "Spot bump is too small; bumped S0 collapses to S0 under current pricing precision.\n"));
    }
    if (compute_vega_et_vomma || compute_vanna) {
        const double vol0 = std::sqrt(m.v0);
        if (volatility_absolute_bump_size <= 0.0 || volatility_absolute_bump_size > vol0)
            throw std::invalid_argument(config_err_msg(
                Keys::MC, "volatility_absolute_bump_size",
                "must be strictly positive and and strictly smaller than sqrt(v0) when Vega/Vomma or Vanna is requested"));
    }
    if (compute_rho && (risk_free_rate_absolute_bump_size <= 0.0 || risk_free_rate_absolute_bump_size > 0.05))
        throw std::invalid_argument(
            config_err_msg(Keys::MC, "volatility_absolute_bump_size", "must be strictly positive and <= 0.05."));
    if (compute_theta && (time_absolute_bump_size <= 0.0 || time_absolute_bump_size >= t_end))
        throw std::invalid_argument(config_err_msg(
            Keys::MC, "time_absolute_bump_size",
            "must be strictly positive and strictly smaller than T_End (t_end) when Theta is requested."));
}

// Config.hpp:640-669
void OutputConfig::validate() const {
    if (!out_dir.empty()) {
        try {
            std::filesystem::create_directories(out_dir);
        } catch (const std::filesystem::filesystem_error &e) {
            throw std::invalid_argument("[Configuration Error] The assigned output directory '" + out_dir +
                                        "' could not be prepared or accessed.\nDetails: " + e.what());
        }
    }
    if (format != KI::IOFormat::BIN)
        throw std::invalid_argument("[Configuration Error] Unsupported output format requested. Currently, only the "
                                    "BIN (Binary) writer is fully implemented and supported.");
}

void UInputs::validate() {
    market.validate();
    options.validate(market.S0);
    model.validate();
    time.validate();
    mc.validate(market, time.t_end);
    output.validate();
}

// Config.hpp:710-723
void UInputs::apply_price_scaling() {
    if (!mc.normalize_prices) return;
    price_scale_factor = market.S0;
    market.S0 = 1.0;
    options.StrikePrice /= price_scale_factor;
    options.BarrierPrice /= price_scale_factor;
}

// Config.hpp:750-758
bool UInputs::requires_paths_allocated() const {
    return !output.filename_paths_out.empty() || options.opt_type == KI::OptType::American;
}

void UInputs::print_summary() const {
    const char *ind = "    ";
    std::cout << "\n" << ind << "========================================================\n"
              << ind << "                      MC CONFIGURATION                    \n"
              << ind << "========================================================\n"
              << ind << "  [Market]  S0=" << market.S0 << " v0=" << market.v0 << " r=" << market.r
              << " q=" << market.q << "\n"
              << ind << "  [Options] " << KI::enum_to_string(options.opt_type) << " "
              << KI::enum_to_string(options.opt_right) << " K=" << options.StrikePrice
              << " B=" << options.BarrierPrice << "\n";
    const auto &h = model.active();
    std::cout << ind << "  [Model]   " << KI::enum_to_string(model.id_model) << " k=" << h.k << " theta=" << h.theta
              << " sigma=" << h.sigma << " rho=" << h.rho;
    if (model.id_model == KI::MathModel::Bates)
        std::cout << " lambda_J=" << model.bates.lambda_J << " mu_J=" << model.bates.mu_J
                  << " sigma_J=" << model.bates.sigma_J;
    std::cout << "\n" << ind << "  [Numerics] " << KI::enum_to_string(scheme.id_scheme) << "  precision "
              << KI::enum_to_string(mc.precision) << "\n"
              << ind << "  [Time]    T=" << time.t_end << " dt=" << time.dt << " N=" << time.N_time_steps << "\n"
              << ind << "  [MC]      N_Paths=" << mc.N_Paths << " seed=" << mc.rng_seed << "\n"
              << ind << "========================================================\n" << std::endl;
}

// ---------------------------------------------------------------------------
// Parser (src/IO/ConfigParser.cpp:36-294).
// ---------------------------------------------------------------------------
namespace {

UInputs from_root(const yaml::Node &root) {
    UInputs conf;
    auto need = [&](std::string_view block) -> const yaml::Node & {
        const yaml::Node &n = root[std::string(block)];
        if (!n) throw std::runtime_error("Config Error: Mandatory block '" + std::string(block) + "' missing.");
        return n;
    };
    {
        const auto &n = need(Keys::Market);
        if (n["Price"]) conf.market.S0 = n["Price"].as<double>();
        if (n["Variance"]) conf.market.v0 = n["Variance"].as<double>();
        if (n["r"]) conf.market.r = n["r"].as<double>();
        if (n["q"]) conf.market.q = n["q"].as<double>();
    }
    {
        const auto &n = need(Keys::Options);
        if (n["OptionType"])
            conf.options.opt_type = KI::string_to_enum<KI::OptType>(n["OptionType"].as<std::string>(), "Options.OptionType");
        if (n["OptionRight"])
            conf.options.opt_right = KI::string_to_enum<KI::OptRight>(n["OptionRight"].as<std::string>(), "Options.OptionRight");
        if (n["StrikePrice"]) conf.options.StrikePrice = n["StrikePrice"].as<double>();
        if (n["BarrierPrice"]) conf.options.BarrierPrice = n["BarrierPrice"].as<double>();
    }
    {
        const auto &n = need(Keys::Model);
        const bool has_h = n["Heston"].IsDefined(), has_b = n["Bates"].IsDefined();
        if (has_h && has_b)
            throw std::runtime_error("Config Error: Cannot define both Heston and Bates models simultaneously.");
        auto read_heston = [](const yaml::Node &h, MathModelConfig::Heston &p) {
            if (h["k"]) p.k = h["k"].as<double>();
            if (h["theta"]) p.theta = h["theta"].as<double>();
            if (h["sigma"]) p.sigma = h["sigma"].as<double>();
            if (h["rho"]) p.rho = h["rho"].as<double>();
        };
        if (has_h) {
            conf.model.id_model = KI::MathModel::Heston;
            read_heston(n["Heston"], conf.model.heston);
        } else if (has_b) {
            conf.model.id_model = KI::MathModel::Bates;
            const auto &b = n["Bates"];
            read_heston(b, conf.model.bates);
            if (b["lambda_J"]) conf.model.bates.lambda_J = b["lambda_J"].as<double>();
            if (b["mu_J"]) conf.model.bates.mu_J = b["mu_J"].as<double>();
            if (b["sigma_J"]) conf.model.bates.sigma_J = b["sigma_J"].as<double>();
        } else {
            throw std::runtime_error("Config Error: Mandatory block 'Model' has requested a non implemented model!");
        }
    }
    if (const auto &n = root["Numerics"]; n) {
        if (n["Scheme"])
            conf.scheme.id_scheme = KI::string_to_enum<KI::NumScheme>(n["Scheme"].as<std::string>(), "Numerics.Scheme");
    }
    {
        const auto &n = need(Keys::Time);
        if (n["T_End"]) conf.time.t_end = n["T_End"].as<double>();
        if (n["Inp_DT"]) conf.time.inp_dt = n["Inp_DT"].as<double>();
    }
    {
        const auto &n = need(Keys::MC);
        if (n["Normalize_prices"]) {
            try {
                conf.mc.normalize_prices = n["Normalize_prices"].as<bool>();
            } catch (...) {
                throw std::runtime_error(config_err_msg(Keys::MC, "Normalize_prices",
                                                        "must be a boolean value (true/false, yes/no, 1/0)."));
            }
        }
        if (n["N_Paths"]) conf.mc.N_Paths = n["N_Paths"].as<int>();
        if (n["Batch_Size"]) conf.mc.batch_size = n["Batch_Size"].as<int>();
        if (n["RNG_Seed"]) conf.mc.rng_seed = n["RNG_Seed"].as<unsigned long long>();
        if (n["Max_VRAM_MB"]) conf.mc.Max_VRAM_MB = n["Max_VRAM_MB"].as<long long>();
        if (n["Max_CPU_RAM_MB"]) conf.mc.Max_CPU_RAM_MB = n["Max_CPU_RAM_MB"].as<long long>();
        if (n["analyze_risk_neutral_payoff_distribution"])
            conf.mc.analyze_risk_neutral_payoff_distribution = n["analyze_risk_neutral_payoff_distribution"].as<bool>();
        if (n["compute_delta_et_gamma"]) conf.mc.compute_delta_et_gamma = n["compute_delta_et_gamma"].as<bool>();
        if (n["spot_price_relative_bump_size"])
            conf.mc.spot_price_relative_bump_size = n["spot_price_relative_bump_size"].as<double>();
        if (n["compute_vega_et_vomma"]) conf.mc.compute_vega_et_vomma = n["compute_vega_et_vomma"].as<bool>();
        if (n["compute_vanna"]) conf.mc.compute_vanna = n["compute_vanna"].as<bool>();
        if (n["volatility_absolute_bump_size"])
            conf.mc.volatility_absolute_bump_size = n["volatility_absolute_bump_size"].as<double>();
        if (n["compute_rho"]) conf.mc.compute_rho = n["compute_rho"].as<bool>();
        if (n["risk_free_rate_absolute_bump_size"])
            conf.mc.risk_free_rate_absolute_bump_size = n["risk_free_rate_absolute_bump_size"].as<double>();
        if (n["compute_theta"]) conf.mc.compute_theta = n["compute_theta"].as<bool>();
        if (n["time_absolute_bump_size"]) conf.mc.time_absolute_bump_size = n["time_absolute_bump_size"].as<double>();
        // This build's addition: arithmetic precision chosen at run time.
        if (n["Precision"]) conf.mc.precision = KI::string_to_enum<KI::Precision>(n["Precision"].as<std::string>(), "MC.Precision");
        // Also this build's: the Longstaff-Schwartz basis chosen at run time.
        if (n["LSM_Basis"])
            conf.mc.lsm_basis = KI::string_to_enum<KI::LSRegressionBasis>(n["LSM_Basis"].as<std::string>(), "MC.LSM_Basis");
        if (n["LSM_Weighted_Basis"]) conf.mc.lsm_weighted_basis = n["LSM_Weighted_Basis"].as<bool>();
    }
    if (const auto &n = root["Output"]; n) {
        if (n["Verbosity"])
            conf.output.verbosity = KI::string_to_enum<KI::VerbosityLevel>(n["Verbosity"].as<std::string>(), "Verbosity");
        if (n["out_dir"]) conf.output.out_dir = n["out_dir"].as<std::string>();
        if (n["Name_Paths_Out_File"]) conf.output.filename_paths_out = n["Name_Paths_Out_File"].as<std::string>();
        if (n["Name_Log_File"]) conf.output.filename_log = n["Name_Log_File"].as<std::string>();
        if (n["Format"]) conf.output.format = KI::string_to_enum<KI::IOFormat>(n["Format"].as<std::string>(), "Output.Format");
    }
    return conf;
}

UInputs finish(UInputs conf, bool quiet, std::chrono::steady_clock::time_point t0) {
    conf.validate();
    if (!quiet) {
        std::cout << "  [ Config ] Input configuration validated successfully\n\n";
        conf.print_summary();
        const std::chrono::duration<double> el = std::chrono::steady_clock::now() - t0;
        std::cout << "  [ Parser ] Time taken to parse the input  : " << el.count() << "s \n";
    }
    return conf;
}

}  // namespace

UInputs Parser::parse(const std::string &file, bool quiet) {
    const auto t0 = std::chrono::steady_clock::now();
    if (!quiet) std::cout << "\n>>> [ Parser ] Start Parsing User Inputs...\n\n";
    if (!std::filesystem::exists(file)) throw std::runtime_error("Config Error: File not found -> " + file);
    yaml::Node root;
    try {
        root = yaml::Node::LoadFile(file);
    } catch (const yaml::ParserException &e) {
        throw std::runtime_error("Config Error: Invalid YAML syntax -> " + std::string(e.what()));
    }
    return finish(from_root(root), quiet, t0);
}

UInputs Parser::parse_text(const std::string &text, bool quiet) {
    const auto t0 = std::chrono::steady_clock::now();
    yaml::Node root;
    try {
        root = yaml::Node::Load(text);
    } catch (const yaml::ParserException &e) {
        throw std::runtime_error("Config Error: Invalid YAML syntax -> " + std::string(e.what()));
    }
    return finish(from_root(root), quiet, t0);
}

}  // namespace Config
}  // namespace quantkos
