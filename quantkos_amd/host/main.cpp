// main.cpp -- the `quantkos` command line, with the reference's argv contract
// (app/main.cpp:51-110): `quantkos <config.yaml>`, results on stdout, and any
// exception reported as "FATAL ERROR" with exit status 1.
//
//   quantkos [--device N] [--fp32 | --fp64] [--json] <config.yaml>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <string>

#include "qk_config.hpp"
#include "qk_engine.hpp"

static void usage() {
    std::cout << "Usage: quantkos [--device N] [--fp32 | --fp64] [--json] <config.yaml>\n"
              << "  Prices the option described by a QuantKos YAML configuration on an MI355X.\n";
}

int main(int argc, char **argv) {
    using namespace quantkos;
    std::string file;
    int device = -1;
    int precision = -1;
    bool json = false;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "-h" || a == "--help") {
            usage();
            return 0;
        } else if (a == "--device" && i + 1 < argc) {
            device = std::atoi(argv[++i]);
        } else if (a == "--fp32") {
            precision = 0;
        } else if (a == "--fp64") {
            precision = 1;
        } else if (a == "--json") {
            json = true;
        } else if (!a.empty() && a[0] == '-') {
            std::cerr << "Unknown option " << a << "\n";
            usage();
            return 1;
        } else {
            file = a;
        }
    }
    if (file.empty()) {
        usage();
        return 1;
    }
    std::cout << "\n  QuantKos Monte Carlo engine -- MI355X (gfx950) build\n";
    try {
        Config::UInputs conf = Config::Parser::parse(file);
        if (precision >= 0)
            conf.mc.precision = precision == 0 ? Implemented::Precision::FP32 : Implemented::Precision::FP64;
        if (device >= 0) Engine::set_device(device);
        const Engine::MCResults r = Engine::MCDispatcher(conf).launch_montecarlo();
        if (json) {
            std::printf("{\"price\": %.10g, \"se\": %.10g, \"prob_itm\": %.10g, \"delta\": %.10g, \"gamma\": %.10g, "
                        "\"vega\": %.10g, \"vomma\": %.10g, \"vanna\": %.10g, \"rho\": %.10g, \"theta\": %.10g, "
                        "\"t_total_s\": %.6g, \"t_pricing_s\": %.6g, \"t_greeks_s\": %.6g}\n",
                        r.OptionPrice.option_price, r.OptionPrice.standard_error, r.OptionPrice.prob_itm,
                        r.Greeks.delta, r.Greeks.gamma, r.Greeks.vega, r.Greeks.vomma, r.Greeks.vanna, r.Greeks.rho,
                        r.Greeks.theta, r.Timing.s_elapsed_total, r.Timing.s_elapsed_for_pricing,
                        r.Timing.s_elapsed_for_greeks);
        }
    } catch (const std::exception &e) {
        std::cerr << "\n  FATAL ERROR: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
