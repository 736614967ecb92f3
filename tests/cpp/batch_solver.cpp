// The reference's batch loop written against the drop-in types
// (ForwardMC.hpp:113-135 with MSolver, RNGManager, PathsMCBatchMem and
// OptionPricer; LSMEngine for the American price of the same market).
// Usage: batch_solver <cfg.yaml>; the configuration must be a Heston
// Andersen-QE Asian call.  Prints the batch plan and both prices.
#include <cinttypes>
#include <cstdio>
#include <exception>

#include "qk_config.hpp"
#include "qk_engine.hpp"
#include "qk_solver.hpp"

namespace E = quantkos::Engine;
namespace KI = quantkos::Implemented;

int main(int argc, char **argv) {
    if (argc != 2) return 2;
    try {
        auto conf = quantkos::Config::Parser::parse(argv[1]);
        conf.apply_price_scaling();
        E::RNGManager rng(conf);
        E::PathsMCBatchMem mem(conf);
        E::OptionPricer pricer(conf);
        E::MSolver<KI::MathModel::Heston, KI::NumScheme::AndersonQE, KI::OptType::Asian, KI::OptRight::Call> solver(conf);
        E::run_forward_all_mc_batches(mem, rng, solver,
                                      [&](int64_t, int64_t n) { pricer.accumulate_batch_payoffs(mem, n); });
        pricer.evaluate_option_price();
        const qk_stats &s = pricer.device_stats();
        std::printf("batches %" PRId64 " of %" PRId64 "\n", mem.total_batch_loops(), mem.n_sims_per_batch);
        std::printf("asian %.17g %.17g %" PRId64 " %" PRId64 "\n", s.price, s.se, s.itm, s.n);
        const qk_stats a = E::LSMEngine<KI::OptRight::Put>::price(conf);
        std::printf("american %.17g %.17g nbasis %zu\n", a.price, a.se, E::LSMEngine<KI::OptRight::Put>::NBasis);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}
