// qk_solver.cpp -- batch plan, batch windows and their device merge.
#include "qk_solver.hpp"

#include <stdexcept>
#include <string>

namespace quantkos::Engine {

void check_status(qk_status st, const char *what) {
    if (st != QK_OK) {
        qk_ctx *ctx = device_context();
        throw std::runtime_error(std::string(what) + " failed: " + qk_strerror(st) + ": " + qk_last_error(ctx));
    }
}

// PathsMCBatchMem.hpp:330-345 (sizes) with the chunk rounding described in
// the header.
PathsMCBatchMem::PathsMCBatchMem(const KC::UInputs &conf) : n_paths(conf.mc.N_Paths) {
    if (n_paths < 1) throw std::runtime_error("Cannot plan batches: Invalid number of paths.");
    const int64_t b = conf.mc.batch_size;
    if (b <= 0) {
        n_sims_per_batch = n_paths;
    } else {
        const int64_t w = QK_WINDOW_CHUNK_PATHS;
        n_sims_per_batch = (b + w - 1) / w * w;
        if (n_sims_per_batch > n_paths) n_sims_per_batch = n_paths;
    }
}

qk_contract batch_contract(const KC::UInputs &config, int model, int scheme, int64_t n_active, uint64_t seed) {
    qk_contract c = to_contract(config);
    c.model = model;
    c.scheme = scheme;
    c.n_paths = n_active;
    c.seed = seed;
    return c;
}

void execute_window(const qk_contract &c, const qk_payoff &p, PathsMCBatchMem &BatchMem) {
    if (!BatchMem.chunks.empty() &&
        (BatchMem.payoff.type != p.type || BatchMem.payoff.right != p.right || BatchMem.payoff.K != p.K ||
         BatchMem.payoff.B != p.B))
        throw std::runtime_error("every batch of a path set must price the same payoff");
    const int64_t nc = (c.n_paths + QK_WINDOW_CHUNK_PATHS - 1) / QK_WINDOW_CHUNK_PATHS;
    const size_t at = BatchMem.chunks.size();
    BatchMem.chunks.resize(at + (size_t)nc);
    check_status(qk_price_window(device_context(), &c, BatchMem.next_path0, &p, 1, BatchMem.chunks.data() + at),
                 "qk_price_window");
    BatchMem.payoff = p;
    BatchMem.next_path0 += c.n_paths;
}

void OptionPricer::accumulate_batch_payoffs(const PathsMCBatchMem &BatchMem, int64_t curr_batch_size) {
    batch_mem = &BatchMem;
    batch_paths += curr_batch_size;
}

void OptionPricer::merge_batches() {
    if (batch_mem == nullptr) return;
    if (batch_paths != config.mc.N_Paths)
        throw std::runtime_error("batches cover " + std::to_string(batch_paths) + " paths of " +
                                 std::to_string(config.mc.N_Paths));
    const qk_contract c = to_contract(config);
    check_status(qk_merge_windows(device_context(), &c, &batch_mem->payoff, 1, batch_mem->chunks.data(),
                                  (int64_t)batch_mem->chunks.size(), &stats),
                 "qk_merge_windows");
    batch_mem = nullptr;
}

qk_stats price_in_batches(const KC::UInputs &scaled) {
    using KI::OptRight;
    RNGManager RNGen(scaled);
    PathsMCBatchMem BatchMem(scaled);
    OptionPricer pricer(scaled);
    const qk_contract c0 = to_contract(scaled);
    const qk_payoff p = to_payoff(scaled);
    // The template policies of the reference's dispatcher are run-time values
    // here; one generic batch body serves every combination.
    struct Solver {
        const KC::UInputs &cfg;
        int model, scheme;
        qk_payoff pay;
        void execute_batch(int64_t n, PathsMCBatchMem &mem, const RNGManager &rng) const {
            execute_window(batch_contract(cfg, model, scheme, n, rng.seed()), pay, mem);
        }
    } solver{scaled, c0.model, c0.scheme, p};
    run_forward_all_mc_batches(BatchMem, RNGen, solver,
                               [&](int64_t, int64_t n) { pricer.accumulate_batch_payoffs(BatchMem, n); });
    pricer.evaluate_option_price();
    return pricer.device_stats();
}

}  // namespace quantkos::Engine
