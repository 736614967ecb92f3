// qk_solver.hpp -- the reference's batch-level types, on the device windows.
//
// Drop-in counterparts of:
//   RNGManager                  (schemes/RandNGenerator.hpp:40-80)
//   PathsMCBatchMem             (memory/PathsMCBatchMem.hpp:49-380): the batch plan
//   MSolver<M,S,T,R>            (schemes/markovian/MSolver.hpp:166-232)
//   run_forward_all_mc_batches  (engine/MCEngine.hpp:489-520)
//   LSMEngine<R, Basis>         (options/LSMAmerican.hpp:217-614)
// A batch is a window of the path set, priced on the device by
// qk_price_window; the windows' first-stage merge chunks are merged on the
// device by qk_merge_windows when OptionPricer evaluates the price.  The host
// only moves the opaque chunk records: no arithmetic on paths or payoffs.
#pragma once

#include <cstdint>
#include <vector>

#include "qk.h"
#include "qk_config.hpp"
#include "qk_engine.hpp"

namespace quantkos::Engine {

// Throws std::runtime_error with the ctx's message when st is not QK_OK.
void check_status(qk_status st, const char *what);

// The reference seeds a pool of per-thread XorShift states; the counter-based
// stream of the kernels needs no device state (a path's random numbers follow
// from the seed and the path's index), so the manager holds the seed.
class RNGManager {
public:
    explicit RNGManager(const KC::UInputs &config) : seed_(config.mc.rng_seed) {}
    RNGManager(const RNGManager &) = delete;
    RNGManager &operator=(const RNGManager &) = delete;
    RNGManager(RNGManager &&) = default;
    RNGManager &operator=(RNGManager &&) = default;
    uint64_t seed() const { return seed_; }

private:
    uint64_t seed_;
};

// The batch plan of MC.batch_size (PathsMCBatchMem.hpp:330-345: -1 every path
// in one batch, 0 the size the memory budget allows, n > 0 batches of n).  The
// kernels keep no per-path memory, so 0 is one batch as well; a positive size
// is rounded up to a whole number of merge chunks (QK_WINDOW_CHUNK_PATHS), so
// every batch starts on a chunk boundary and the batched price equals the
// one-launch price bit for bit.  The batches' device results accumulate here.
class PathsMCBatchMem {
public:
    explicit PathsMCBatchMem(const KC::UInputs &conf);
    PathsMCBatchMem(const PathsMCBatchMem &) = delete;
    PathsMCBatchMem &operator=(const PathsMCBatchMem &) = delete;
    PathsMCBatchMem(PathsMCBatchMem &&) = default;

    int64_t n_sims_per_batch = 0;
    int64_t n_full_batches() const { return n_paths / n_sims_per_batch; }
    int64_t n_sims_left_over_after_full_batches() const { return n_paths % n_sims_per_batch; }
    int64_t total_batch_loops() const {
        return n_full_batches() + (n_sims_left_over_after_full_batches() > 0 ? 1 : 0);
    }
    int64_t get_curr_batch_size(int64_t batch_idx) const {
        return batch_idx < n_full_batches() ? n_sims_per_batch : n_sims_left_over_after_full_batches();
    }

    // Filled by MSolver::execute_batch: the merge chunks of every batch run so
    // far, in path order, the payoff they were priced with, and the first path
    // of the next batch.
    std::vector<qk_moments> chunks;
    qk_payoff payoff{};
    int64_t next_path0 = 0;

private:
    int64_t n_paths;
};

// Contract of a batch: the configuration's market, time grid, precision and
// seed, with the solver's model and scheme and the batch's path count.
qk_contract batch_contract(const KC::UInputs &config, int model, int scheme, int64_t n_active, uint64_t seed);

// Prices the batch window [BatchMem.next_path0, + n_active) on the device and
// appends its merge chunks to BatchMem (qk_price_window).
void execute_window(const qk_contract &c, const qk_payoff &p, PathsMCBatchMem &BatchMem);

template <KI::MathModel ModelPolicy, KI::NumScheme SchemePolicy, KI::OptType OptType, KI::OptRight OptRight>
class MSolver {
public:
    static_assert(OptType != KI::OptType::American,
                  "the forward solver prices a European-style payoff; American goes through LSMEngine");
    explicit MSolver(const KC::UInputs &config) : config(config) {}
    MSolver(const MSolver &) = delete;
    MSolver &operator=(const MSolver &) = delete;
    MSolver(MSolver &&) = default;

    // One batch of n_active_sims_in_batch paths: the fused path kernel with
    // the template's scheme and payoff over the batch's window of the stream.
    void execute_batch(const int64_t n_active_sims_in_batch, PathsMCBatchMem &BatchMem, const RNGManager &RNGen) const {
        const qk_contract c = batch_contract(config, (int)ModelPolicy, (int)SchemePolicy, n_active_sims_in_batch,
                                             RNGen.seed());
        const qk_payoff p{(int32_t)OptType, OptRight == KI::OptRight::Call ? QK_CALL : QK_PUT,
                          config.options.StrikePrice, config.options.BarrierPrice};
        execute_window(c, p, BatchMem);
    }

private:
    const KC::UInputs &config;
};

// Every batch of the plan through the solver, then the per-batch callback
// (the reference's host copy / accumulation hook).
template <typename SolverType, typename PerBatch>
void run_forward_all_mc_batches(PathsMCBatchMem &BatchMem, const RNGManager &RNGen, const SolverType &Solver,
                                PerBatch &&per_batch) {
    const int64_t loops = BatchMem.total_batch_loops();
    for (int64_t b = 0; b < loops; ++b) {
        const int64_t n = BatchMem.get_curr_batch_size(b);
        Solver.execute_batch(n, BatchMem, RNGen);
        per_batch(b, n);
    }
}

// Longstaff-Schwartz with the basis of the template (LSMAmerican.hpp:217-233;
// the reference's default LaguerreP03).  The whole backward induction runs on
// the device: path store in HBM, per-date regressions and cash-flow updates
// (qk_price_american_ex).
template <KI::OptRight OptRight, KI::LSRegressionBasis Basis = KI::LSRegressionBasis::LaguerreP03>
class LSMEngine {
public:
    static constexpr size_t get_n_basis() {
        if constexpr (Basis == KI::LSRegressionBasis::LaguerreP02) return 3;
        else if constexpr (Basis == KI::LSRegressionBasis::LaguerreP03) return 4;
        else return 5;
    }
    static constexpr size_t NBasis = get_n_basis();

    // Price of the American option of a scaled configuration; coeffs (if not
    // null) receives the regression coefficients [N_time_steps][NBasis].
    static qk_stats price(const KC::UInputs &scaled, bool weighted = false, double *coeffs = nullptr) {
        qk_contract c = to_contract(scaled);
        const qk_lsm_options opt{(int32_t)Basis, weighted ? 1 : 0};
        qk_stats s{};
        check_status(qk_price_american_ex(device_context(), &c, OptRight == KI::OptRight::Call ? QK_CALL : QK_PUT,
                                          scaled.options.StrikePrice, &opt, &s, coeffs),
                     "qk_price_american_ex");
        return s;
    }
};

// The forward pricing of a scaled configuration in the batches of its plan
// (ForwardMCRunner's batch loop, ForwardMC.hpp:113-135): equal to the
// one-launch price bit for bit.
qk_stats price_in_batches(const KC::UInputs &scaled);

}  // namespace quantkos::Engine
