// qk_types.hpp -- structures shared by the host launch code and the kernels.
#pragma once

#include <stdint.h>

#include "../../include/qk.h"

namespace qk {

constexpr int SCHEME_EULER = QK_SCHEME_EULER;
constexpr int SCHEME_MILSTEIN = QK_SCHEME_MILSTEIN;
constexpr int SCHEME_QE = QK_SCHEME_QE;
constexpr int MODEL_HESTON = QK_MODEL_HESTON;
constexpr int MODEL_BATES = QK_MODEL_BATES;

constexpr int TILE = 64;          // paths per tile = one wavefront
constexpr int WAVES_PER_BLOCK = 4;
constexpr int BLOCK = TILE * WAVES_PER_BLOCK;
constexpr int MAX_PAYOFFS = 16;
constexpr int MAX_LANES = 64;
constexpr int MERGE_CHUNK = 1024;   // tiles per first-stage merge block
constexpr int64_t MAX_PATH_STEPS = (int64_t)1 << 56;   // n_paths * n_steps of one contract
// Blocks of one k_paths launch: HIP caps a launch at 2^32 - 1 work-items, so
// larger tile counts go in windows of this many blocks (2^24 tiles).
constexpr int64_t PATHS_WINDOW_BLOCKS = (int64_t)1 << 22;

// Per-lane scheme constants, precomputed on the host in FP64 (the reference's
// SDEScheme constructors, MSchemes.hpp:104-122, :199-216, :296-352, :485-491)
// and stored in the kernel's precision.
template <typename Real>
struct LaneC {
    Real S0, v0;
    Real rqdt;                       // (r - q') dt, q' includes the Bates compensator
    Real half_dt, sqrt_dt, rho, rho_c;
    Real omkdt, ktdt, ssdt, qs2dt, iden;     // Euler / implicit Milstein
    Real ekdt, tmf, s2f1, s2f2, K2, K3, K4, A, cv, eps;   // Andersen QE (cv = -K3/2)
    Real eml_dt, mu_j, sigma_j;      // Bates: exp(-lambda dt), jump mean and vol
    uint32_t jump_w;                 // Bates: a step jumps iff its Poisson word > jump_w
};

template <typename Real>
struct PayD {
    int32_t type, right;
    Real K, B;
};

template <typename Real>
struct PathArgs {
    const LaneC<Real> *lanes;   // [n_lanes padded to a multiple of L]
    const PayD<Real> *pays;     // [n_pay]
    int32_t n_lanes, n_pay, n_steps, pad0;
    int64_t n_paths, n_tiles;
    int64_t path0;              // global index of path 0 (random-stream counter offset)
    int64_t tile_off;           // tile of this launch's block 0 (windowed launches)
    uint32_t key0, key1;
    Real itm_eps;
    double *part_sum;           // [(lane*n_pay + j) * n_tiles + tile]
    double *part_m2;
    uint32_t *part_itm;
    Real *store;                // [step][path] or null
    Real *pay_out;              // [path]: payoff 0 of lane 0 per path, or null
};

}  // namespace qk
