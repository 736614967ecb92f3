// qk_launch.hpp -- kernel launchers (qk_kernels.hip) called by the C ABI layer.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "qk_types.hpp"

namespace qk {

// Lane indices of the Greek stencil (-1 = not priced) and the real-scale bumps.
struct GreekMap {
    int32_t base, s_up, s_dn, v_up, v_dn, r_up, r_dn, t_dn, t_up, uu, du, ud, dd, n_lanes;
    double h_S_real, h_vol, h_r, h_T;
};

template <typename Real>
struct LsmArgs {
    const Real *store;     // [n_steps][n_paths]
    Real *cf;              // [n_paths]
    double *coeffs;        // [n_steps][4]
    double *bpart;         // [lsm_blocks(n_paths)][16]
    double *sums_out;      // [n_steps][16] or null
    unsigned int *counter; // unused (kept zero)
    int64_t n_paths;
    int32_t n_steps, right;
    int32_t t_upd, t_gram;
    Real K, disc;
};

// A request of a serving batch (one payoff, or a path store for LSM).
constexpr int32_t REQ_PATHDEP = 1, REQ_STORE = 2;
template <typename Real>
struct ReqD {
    LaneC<Real> lc;
    PayD<Real> pay;
    int32_t model, scheme, n_steps, flags;
    int64_t n_paths, n_tiles;
    uint32_t key0, key1;
    Real itm_eps;
    double *psum, *pm2;  // this request's tile moments [n_tiles]
    uint32_t *pitm;
    Real *store;         // [step][path] for REQ_STORE
};
struct ItemD {
    int32_t req, n_tiles;
    int64_t tile0;
};
struct MergeReqD {
    const double *psum, *pm2;
    const uint32_t *pitm;
    void *scratch;       // [n_chunks] first-stage moments of this request
    int64_t n_tiles, n_paths, first_chunk, n_chunks;
    double factor;       // discount x price scale
};

template <typename Real>
hipError_t launch_batch(const ReqD<Real> *reqs, const ItemD *items, int32_t n_items, unsigned int *counter,
                        int blocks, hipStream_t s);
hipError_t launch_batch_merge(const MergeReqD *mr, int n_reqs, const int32_t *chunk_req, int64_t n_chunks,
                              qk_stats *out, hipStream_t s);

template <typename Real>
hipError_t launch_paths(const PathArgs<Real> &a, int model, int scheme, bool pathdep, bool store, int L,
                        hipStream_t s);
// Bytes of merge scratch for n_tiles tiles and `groups` (lane, payoff) pairs.
int64_t merge_chunks(int64_t n_tiles);
constexpr size_t MERGE_MOM_BYTES = 32;
hipError_t launch_merge(const double *ps, const double *pm2, const uint32_t *pitm, int64_t n_tiles, int64_t n_paths,
                        int n_lanes, int n_pay, const double *disc, const double *scale, void *scratch, qk_stats *out,
                        hipStream_t s);
hipError_t launch_greeks(const qk_stats *st, int n_pay, const GreekMap &m, qk_greeks *out, hipStream_t s);
// Blocks of one LSM sweep: one path per thread up to LSM_MAX_BLOCKS blocks,
// then grid-stride.
constexpr int LSM_MAX_BLOCKS = 2048;
int lsm_blocks(int64_t n_paths);
template <typename Real>
hipError_t launch_lsm_sweep(const LsmArgs<Real> &a, hipStream_t s);
template <typename Real>
hipError_t launch_lsm_final(const LsmArgs<Real> &a, double *ps, double *pm2, uint32_t *pitm, int64_t n_tiles,
                            Real itm_eps, hipStream_t s);

}  // namespace qk
