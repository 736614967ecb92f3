// qk_launch.hpp -- kernel launchers (qk_kernels.hip) called by the C ABI layer.
#pragma once

#include <algorithm>
#include <cmath>

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "qk_types.hpp"

namespace qk {

// Lane indices of the Greek stencil (-1 = not priced) and the real-scale bumps.
struct GreekMap {
    int32_t base, s_up, s_dn, v_up, v_dn, r_up, r_dn, t_dn, t_up, uu, du, ud, dd, n_lanes;
    double h_S_real, h_vol, h_r, h_T;
};

// Longstaff-Schwartz regression of one date with NB basis functions.  The
// left-hand side (the in-the-money count and sum L L^T) depends only on the
// stored spots, so a pre-pass sums and factorises it for every date before the
// backward sweep; the sweep's chain of dates carries only the right-hand side
// sum L Y, which depends on the cash flows.
// Pre-pass partial: count, then the upper triangle of sum L L^T row by row.
template <int NB>
struct LsmGram {
    static constexpr int NM = NB * (NB + 1) / 2;
    static constexpr int NG = 1 + NM;
    static constexpr int STRIDE = 16;                   // doubles per partial (NG <= 16)
    __host__ __device__ static constexpr int m(int i, int j) { return 1 + i * NB - i * (i - 1) / 2 + (j - i); }
};
// Right-hand side partial: NB components padded to a power of two.
template <int NB>
struct LsmRhs {
    static constexpr int STRIDE = NB <= 4 ? 4 : 8;
};
// Factor record of one date (doubles): [0] in-the-money count, [1] 1 when the
// date takes the mean-payoff fallback (LSMAmerican.hpp:536-545, :605-613), then
// the strictly lower Cholesky entries of M + 1e-8 I row by row, then the NB
// reciprocals of its diagonal.
constexpr int LSM_FACTOR_STRIDE = 32;
template <int NB>
struct LsmFactor {
    static constexpr int NL = NB * (NB - 1) / 2;
    static constexpr int NF = 2 + NL + NB;              // entries in use
    __host__ __device__ static constexpr int l(int i, int j) { return 2 + i * (i - 1) / 2 + j; }   // j < i
    __host__ __device__ static constexpr int inv(int i) { return 2 + NL + i; }
};
// Partial of the request server's pump, which accumulates the Gram matrix
// together with the right-hand side while it reads S(t-1) (no second read of
// the path store; its solve block factorises): the pre-pass layout, then
// sum L Y.
template <int NB>
struct LsmSums {
    static constexpr int NG = LsmGram<NB>::NG;
    static constexpr int NS = NG + NB;
    static constexpr int STRIDE = NS <= 16 ? 16 : 32;   // doubles per partial
    __host__ __device__ static constexpr int b(int k) { return NG + k; }
};
constexpr int LSM_MAX_BASIS = 5;
constexpr int LSM_STRIDE = 32;          // partial allocation in doubles (>= every LsmRhs / LsmSums STRIDE)

// Gram pre-pass of one request: blocks of 256 threads over (chunk of paths,
// date); dates are launched in windows of at most lsm_gram_window() dates,
// whose partials [window][gchunks][LsmGram::STRIDE] one k_lsm_factor launch
// sums (chunk order) and factorises into factor[date].
template <typename Real>
struct LsmGramArgs {
    const Real *store;    // [n_steps][n_paths]
    double *gpart;        // [window][gchunks][LsmGram<NB>::STRIDE]
    double *factor;       // [n_steps][LSM_FACTOR_STRIDE]; dates 0 .. n_steps-2 are written
    int64_t n_paths, gchunk;
    int32_t n_steps, right, gchunks, pad;
    Real K, pad2;
};
// Paths per pre-pass chunk and the number of chunks.
int lsm_gram_chunks(int64_t n_paths, int64_t *gchunk);
// Dates per pre-pass launch for the given chunk count.
int lsm_gram_window(int gchunks);
// Device bytes of the pre-pass: the window's partials and the factor records.
size_t lsm_prepass_bytes(int64_t n_paths, int32_t n_steps);
template <typename Real>
hipError_t launch_lsm_gram(const LsmGramArgs<Real> &a, int n_basis, bool weighted, hipStream_t s);

// One request of a Longstaff-Schwartz date step (k_lsm_step).
constexpr int PUMP_MAX_REQ = 24;        // requests per pump launch (kernel-argument size)
constexpr int PUMP_MAX_BLOCKS = 256;    // blocks per request per date (each re-reads all partials)
template <typename Real>
struct PumpReq {
    const Real *store;    // [n_steps][n_paths]
    Real *cf;             // [n_paths]
    double *bpart;        // [2][nb][STRIDE] partials, date d in buffer d & 1: LsmRhs (k_lsm_step) or
                          // LsmSums (the server's k_lsm_solve / k_lsm_paths)
    const double *factor; // [n_steps][LSM_FACTOR_STRIDE] from the Gram pre-pass (k_lsm_step only)
    double *coeffs;       // [n_steps][n_basis]
    double *psum, *pm2;   // tile moments of the final cash flows
    uint32_t *pitm;
    int64_t n_paths, chunk;
    int32_t n_steps, right, nb, t;   // t: the date this step processes
    Real K, disc, itm_eps, pad;
};
template <typename Real>
struct PumpArgs {
    PumpReq<Real> r[PUMP_MAX_REQ];
    int32_t block_start[PUMP_MAX_REQ + 1];
    int32_t n_req;
};
// One pass of the order-statistic selection (k_select_hist): the open key
// prefixes and the bit position of the 8-bit digit counted under each.
constexpr int SELECT_MAX_GROUPS = 16;
struct SelectArgs {
    uint64_t prefix[SELECT_MAX_GROUPS];
    int32_t n_groups, shift;
};
template <typename Real>
hipError_t launch_select_hist(const Real *vals, int64_t n, const SelectArgs &a, uint32_t *hist, int blocks,
                              hipStream_t s);

// Threads per Longstaff-Schwartz block (k_lsm_step, k_lsm_coop): one block
// per chunk of paths and one Gram partial per block.  16 wavefronts (FP32) or
// 8 (FP64, whose date kernel needs up to 256 registers) per block keep 4 or 2
// wavefronts on every SIMD when the 256 partials' blocks cover the GPU.
template <typename Real>
constexpr int lsm_group() { return sizeof(Real) == 4 ? 1024 : 512; }
// Threads per block of the request server's pump, which shares the GPU with
// batch launches on other streams: smaller blocks find room on partly
// occupied CUs (measured: 1,024-thread pump blocks cut the config-4 burst rate
// by 13 %).
constexpr int LSM_SHARED_GROUP = 256;
// One request's whole backward pass in one cooperative launch (k_lsm_coop):
// one block of lsm_group<Real>() threads per chunk.
constexpr unsigned int LSM_COOP_SPIN_LIMIT = 1u << 22;   // polls per wait before giving up
constexpr int LSM_COOP_SYNC_ERR_WORD = 16 * 8 + 32;       // index of the error word in LsmCoopArgs::sync
template <typename Real>
struct LsmCoopArgs {
    const Real *store;    // [n_steps][n_paths]
    Real *cf;             // [n_paths], final cash flows
    double *bpart;        // [2][nb][LsmRhs<NB>::STRIDE], by date parity
    const double *factor; // [n_steps][LSM_FACTOR_STRIDE] from the Gram pre-pass
    double *coeffs;       // [n_steps][n_basis]
    double *psum, *pm2;
    uint32_t *pitm;
    unsigned int *sync;   // 2 KiB of counters, flags and published coefficients; zero before the launch
    int64_t n_paths, chunk;
    int32_t n_steps, right, nb, pad;
    Real K, disc, itm_eps, pad2;
};
template <typename Real>
int lsm_coop_occupancy(int n_basis, bool weighted, size_t lds);
template <typename Real>
hipError_t launch_lsm_coop(const LsmCoopArgs<Real> &a, int n_basis, bool weighted, int blocks, size_t lds,
                           hipStream_t s);

// Blocks and paths per block of one request's pump work (group = threads per
// block; at most max_blocks blocks, each a whole number of 256-path rows).
int pump_blocks(int64_t n_paths, int group, int max_blocks, int64_t *chunk);
// Blocks of a request in the server's two-launch pump step (k_lsm_solve +
// k_lsm_paths): chunks of SERVER_PUMP_CHUNK paths (one pass of 8 rows of 256
// threads), larger when the request would need more than
// SERVER_PUMP_MAX_BLOCKS blocks, so that the one solve block per request
// reads at most 256 partials (32 KiB) per date.  Beside batch kernels about
// one pump block fits per CU, so a large request's date runs in rounds of
// ~256 blocks whatever its block count, and fewer, longer blocks cost the
// date kernel nothing while they cut the solve's read by 4x (4,096-request
// stream, against 1,024: American p50 / p99 9.0-9.9 / 101-113 ms ->
// 6.6-8.6 / 94-100 ms, burst throughput 0.4 % higher).
constexpr int64_t SERVER_PUMP_CHUNK = 2048;
constexpr int64_t SERVER_PUMP_MAX_BLOCKS = 256;
constexpr int LSM_SOLVE_GROUP = 256;    // threads of a k_lsm_solve block
int server_pump_blocks(int64_t n_paths, int64_t *chunk);
template <typename Real>
hipError_t launch_pump_two_phase(const PumpArgs<Real> &a, int n_basis, bool weighted, hipStream_t s);

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
                        int blocks, bool raised, hipStream_t s);
hipError_t launch_batch_merge(const MergeReqD *mr, int n_reqs, const int32_t *chunk_req, int64_t n_chunks,
                              qk_stats *out, hipStream_t s);

// Paths [p0, p0 + n) of a [step][path] store (leading dimension n_paths) into
// out[(i - p0) * n_steps + t]; grid.x = ceil(n / 64) <= 2^32 - 1 blocks.
template <typename Real>
hipError_t launch_transpose_paths(const Real *store, int64_t n_paths, int32_t n_steps, int64_t p0, int64_t n,
                                  Real *out, hipStream_t s);

template <typename Real>
hipError_t launch_paths(const PathArgs<Real> &a, int model, int scheme, bool pathdep, bool store, int L,
                        hipStream_t s);
// First-stage merge chunks of n_tiles tiles (MERGE_MOM_BYTES each per
// (lane, payoff) pair).
int64_t merge_chunks(int64_t n_tiles);
constexpr size_t MERGE_MOM_BYTES = 32;
// Both merge stages and the finalisation into out; with chunks_only, only the
// first stage, whose chunk moments stay in scratch [group][chunk].
hipError_t launch_merge(const double *ps, const double *pm2, const uint32_t *pitm, int64_t n_tiles, int64_t n_paths,
                        int n_lanes, int n_pay, const double *disc, const double *scale, void *scratch, qk_stats *out,
                        hipStream_t s, bool chunks_only = false);
// The second stage and finalisation over chunk moments [group][n_chunks].
hipError_t launch_merge_final(const void *chunks, int64_t n_chunks, int groups, int n_pay, const double *disc,
                              const double *scale, qk_stats *out, hipStream_t s);
hipError_t launch_greeks(const qk_stats *st, int n_pay, const GreekMap &m, qk_greeks *out, hipStream_t s);
// One k_lsm_step launch of lsm_group<Real>() threads per block (the blocks of
// every request must come from pump_blocks with that group).
template <typename Real>
hipError_t launch_pump(const PumpArgs<Real> &a, int n_basis, bool weighted, hipStream_t s);
}  // namespace qk
