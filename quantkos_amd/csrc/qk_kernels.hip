// qk_kernels.hip -- hand-written gfx950 kernels of the Monte Carlo hot path.
//
//  k_paths      fused path generation: Philox normals -> Heston/Bates step ->
//               payoff tracking -> per-tile (64-path) moments of every payoff.
//               Replaces MarkovianIntegrationKernel (MSolver.hpp:98-130) and the
//               two reductions of OptionPricer::accumulate_batch_payoffs
//               (OptionPricer.hpp:129-146).  L lanes (bumped contracts) per
//               thread share each step's Philox words (common random numbers);
//               Heston Euler/Milstein take two steps from one block, Heston QE
//               four steps from three.
//  k_merge      fixed-order Chan merge of tile moments + FP64 finalisation
//               (OptionPricer.hpp:152-172, :204-249).
//  k_greeks     FP64 central differences of lane prices (MCEngine.hpp:586-730).
//  k_lsm_gram   pre-pass of a Longstaff-Schwartz request: every date's
//  k_lsm_factor in-the-money count and Laguerre Gram matrix, summed in a fixed
//               order and Cholesky-factorised in FP64 (LSMAmerican.hpp:433-473,
//               :516-581).
//  k_lsm_step   one Longstaff-Schwartz date for every listed request: solve of
//               the date's normal equations with its factor, cash-flow update,
//               right-hand side of the previous date, and at date 0 the final
//               discount and tile moments (LSMAmerican.hpp:310-341, :393-419,
//               :433-473, :484-502, :605-613).
//  k_lsm_coop   the same dates of one request in one cooperative launch.
//  k_lsm_solve  the request server's pump step (Gram and right-hand side
//  k_lsm_paths  accumulated together, factorised per date by one block).
//  k_select_hist  order statistics of per-path payoffs (distribution).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "qk_device.hpp"
#include "qk_launch.hpp"

namespace qk {

// ---------------------------------------------------------------------------
// Fused path kernel.  One wavefront = one tile of 64 consecutive paths, one
// path per lane; blockIdx.y selects a group of L lanes.  Tile moments are
// stored per (lane, payoff, tile) so the merge order is fixed by the tile
// index: results do not depend on the grid shape or on other work in flight.
// ---------------------------------------------------------------------------
// One tile (64 paths, one per lane) for L lanes sharing each step's Philox
// block.  Lane l's moments for payoff j go to part_*[l*lane_stride + j*pay_stride
// + out_index]; path p of the tile draws counter (path0 + p, step, slot).
template <typename Real, int MODEL, int SCHEME, bool PATHDEP, bool STORE, int L, bool HIU = true>
__device__ __forceinline__ void run_tile(const LaneC<Real> *__restrict__ lc, int n_valid, const PayD<Real> *__restrict__ pays,
                                         int n_pay, int n_steps, int64_t n_paths, int64_t path0, uint32_t key0,
                                         uint32_t key1, Real itm_eps, int64_t tile, Real *__restrict__ store,
                                         double *__restrict__ part_sum, double *__restrict__ part_m2,
                                         uint32_t *__restrict__ part_itm, int64_t out_index, int64_t lane_stride,
                                         int64_t pay_stride, Real *__restrict__ pay_out = nullptr) {
    using M = Math<Real>;
    const int lid = threadIdx.x & (TILE - 1);
    const int64_t path = tile * TILE + lid;
    const bool active = path < n_paths;
    const uint64_t gpath = (uint64_t)(path0 + path);     // counter of the random stream
    // HIU: the caller guarantees path0 % 64 == 0, so the tile's 64 paths share
    // the high counter word; held in a scalar register it makes the first two
    // Philox rounds' products uniform or loop-invariant (scalar unit, hoisted)
    // instead of per-lane work at every step.
    const uint32_t p0 = (uint32_t)gpath;
    const uint32_t p1 = HIU ? __builtin_amdgcn_readfirstlane((uint32_t)(gpath >> 32)) : (uint32_t)(gpath >> 32);
    constexpr int PH = HIU ? 2 : 0;

    // Lane constants held in registers for the whole path loop (scalar
    // registers: the lane block is uniform per wavefront).
    LaneC<Real> cl[L];
#pragma unroll
    for (int l = 0; l < L; ++l) {
        cl[l] = lc[l];
        // The QE moments m = v e^{-k dt} + tmf and s^2 = v s2f1 + s2f2 take two
        // uniform operands; with the addends held in vector registers the
        // step needs no copy of them (one scalar operand per instruction).
        if constexpr (SCHEME == SCHEME_QE) __asm__("" : "+v"(cl[l].tmf), "+v"(cl[l].s2f2));
        // Eight lanes' constants exceed the scalar register file: the rest
        // are spilled to vector-register lanes and read back with a
        // v_readlane per use.  Holding these in vector registers instead
        // (measured, 13 Greek lanes, 2M x 252): Milstein 3.73 -> 3.27 ms,
        // Euler 3.06 -> 2.96 ms.
        if constexpr (SCHEME == SCHEME_MILSTEIN && L >= 8) __asm__("" : "+v"(cl[l].iden), "+v"(cl[l].qs2dt));
        if constexpr (SCHEME == SCHEME_EULER && L >= 8)
            __asm__("" : "+v"(cl[l].rho), "+v"(cl[l].rho_c), "+v"(cl[l].sqrt_dt), "+v"(cl[l].half_dt));
    }
    Real S[L], v[L], mx[L], mn[L];
    double sm[L];
#pragma unroll
    for (int l = 0; l < L; ++l) {
        S[l] = PATHDEP ? cl[l].S0 : Real(0);     // log-increment accumulator when !PATHDEP
        v[l] = cl[l].v0;
        mx[l] = mn[l] = cl[l].S0;
        sm[l] = 0.0;
    }

    // One step from the words (a0, a1) for the Box-Muller pair, a2 for the QE
    // exponential branch and a3 for the first Bates Poisson uniform.
    auto step = [&](int t, uint32_t a0, uint32_t a1, uint32_t a2, uint32_t a3) {
        const Real u0 = M::uniform(a0), u1 = M::uniform(a1);
        Real sn, cs;
        M::sincos2pi(u1, sn, cs);
        const Real R = M::sqrt(M::neg2log(u0));
        const Real z1 = R * cs, z2 = R * sn;
#pragma unroll
        for (int l = 0; l < L; ++l) {
            Real inc = heston_step<Real, SCHEME>(cl[l], z1, z2, a2, v[l]);
            if constexpr (MODEL == MODEL_BATES) {
                // U(a3) >= e^{-lambda dt} as a compare of the word itself
                if (a3 > cl[l].jump_w)
                    inc += bates_jump_slow<Real>(cl[l], M::uniform(a3), p0, p1, (uint32_t)t, key0, key1);
            }
            if constexpr (PATHDEP) {
                S[l] *= M::exp(inc);
                sm[l] += (double)S[l];
                mx[l] = M::track_max(mx[l], S[l]);
                mn[l] = M::track_min(mn[l], S[l]);
            } else {
                S[l] += inc;
            }
            if constexpr (STORE) {
                if (active) store[(int64_t)t * n_paths + path] = S[l];
            }
        }
    };
    if constexpr (MODEL == MODEL_HESTON && (SCHEME == SCHEME_EULER || SCHEME == SCHEME_MILSTEIN)) {
        // Two words per step: steps 2i and 2i+1 take words (0, 1) and (2, 3)
        // of block (path, i, slot 0).
        int t = 0;
        for (; t + 1 < n_steps; t += 2) {
            const Words w = philox<PH>(p0, p1, (uint32_t)t >> 1, 0u, key0, key1);
            step(t, w.w0, w.w1, 0u, 0u);
            step(t + 1, w.w2, w.w3, 0u, 0u);
        }
        if (t < n_steps) {
            const Words w = philox<PH>(p0, p1, (uint32_t)t >> 1, 0u, key0, key1);
            step(t, w.w0, w.w1, 0u, 0u);
        }
    } else if constexpr (MODEL == MODEL_HESTON && SCHEME == SCHEME_QE) {
        // Three words per step, taken in order from the blocks (path, b, 0),
        // b = 0, 1, ...: step t uses words 3t, 3t+1, 3t+2 of that sequence
        // (word k = word k mod 4 of block k / 4), so steps 4i..4i+3 share
        // blocks 3i..3i+2.
        if constexpr (sizeof(Real) == 8) {
            // FP64: one step per iteration (four steps unrolled exceed the
            // register budget), carrying the block that holds word 3t.
            Words cur{0u, 0u, 0u, 0u};
            for (int t = 0; t < n_steps; ++t) {
                const uint32_t k = 3u * (uint32_t)t, r = k & 3u;
                uint32_t a0, a1, a2;
                if (r == 0u) {
                    cur = philox<PH>(p0, p1, k >> 2, 0u, key0, key1);
                    a0 = cur.w0; a1 = cur.w1; a2 = cur.w2;
                } else if (r == 1u) {
                    a0 = cur.w1; a1 = cur.w2; a2 = cur.w3;
                } else {
                    const Words nx = philox<PH>(p0, p1, (k >> 2) + 1u, 0u, key0, key1);
                    if (r == 3u) {
                        a0 = cur.w3; a1 = nx.w0; a2 = nx.w1;
                    } else {
                        a0 = cur.w2; a1 = cur.w3; a2 = nx.w0;
                    }
                    cur = nx;
                }
                step(t, a0, a1, a2, 0u);
            }
        } else {
            int t = 0;
            uint32_t b = 0;
            for (; t + 3 < n_steps; t += 4, b += 3) {
                const Words x = philox<PH>(p0, p1, b, 0u, key0, key1);
                step(t, x.w0, x.w1, x.w2, 0u);
                const Words y = philox<PH>(p0, p1, b + 1, 0u, key0, key1);
                step(t + 1, x.w3, y.w0, y.w1, 0u);
                const Words z = philox<PH>(p0, p1, b + 2, 0u, key0, key1);
                step(t + 2, y.w2, y.w3, z.w0, 0u);
                step(t + 3, z.w1, z.w2, z.w3, 0u);
            }
            for (; t < n_steps; ++t) {      // at most 3 steps; word 3t is word 0, 3 or 2 of its block
                const uint32_t k = 3u * (uint32_t)t;
                const Words x = philox<PH>(p0, p1, k >> 2, 0u, key0, key1);
                uint32_t a0 = x.w0, a1 = x.w1, a2 = x.w2;
                if ((k & 3u) != 0u) {
                    const Words y = philox<PH>(p0, p1, (k >> 2) + 1u, 0u, key0, key1);
                    if ((k & 3u) == 3u) {
                        a0 = x.w3; a1 = y.w0; a2 = y.w1;
                    } else {
                        a0 = x.w2; a1 = x.w3; a2 = y.w0;
                    }
                }
                step(t, a0, a1, a2, 0u);
            }
        }
    } else {
        for (int t = 0; t < n_steps; ++t) {
            const Words w = philox<PH>(p0, p1, (uint32_t)t, 0u, key0, key1);
            step(t, w.w0, w.w1, w.w2, w.w3);
        }
    }

    const unsigned long long amask = __ballot(active);
    const double n_act = (double)__popcll(amask);
#pragma unroll
    for (int l = 0; l < L; ++l) {
        if (l >= n_valid) break;
        const Real ST = PATHDEP ? S[l] : cl[l].S0 * M::exp(S[l]);
        for (int j = 0; j < n_pay; ++j) {
            const Real x = active ? payoff_eval<Real>(pays[j], ST, sm[l], mx[l], mn[l], n_steps) : Real(0);
            if (pay_out != nullptr && l == 0 && j == 0 && active) pay_out[path] = x;
            const double xd = (double)x;
            const double s = wave_sum(xd);
            const double mean = s / n_act;
            const double d = active ? xd - mean : 0.0;
            const double m2 = wave_sum(d * d);
            const unsigned long long im = __ballot(active && x > itm_eps);
            if (lid == 0) {
                const int64_t idx = l * lane_stride + j * pay_stride + out_index;
                part_sum[idx] = s;
                part_m2[idx] = m2;
                part_itm[idx] = (uint32_t)__popcll(im);
            }
        }
    }
}

// One wavefront = one tile of 64 consecutive paths; blockIdx.y selects a group
// of L lanes.  Tile moments are stored per (lane, payoff, tile) so the merge
// order is fixed by the tile index: results do not depend on the grid shape.
template <typename Real, int MODEL, int SCHEME, bool PATHDEP, bool STORE, int L>
__global__ __launch_bounds__(BLOCK) void k_paths(PathArgs<Real> a) {
    const int64_t tile = a.tile_off + (int64_t)blockIdx.x * WAVES_PER_BLOCK + (threadIdx.x >> 6);
    if (tile >= a.n_tiles) return;                       // uniform per wavefront
    const int lbase = blockIdx.y * L;
    // STORE kernels also serve windows of the stream at any path0
    run_tile<Real, MODEL, SCHEME, PATHDEP, STORE, L, !STORE>(
        a.lanes + lbase, a.n_lanes - lbase, a.pays, a.n_pay, a.n_steps, a.n_paths, a.path0, a.key0, a.key1,
        a.itm_eps, tile, a.store, a.part_sum, a.part_m2, a.part_itm, (int64_t)lbase * a.n_pay * a.n_tiles + tile,
        (int64_t)a.n_pay * a.n_tiles, a.n_tiles, lbase == 0 ? a.pay_out : nullptr);
}

// ---------------------------------------------------------------------------
// Fixed-order merge of tile moments and FP64 finalisation, in two stages.
// Stage 1: block (chunk, group) merges MERGE_CHUNK consecutive tiles; thread i
// merges tiles [i*4, i*4+4) of the chunk, then a fixed LDS tree.  Stage 2: one
// block per group merges the chunk results the same way.  The partition depends
// only on the tile count, so the result is independent of the launch shape.
// ---------------------------------------------------------------------------
struct Mom { double n, mean, m2; unsigned long long itm; };

__device__ __forceinline__ Mom block_tree_merge(Mom m) {
    __shared__ double sn[256], smu[256], sq[256];
    __shared__ unsigned long long si[256];
    const int t = threadIdx.x;
    sn[t] = m.n; smu[t] = m.mean; sq[t] = m.m2; si[t] = m.itm;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s) {
            double na = sn[t], ma = smu[t], qa = sq[t];
            chan_merge(na, ma, qa, sn[t + s], smu[t + s], sq[t + s]);
            sn[t] = na; smu[t] = ma; sq[t] = qa;
            si[t] += si[t + s];
        }
        __syncthreads();
    }
    return Mom{sn[0], smu[0], sq[0], si[0]};
}

// With one chunk per group (n_tiles <= MERGE_CHUNK) the stage-1 block holds
// the group's whole merge and finalises it (final != null): the stage-2 merge
// of one chunk would copy it unchanged.
__device__ __forceinline__ void finalize_group(const Mom &m, int g, int n_pay, const double *__restrict__ disc,
                                               const double *__restrict__ scale, qk_stats *__restrict__ out);

__global__ __launch_bounds__(256) void k_merge_chunks(const double *__restrict__ ps, const double *__restrict__ pm2,
                                                      const uint32_t *__restrict__ pitm, int64_t n_tiles,
                                                      int64_t n_paths, int64_t n_chunks, Mom *__restrict__ out,
                                                      int n_pay, const double *__restrict__ disc,
                                                      const double *__restrict__ scale, qk_stats *__restrict__ final) {
    const int64_t g = blockIdx.y, chunk = blockIdx.x;
    const int64_t base = g * n_tiles;
    Mom m{0.0, 0.0, 0.0, 0ull};
    const int64_t t0 = chunk * MERGE_CHUNK + (int64_t)threadIdx.x * (MERGE_CHUNK / 256);
#pragma unroll
    for (int k = 0; k < MERGE_CHUNK / 256; ++k) {
        const int64_t t = t0 + k;
        if (t < n_tiles) {
            const int64_t rem = n_paths - t * TILE;
            const double nt = (double)(rem < TILE ? rem : TILE);
            chan_merge(m.n, m.mean, m.m2, nt, ps[base + t] / nt, pm2[base + t]);
            m.itm += pitm[base + t];
        }
    }
    m = block_tree_merge(m);
    if (threadIdx.x == 0) {
        if (final) finalize_group(m, (int)g, n_pay, disc, scale, final);
        else out[g * n_chunks + chunk] = m;
    }
}

// Price, SE and ITM share of group g = lane * n_pay + payoff from its merged
// moments (OptionPricer.hpp:204-249).
__device__ __forceinline__ void finalize_group(const Mom &m, int g, int n_pay, const double *__restrict__ disc,
                                               const double *__restrict__ scale, qk_stats *__restrict__ out) {
    const int lane = g / n_pay, j = g % n_pay;
    const double N = m.n;
    const double var = N > 1.0 ? m.m2 / (N - 1.0) : 0.0;
    const double f = disc[lane] * scale[j];
    qk_stats o;
    o.price = m.mean * f;
    o.se = sqrt(var / N) * f;
    o.prob_itm = (double)m.itm / N;
    o.mean = m.mean;
    o.m2 = m.m2;
    o.n = (int64_t)N;
    o.itm = (int64_t)m.itm;
    out[g] = o;
}

__global__ __launch_bounds__(256) void k_merge_final(const Mom *__restrict__ chunks, int64_t n_chunks, int n_pay,
                                                     const double *__restrict__ disc, const double *__restrict__ scale,
                                                     qk_stats *__restrict__ out) {
    const int g = blockIdx.x;
    Mom m{0.0, 0.0, 0.0, 0ull};
    for (int64_t c = threadIdx.x; c < n_chunks; c += 256) {
        const Mom x = chunks[(int64_t)g * n_chunks + c];
        chan_merge(m.n, m.mean, m.m2, x.n, x.mean, x.m2);
        m.itm += x.itm;
    }
    m = block_tree_merge(m);
    if (threadIdx.x == 0) finalize_group(m, g, n_pay, disc, scale, out);
}

// ---------------------------------------------------------------------------
// Greeks: FP64 stencils of MCEngine.hpp:586-730 over lane prices.
// ---------------------------------------------------------------------------
__global__ void k_greeks(const qk_stats *__restrict__ st, int n_pay, GreekMap m, qk_greeks *__restrict__ out) {
    const int j = threadIdx.x;
    if (j >= n_pay) return;
    auto P = [&](int lane) { return st[lane * n_pay + j].price; };
    qk_greeks g = {};
    const double P0 = P(m.base);
    g.base_price = P0;
    g.base_se = st[m.base * n_pay + j].se;
    g.n_lanes = m.n_lanes;
    if (m.s_up >= 0) {
        const double pu = P(m.s_up), pd = P(m.s_dn);
        g.delta = (pu - pd) / (2.0 * m.h_S_real);
        g.gamma = (pu - 2.0 * P0 + pd) / (m.h_S_real * m.h_S_real);
    }
    if (m.v_up >= 0) {
        const double pu = P(m.v_up), pd = P(m.v_dn);
        g.vega = (pu - pd) / (2.0 * m.h_vol);
        g.vomma = (pu - 2.0 * P0 + pd) / (m.h_vol * m.h_vol);
    }
    if (m.r_up >= 0) g.rho = (P(m.r_up) - P(m.r_dn)) / (2.0 * m.h_r);
    if (m.t_dn >= 0) {
        if (m.t_up >= 0) g.theta = (P(m.t_dn) - P(m.t_up)) / (2.0 * m.h_T);
        else g.theta = (P(m.t_dn) - P0) / m.h_T;
    }
    if (m.uu >= 0) g.vanna = (P(m.uu) - P(m.du) - P(m.ud) + P(m.dd)) / (4.0 * m.h_S_real * m.h_vol);
    out[j] = g;
}

// ---------------------------------------------------------------------------
// Longstaff-Schwartz.
// ---------------------------------------------------------------------------
// Regression basis at X = S/K: Laguerre polynomials L0..L(NB-1) in Horner form
// (LSMAmerican.hpp:172-198; NB = 3, 4, 5 are the reference's LaguerreP02/P03/
// P04), times exp(-X/2) for the weighted Laguerre functions of Longstaff and
// Schwartz (2001) when W.
template <typename Real, int NB, bool W>
__device__ __forceinline__ void lsm_basis(Real X, Real L[NB]) {
    L[0] = Real(1);
    if constexpr (NB > 1) L[1] = Real(1) - X;
    if constexpr (NB > 2) L[2] = ((Real(0.5) * X - Real(2)) * X) + Real(1);
    if constexpr (NB > 3) L[3] = (((Real(-1.0 / 6.0) * X + Real(1.5)) * X - Real(3)) * X) + Real(1);
    if constexpr (NB > 4)
        L[4] = ((((Real(1.0 / 24.0) * X - Real(2.0 / 3.0)) * X + Real(3)) * X - Real(4)) * X) + Real(1);
    if constexpr (W) {
        const Real w = Math<Real>::exp(Real(-0.5) * X);
#pragma unroll
        for (int k = 0; k < NB; ++k) L[k] *= w;
    }
}

// Cholesky factorisation of one date's Gram matrix with the reference's 1e-8
// ridge, 1e-14 pivot limit and mean-payoff fallback (LSMAmerican.hpp:516-581,
// :605-613).  s is laid out as LsmGram<NB> {count, upper triangle of M row by
// row}; F receives the factor record (LsmFactor<NB>): count, fallback flag,
// the strictly lower entries and the reciprocal diagonal.  Fallback dates are
// those with fewer than NB in-the-money paths or a pivot below 1e-14.
template <int NB>
__device__ void lsm_factor(const double *s, double F[LSM_FACTOR_STRIDE]) {
    using S = LsmGram<NB>;
    using FR = LsmFactor<NB>;
    const double count = s[0];
#pragma unroll
    for (int k = 0; k < LSM_FACTOR_STRIDE; ++k) F[k] = 0.0;
    F[0] = count;
    if (count < (double)NB) {
        F[1] = 1.0;
        return;
    }
    double Lm[NB][NB] = {};
    double inv[NB];
#pragma unroll
    for (int i = 0; i < NB; ++i) {
#pragma unroll
        for (int j = 0; j <= i; ++j) {
            double acc = 0.0;
#pragma unroll
            for (int k = 0; k < j; ++k) acc += Lm[i][k] * Lm[j][k];
            const double Mij = s[S::m(j, i)] + (i == j ? 1e-8 : 0.0);
            if (i == j) {
                const double val = Mij - acc;
                if (val < 1e-14) {
                    F[1] = 1.0;
                    return;
                }
                Lm[i][i] = Math<double>::sqrt(val);
                inv[i] = Math<double>::rcp(Lm[i][i]);
            } else {
                Lm[i][j] = (Mij - acc) * inv[j];
            }
        }
    }
#pragma unroll
    for (int i = 0; i < NB; ++i) {
#pragma unroll
        for (int j = 0; j < i; ++j) F[FR::l(i, j)] = Lm[i][j];
        F[FR::inv(i)] = inv[i];
    }
}

// Coefficients of one date from its factor record F and right-hand side B:
// forward and back substitution with the factor, or the fallback C0 = B0 /
// count (LSMAmerican.hpp:605-613; all zero without in-the-money paths).
template <int NB>
__device__ __forceinline__ void lsm_solve_factored(const double *F, const double *B, double C[NB]) {
    using FR = LsmFactor<NB>;
#pragma unroll
    for (int k = 0; k < NB; ++k) C[k] = 0.0;
    if (F[1] != 0.0) {
        if (F[0] > 0.0) C[0] = B[0] / F[0];
        return;
    }
    double y[NB];
#pragma unroll
    for (int i = 0; i < NB; ++i) {
        double acc = 0.0;
#pragma unroll
        for (int k = 0; k < i; ++k) acc += F[FR::l(i, k)] * y[k];
        y[i] = (B[i] - acc) * F[FR::inv(i)];
    }
#pragma unroll
    for (int i = NB - 1; i >= 0; --i) {
        double acc = 0.0;
#pragma unroll
        for (int k = i + 1; k < NB; ++k) acc += F[FR::l(k, i)] * C[k];
        C[i] = (y[i] - acc) * F[FR::inv(i)];
    }
}

// Wavefront sums of KP (4, 8, 16 or 32) components held by every lane, by a
// reduce-scatter butterfly: at each level every lane keeps half of its values
// and adds the partner's copy of them, so KP components cost KP - 1
// exchanges (plus 6 - log2 KP full steps) instead of 6 per component.  The
// partners differ in lane bit 5, 4 (permlane swaps, which need no select), 3,
// 2 (row mirrors) and 1, 0 (quad permutations); those masks span all six
// bits, so every sum covers the 64 lanes once.  Lane l returns component
// l >> (6 - log2 KP) summed over the wavefront (the same bits in every lane
// of its group).  v is clobbered.
template <int KP>
__device__ __forceinline__ double wave_scatter_sum(double (&v)[KP], int lid) {
    static_assert(KP == 4 || KP == 8 || KP == 16 || KP == 32, "component count");
#pragma unroll
    for (int k = 0; k < KP / 2; ++k) {            // bit 5: lanes >= 32 keep k + KP/2
        swap_halves_f64<false>(v[k], v[k + KP / 2]);
        v[k] += v[k + KP / 2];
    }
#pragma unroll
    for (int k = 0; k < KP / 4; ++k) {            // bit 4
        swap_halves_f64<true>(v[k], v[k + KP / 4]);
        v[k] += v[k + KP / 4];
    }
    // bits 3, 2, 1: the lane with the bit set keeps the upper half and sends
    // the lower one.
    auto level = [&](auto ctrl, int h, int bit) {
        const bool up = (lid & bit) != 0;
#pragma unroll
        for (int k = 0; k < h; ++k) {
            const double keep = up ? v[k + h] : v[k];
            const double send = up ? v[k] : v[k + h];
            v[k] = keep + dpp_f64<decltype(ctrl)::value>(send);
        }
    };
    if constexpr (KP >= 8) level(std::integral_constant<int, DPP_MIRROR>{}, KP / 8, 8);
    if constexpr (KP >= 16) level(std::integral_constant<int, DPP_HALF_MIRROR>{}, KP / 16, 4);
    if constexpr (KP >= 32) level(std::integral_constant<int, DPP_QP_XOR2>{}, 1, 2);
    double sx = v[0];
    if constexpr (KP < 8) sx += dpp_f64<DPP_MIRROR>(sx);
    if constexpr (KP < 16) sx += dpp_f64<DPP_HALF_MIRROR>(sx);
    if constexpr (KP < 32) sx += dpp_f64<DPP_QP_XOR2>(sx);
    sx += dpp_f64<DPP_QP_XOR1>(sx);
    return sx;
}
template <int KP>
constexpr int scatter_shift() { return KP == 4 ? 4 : KP == 8 ? 3 : KP == 16 ? 2 : 1; }

// Partial layout of the date step: the right-hand side only (GRAM = false,
// the Gram matrix comes from the pre-pass) or Gram and right-hand side
// together (GRAM = true, the request server's pump).
template <int NB, bool GRAM>
constexpr int lsm_part_stride() { return GRAM ? LsmSums<NB>::STRIDE : LsmRhs<NB>::STRIDE; }

// Sums the block partials [nb][lsm_part_stride<NB, GRAM>()] of one date in a
// fixed order (per thread, then across the wavefront, then the wavefronts in
// order).  Every thread of the
// block must call it (it holds one block barrier); every wavefront returns
// the NV = lsm_part_count<NB, GRAM>() sums in sm, the same bits in all (each
// wavefront sums the wavefront totals itself, so no second barrier is needed
// to share them).
template <int NB, bool GRAM>
constexpr int lsm_part_count() { return GRAM ? LsmSums<NB>::NS : NB; }
template <int NB, int G, bool GRAM>
__device__ __forceinline__ void lsm_reduce_parts(const double *__restrict__ bpart, int nb, double *red,
                                                 double sm[lsm_part_count<NB, GRAM>()]) {
    constexpr int ST = lsm_part_stride<NB, GRAM>();
    const int lt = threadIdx.x & (G - 1), lid = lt & 63, wv = lt >> 6;
    if constexpr (ST >= 16) {
        // Wide partials as one coalesced stream: thread i adds elements i,
        // i + G, ... of [nb][ST] (ST divides G, so always component i % ST),
        // then lanes l, l + ST, ... of the wavefront are added by one or two
        // exchanges across lane bits 4 and 5.
        constexpr int UR = 16;   // loads in flight per thread, combined in a fixed order
        double acc[UR];
#pragma unroll
        for (int u = 0; u < UR; ++u) acc[u] = 0.0;
        const int64_t total = (int64_t)nb * ST;
        for (int64_t e0 = lt; e0 < total; e0 += (int64_t)G * UR) {
#pragma unroll
            for (int u = 0; u < UR; ++u) {
                const int64_t e = e0 + (int64_t)u * G;
                if (e < total) acc[u] += bpart[e];
            }
        }
        double tot = 0.0;
#pragma unroll
        for (int u = 0; u < UR; ++u) tot += acc[u];
        if constexpr (ST == 16) tot = add_across<true>(tot);
        tot = add_across<false>(tot);
        if (lid < ST) red[wv * ST + lid] = tot;
    } else {
        // Narrow partials: thread i adds whole partials i, i + G, ..., then
        // the wavefront reduce-scatter.
        double v[ST];
#pragma unroll
        for (int k = 0; k < ST; ++k) v[k] = 0.0;
        for (int p = lt; p < nb; p += G) {
#pragma unroll
            for (int k = 0; k < ST; ++k) v[k] += bpart[(int64_t)p * ST + k];
        }
        const double sx = wave_scatter_sum<ST>(v, lid);
        constexpr int SH = scatter_shift<ST>();
        if ((lid & ((1 << SH) - 1)) == 0) red[wv * ST + (lid >> SH)] = sx;
    }
    __syncthreads();
    double b = 0.0;
    if (lid < ST)
        for (int w = 0; w < G / 64; ++w) b += red[w * ST + lid];
    const uint64_t bb = (uint64_t)__double_as_longlong(b);
#pragma unroll
    for (int k = 0; k < lsm_part_count<NB, GRAM>(); ++k) {   // lane k's component into scalar registers
        const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)bb, k);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(bb >> 32), k);
        sm[k] = __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
    }
}

// The right-hand-side partials of one date summed as above and solved with
// the date's pre-computed factor record F; every thread returns the
// coefficients in C, the same bits in all.
template <int NB, int G>
__device__ __forceinline__ void lsm_reduce_solve(const double *__restrict__ bpart, int nb,
                                                 const double *__restrict__ F, double *red, double C[NB]) {
    constexpr int NF = LsmFactor<NB>::NF;
    double Fr[NF];   // uniform: loaded while the partials are in flight
#pragma unroll
    for (int k = 0; k < NF; ++k) Fr[k] = F[k];
    double sm[NB];
    lsm_reduce_parts<NB, G, false>(bpart, nb, red, sm);
    lsm_solve_factored<NB>(Fr, sm, C);
}

// One Longstaff-Schwartz date t over the paths [p0, p1) by a group of G
// threads (threadIdx.x & (G - 1) is the thread's place in the group): (1) updates
// the cash flows at date t with the coefficients C of date t (initialises them
// at the last date); (2) for t >= 1 accumulates the in-the-money right-hand
// side sum L(S(t-1)) Y of date t-1 in FP64 (Y the discounted new cash flow;
// the Gram matrix comes from the pre-pass) and stores the group's sums in
// part[0 .. lsm_part_stride<NB, GRAM>()) (nothing if part is null); with GRAM
// the in-the-money count and sum L L^T of date t-1 are accumulated too (the
// request server's pump, which has no pre-pass); (3) at date 0 applies the
// final one-step discount and
// writes the per-tile moments of the cash flows.  cfp[i - p0] is the cash flow
// of path i (device memory, or LDS when CF_LDS).  Holds a block barrier when t >= 1: every
// thread of the block must call it for the same t.
template <typename Real, int NB, bool W, bool CF_LDS, int G, int U, bool GRAM, typename Pre>
__device__ __forceinline__ void lsm_chunk_date(const Real *__restrict__ store, Real *cfp, int64_t n, int64_t p0,
                                               int64_t p1, int t, int N, int right, Real K, Real disc, Real itm_eps,
                                               double C[NB], double *part, double *psum, double *pm2,
                                               uint32_t *pitm, double (*wpart)[lsm_part_stride<NB, GRAM>()],
                                               Pre &&pre) {
    constexpr int KP = lsm_part_stride<NB, GRAM>(), NV = lsm_part_count<NB, GRAM>();
    const int lt = threadIdx.x & (G - 1), lid = lt & 63, wv = lt >> 6;
    double g[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) g[k] = 0.0;
    // U rows of G paths per pass; a pass's loads are issued together, and
    // the next pass's loads are issued before this pass is computed (they are
    // other paths, so the cash-flow writes of this pass do not touch them).
    // Thread lt meets paths lt, lt + G, lt + 2G, ... in that order whatever U
    // is, so U changes no result.
    // Paths are addressed by 32-bit offsets o = i - p0 through buffer
    // resources over [p0, p1) of the date's row: rows past p1 read 0 without a
    // branch (they are then masked out).  Cash flows in LDS (CF_LDS) are read
    // with a bounds test instead.
    constexpr uint32_t SZ = sizeof(Real);
    const uint32_t len = (uint32_t)(p1 - p0);
    const __amdgpu_buffer_rsrc_t r_s0 = buf_rsrc(store + (int64_t)t * n + p0, len * SZ);
    const __amdgpu_buffer_rsrc_t r_s1 = buf_rsrc(store + (int64_t)(t >= 1 ? t - 1 : 0) * n + p0, len * SZ);
    const __amdgpu_buffer_rsrc_t r_cf = buf_rsrc(cfp, CF_LDS ? 0u : len * SZ);
    Real Sn[U], S1n[U], cfn[U];
    auto fetch = [&](uint32_t r0) {
#pragma unroll
        for (int u = 0; u < U; ++u) Sn[u] = buf_load<Real>(r_s0, (r0 + (uint32_t)(u * G + lt)) * SZ);
        if (t >= 1) {
#pragma unroll
            for (int u = 0; u < U; ++u) S1n[u] = buf_load<Real>(r_s1, (r0 + (uint32_t)(u * G + lt)) * SZ);
        } else {
#pragma unroll
            for (int u = 0; u < U; ++u) S1n[u] = Real(0);
        }
        if (t < N - 1) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t o = r0 + (uint32_t)(u * G + lt);
                if constexpr (CF_LDS) cfn[u] = o < len ? cfp[o] : Real(0);
                else cfn[u] = buf_load<Real>(r_cf, o * SZ);
            }
        } else {
#pragma unroll
            for (int u = 0; u < U; ++u) cfn[u] = Real(0);
        }
    };
    if (len > 0) fetch(0);
    // The first pass's loads are in flight while pre(C) provides the date's
    // coefficients (every thread calls it; it may hold block barriers).
    pre(C);
    for (uint32_t r0 = 0; r0 < len; r0 += (uint32_t)G * U) {
        Real Sv[U], S1v[U], cfv[U];
        bool act[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            Sv[u] = Sn[u];
            S1v[u] = S1n[u];
            cfv[u] = cfn[u];
            act[u] = r0 + (uint32_t)(u * G + lt) < len;
        }
        if (r0 + (uint32_t)G * U < len) fetch(r0 + (uint32_t)G * U);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t o = r0 + (uint32_t)(u * G + lt);
            Real c = Real(0);
            if (act[u]) {
                const Real Sx = Sv[u];
                const Real intrinsic = vanilla<Real>(right, Sx, K);
                if (t == N - 1) {
                    c = intrinsic;
                } else {
                    const Real dcf = cfv[u] * disc;
                    c = dcf;
                    if (intrinsic > Real(0)) {
                        Real L[NB];
                        lsm_basis<Real, NB, W>(Math<Real>::div(Sx, K), L);
                        double cont = 0.0;
#pragma unroll
                        for (int k = 0; k < NB; ++k) cont += C[k] * (double)L[k];
                        if ((double)intrinsic > cont) c = intrinsic;
                    }
                }
                if (t >= 1) {
                    const Real S1 = S1v[u];
                    if (vanilla<Real>(right, S1, K) > Real(0)) {
                        Real Lr[NB];
                        lsm_basis<Real, NB, W>(Math<Real>::div(S1, K), Lr);
                        const double Y = (double)(c * disc);
                        if constexpr (GRAM) {
                            using S = LsmSums<NB>;
                            double L[NB];
#pragma unroll
                            for (int k = 0; k < NB; ++k) L[k] = (double)Lr[k];
                            g[0] += 1.0;
#pragma unroll
                            for (int a2 = 0; a2 < NB; ++a2) {
#pragma unroll
                                for (int b2 = a2; b2 < NB; ++b2) g[LsmGram<NB>::m(a2, b2)] += L[a2] * L[b2];
                                g[S::b(a2)] += L[a2] * Y;
                            }
                        } else {
#pragma unroll
                            for (int k = 0; k < NB; ++k) g[k] += (double)Lr[k] * Y;
                        }
                    }
                } else {
                    c = c * disc;   // LSMAmerican.hpp:393-405
                }
                if constexpr (CF_LDS) cfp[o] = c;
                else buf_store(c, r_cf, o * SZ);
            }
            if (t == 0) {
                const unsigned long long am = __ballot(act[u]);
                if (am) {
                    const double n_act = (double)__popcll(am);
                    const double xd = (double)c;
                    const double sx = wave_sum(xd);
                    const double mean = sx / n_act;
                    const double d = act[u] ? xd - mean : 0.0;
                    const double m2 = wave_sum(d * d);
                    const unsigned long long im = __ballot(act[u] && c > itm_eps);
                    if (lid == 0) {
                        const int64_t tile = (p0 + (int64_t)r0 + u * G) / TILE + wv;
                        psum[tile] = sx;
                        pm2[tile] = m2;
                        pitm[tile] = (uint32_t)__popcll(im);
                    }
                }
            }
        }
    }
    if (t >= 1) {
        // Wavefront sums of the NV components padded to KP, then the group's
        // wavefronts in order.
        double v[KP];
#pragma unroll
        for (int k = 0; k < KP; ++k) v[k] = k < NV ? g[k] : 0.0;
        const double sx = wave_scatter_sum<KP>(v, lid);
        constexpr int SH = scatter_shift<KP>();
        const int comp = lid >> SH;
        if ((lid & ((1 << SH) - 1)) == 0) wpart[wv][comp] = sx;
        __syncthreads();
        if (part != nullptr && lt < KP) {
            double sx = 0.0;
            for (int w = 0; w < G / 64; ++w) sx += wpart[w][lt];
            part[lt] = sx;
        }
    }
}

// One Longstaff-Schwartz date for every request listed in the arguments, one
// launch per date: every block of request q first sums the right-hand-side
// partials of date t (written by the previous launch into buffer t & 1) in
// block order and solves with the date's pre-computed factor itself (block 0
// also stores the coefficients), then runs the date over its `chunk` of paths
// and writes its partial of date t-1 into buffer (t-1) & 1.  The redundant
// solves read nb partials per block; nb is kept at most PUMP_MAX_BLOCKS for
// that reason.
template <typename Real, int NB, bool W, int G, int U = 2>
__global__ __launch_bounds__(G) void k_lsm_step(PumpArgs<Real> a) {
    constexpr int ST = LsmRhs<NB>::STRIDE;
    int r = 0;
    while (r + 1 < a.n_req && (int)blockIdx.x >= a.block_start[r + 1]) ++r;
    const PumpReq<Real> &q = a.r[r];
    const int b = (int)blockIdx.x - a.block_start[r];
    const int t = q.t, N = q.n_steps;
    __shared__ double red[G / 64 * ST];
    __shared__ double wpart[G / 64][ST];
    auto pre = [&](double C[NB]) {
        double Cn[NB];
#pragma unroll
        for (int k = 0; k < NB; ++k) Cn[k] = 0.0;
        if (t < N - 1) {
            lsm_reduce_solve<NB, G>(q.bpart + (int64_t)(t & 1) * q.nb * ST, q.nb,
                                    q.factor + (int64_t)t * LSM_FACTOR_STRIDE, red, Cn);
        }
        if (b == 0 && threadIdx.x == 0)   // the last date's row is zero (no regression there)
            for (int k = 0; k < NB; ++k) q.coeffs[(int64_t)t * NB + k] = Cn[k];
#pragma unroll
        for (int k = 0; k < NB; ++k) C[k] = Cn[k];
    };
    double C[NB];
    const int64_t p0 = (int64_t)b * q.chunk;
    const int64_t p1 = p0 + q.chunk < q.n_paths ? p0 + q.chunk : q.n_paths;
    double *part = t >= 1 ? q.bpart + (int64_t)((t - 1) & 1) * q.nb * ST + (int64_t)b * ST : nullptr;
    lsm_chunk_date<Real, NB, W, false, G, U, false>(q.store, q.cf + p0, q.n_paths, p0, p1, t, N, q.right, q.K, q.disc, q.itm_eps, C,
                                          part, q.psum, q.pm2, q.pitm, wpart, pre);
}

// The request server's Longstaff-Schwartz step in two launches, for requests
// split into many blocks: k_lsm_solve has one block per request, which sums
// the request's nb Gram and right-hand-side partials of date t in block
// order, factorises, solves and stores the date's coefficients; k_lsm_paths
// then runs the date over every block's chunk with those coefficients (read
// after the kernel boundary) and writes its Gram and right-hand-side partial
// of date t-1 (the server has no pre-pass: a second read of every American
// path store costs the stream more bandwidth than the Gram sums cost here).  The solve does
// not sit in every block, so a request can take thousands of blocks (the
// server's pump steps share the GPU with batch launches and are long enough
// that one more launch per step costs nothing measurable).
template <typename Real, int NB, int G>
__global__ __launch_bounds__(G) void k_lsm_solve(PumpArgs<Real> a) {
    __builtin_amdgcn_s_setprio(3);   // see k_lsm_paths
    constexpr int ST = lsm_part_stride<NB, true>();
    const PumpReq<Real> &q = a.r[blockIdx.x];
    const int t = q.t;
    if (t >= q.n_steps - 1) return;                   // uniform per block: the last date needs no solve
    __shared__ double red[G / 64 * ST];
    double sm[LsmSums<NB>::NS];
    lsm_reduce_parts<NB, G, true>(q.bpart + (int64_t)(t & 1) * q.nb * ST, q.nb, red, sm);
    if (threadIdx.x >= 64) return;                    // one wavefront factorises and solves
    double F[LSM_FACTOR_STRIDE], C[NB];
    lsm_factor<NB>(sm, F);
    lsm_solve_factored<NB>(F, sm + LsmSums<NB>::NG, C);
    if (threadIdx.x == 0)
        for (int k = 0; k < NB; ++k) q.coeffs[(int64_t)t * NB + k] = C[k];
}

template <typename Real, int NB, bool W, int G>
__global__ __launch_bounds__(G) void k_lsm_paths(PumpArgs<Real> a) {
    // The pump's wavefronts share CUs with batch wavefronts that always have
    // an arithmetic instruction ready; the SIMD issues the oldest ready
    // wavefront first, so without a higher priority a pump step waits behind
    // the batch (measured: tens of milliseconds per date beside a batch).
    // Its memory-bound wavefronts issue few instructions, so the batch loses
    // little.
    __builtin_amdgcn_s_setprio(3);
    constexpr int ST = lsm_part_stride<NB, true>();
    int r = 0;
    while (r + 1 < a.n_req && (int)blockIdx.x >= a.block_start[r + 1]) ++r;
    const PumpReq<Real> &q = a.r[r];
    const int b = (int)blockIdx.x - a.block_start[r];
    const int t = q.t, N = q.n_steps;
    __shared__ double wpart[G / 64][ST];
    auto pre = [&](double C[NB]) {
#pragma unroll
        for (int k = 0; k < NB; ++k) C[k] = t < N - 1 ? q.coeffs[(int64_t)t * NB + k] : 0.0;
    };
    double C[NB];
    const int64_t p0 = (int64_t)b * q.chunk;
    const int64_t p1 = p0 + q.chunk < q.n_paths ? p0 + q.chunk : q.n_paths;
    double *part = t >= 1 ? q.bpart + (int64_t)((t - 1) & 1) * q.nb * ST + (int64_t)b * ST : nullptr;
    lsm_chunk_date<Real, NB, W, false, G, 2, true>(q.store, q.cf + p0, q.n_paths, p0, p1, t, N, q.right, q.K, q.disc, q.itm_eps, C,
                                          part, q.psum, q.pm2, q.pitm, wpart, pre);
}

// Gram pre-pass, block (chunk c, date d0 + d) of 256 threads: the count and
// the upper triangle of sum L L^T over the in-the-money paths of chunk c at
// that date (the same test and basis as the date step: LSMAmerican.hpp:
// 433-473), summed in FP64 in a fixed order (each thread's paths in order,
// the wavefront butterfly, the four wavefronts in order) into
// gpart[d][c][0 .. 16).
template <typename Real, int NB, bool W>
__global__ __launch_bounds__(256) void k_lsm_gram(LsmGramArgs<Real> a, int d0) {
    using S = LsmGram<NB>;
    constexpr int G = 256, U = 8, KP = S::STRIDE;
    constexpr uint32_t SZ = sizeof(Real);
    const int lt = (int)threadIdx.x, lid = lt & 63, wv = lt >> 6;
    const int c = (int)(blockIdx.x % (unsigned)a.gchunks);
    const int d = (int)(blockIdx.x / (unsigned)a.gchunks);
    const int64_t p0 = (int64_t)c * a.gchunk;
    const int64_t p1 = p0 + a.gchunk < a.n_paths ? p0 + a.gchunk : a.n_paths;
    const uint32_t len = (uint32_t)(p1 - p0);
    const __amdgpu_buffer_rsrc_t rs = buf_rsrc(a.store + (int64_t)(d0 + d) * a.n_paths + p0, len * SZ);
    double g[KP];
#pragma unroll
    for (int k = 0; k < KP; ++k) g[k] = 0.0;
    for (uint32_t r0 = 0; r0 < len; r0 += (uint32_t)G * U) {
        Real Sv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) Sv[u] = buf_load<Real>(rs, (r0 + (uint32_t)(u * G + lt)) * SZ);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (r0 + (uint32_t)(u * G + lt) < len && vanilla<Real>(a.right, Sv[u], a.K) > Real(0)) {
                Real Lr[NB];
                lsm_basis<Real, NB, W>(Math<Real>::div(Sv[u], a.K), Lr);
                double L[NB];
#pragma unroll
                for (int k = 0; k < NB; ++k) L[k] = (double)Lr[k];
                g[0] += 1.0;
#pragma unroll
                for (int i = 0; i < NB; ++i) {
#pragma unroll
                    for (int j = i; j < NB; ++j) g[S::m(i, j)] += L[i] * L[j];
                }
            }
        }
    }
    __shared__ double wpart[G / 64][KP];
    const double sx = wave_scatter_sum<KP>(g, lid);
    constexpr int SH = scatter_shift<KP>();
    if ((lid & ((1 << SH) - 1)) == 0) wpart[wv][lid >> SH] = sx;
    __syncthreads();
    if (lt < KP) {
        double v = 0.0;
        for (int w = 0; w < G / 64; ++w) v += wpart[w][lt];
        a.gpart[((int64_t)d * a.gchunks + c) * KP + lt] = v;
    }
}

// One block of 256 threads per date d0 + blockIdx.x: sums the date's
// pre-pass partials in a fixed order (thread (g, k) adds component k of
// chunks g, g + 16, g + 32, ... in order, then component k's 16 group sums
// in order) and writes the date's factor record.
template <int NB>
__global__ __launch_bounds__(256) void k_lsm_factor(const double *__restrict__ gpart, int gchunks,
                                                    double *__restrict__ factor, int d0) {
    using S = LsmGram<NB>;
    constexpr int KS = S::STRIDE, NGRP = 256 / KS, UR = 4;
    static_assert(KS == 16, "16 components per partial");
    const int d = (int)blockIdx.x, lt = (int)threadIdx.x, k = lt % KS, grp = lt / KS;
    const double *gp = gpart + (int64_t)d * gchunks * KS;
    double v = 0.0;
    for (int c0 = grp; c0 < gchunks; c0 += NGRP * UR) {
        double x[UR];
#pragma unroll
        for (int u = 0; u < UR; ++u) {
            const int c = c0 + u * NGRP;
            x[u] = c < gchunks ? gp[(int64_t)c * KS + k] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < UR; ++u) v += x[u];
    }
    __shared__ double red[NGRP][KS];
    red[grp][k] = v;
    __syncthreads();
    if (lt >= 64) return;
    double tot = 0.0;
    if (lt < KS)
#pragma unroll
        for (int g2 = 0; g2 < NGRP; ++g2) tot += red[g2][lt];
    const uint64_t vb = (uint64_t)__double_as_longlong(tot);
    double sm[S::NG];
#pragma unroll
    for (int j = 0; j < S::NG; ++j) {
        const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)vb, j);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(vb >> 32), j);
        sm[j] = __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
    }
    double F[LSM_FACTOR_STRIDE];
    lsm_factor<NB>(sm, F);
    if (lt == 0)
        for (int j = 0; j < LsmFactor<NB>::NF; ++j) factor[(int64_t)(d0 + d) * LSM_FACTOR_STRIDE + j] = F[j];
}

// Synchronisation of the cooperative backward pass (sync words, agent scope):
// arrivals are counted per group of blocks (block b in group b % G_R, one
// counter per 64-byte line) and groups on a top counter, so no address takes
// more than max(G_R, G / G_R) increments per date.  The block whose arrival
// completes the count sums the partials and solves; it publishes the
// coefficients and a generation number, which every other block polls.  The
// polls are bounded: on timeout the error word is raised and each block
// leaves at its next wait, so the kernel always ends.
constexpr int COOP_GROUPS = 8;
constexpr int SYNC_GRP = 0, SYNC_TOP = 16 * COOP_GROUPS, SYNC_GEN = SYNC_TOP + 16, SYNC_ERR = SYNC_GEN + 16;
constexpr int SYNC_COEF_BYTES = 1024;   // published coefficients: double [2][8] at this byte offset
static_assert(SYNC_ERR == LSM_COOP_SYNC_ERR_WORD, "error word index shared with the host");

// Returns true in every thread of the block that completed date k's count.
__device__ __forceinline__ bool coop_arrive(unsigned int *sync, unsigned int k) {
    __shared__ int last;
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned int G = gridDim.x, GR = G < (unsigned)COOP_GROUPS ? G : (unsigned)COOP_GROUPS;
        const unsigned int g = blockIdx.x % GR, in_group = (G - g + GR - 1) / GR;
        int is_last = 0;
        const unsigned int old =
            __hip_atomic_fetch_add(&sync[SYNC_GRP + 16 * g], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (old + 1 == k * in_group) {
            const unsigned int top =
                __hip_atomic_fetch_add(&sync[SYNC_TOP], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
            is_last = top + 1 == k * GR;
        }
        last = is_last;
    }
    __syncthreads();
    return last != 0;
}

// Waits for generation k; false (in every thread) on timeout or error.
__device__ __forceinline__ bool coop_wait(unsigned int *sync, unsigned int k) {
    __shared__ int ok;
    if (threadIdx.x == 0) {
        int good = 1;
        unsigned int spins = 0;
        // Relaxed polls (an acquire per poll would invalidate the cache each
        // time); one acquire fence once the generation is seen.
        while (__hip_atomic_load(&sync[SYNC_GEN], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < k) {
            if (__hip_atomic_load(&sync[SYNC_ERR], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) {
                good = 0;
                break;
            }
            if (++spins > LSM_COOP_SPIN_LIMIT) {
                __hip_atomic_store(&sync[SYNC_ERR], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                good = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        ok = good;
    }
    __syncthreads();
    return ok != 0;
}

// The whole backward pass of one request in one cooperative launch: block b
// holds the chunk of pump block b, the same partition, arithmetic and
// reduction order as the launch-per-date k_lsm_step (so the results are
// bit-identical), with the cash flows in LDS.  Per date: partials, arrival,
// and either the solve (the block that completes the count) or the wait for
// its published coefficients.  Partials and published coefficients alternate
// between two buffers by date parity.
template <typename Real, int NB, bool W>
__global__ __launch_bounds__(lsm_group<Real>()) void k_lsm_coop(LsmCoopArgs<Real> a) {
    constexpr int ST = LsmRhs<NB>::STRIDE;
    constexpr int G = lsm_group<Real>();
    extern __shared__ unsigned char lds_raw[];
    Real *cf_lds = reinterpret_cast<Real *>(lds_raw);
    __shared__ double wpart[G / 64][ST];
    __shared__ double red[G / 64 * ST];
    __shared__ double Cs[NB];
    double *cpub = reinterpret_cast<double *>(reinterpret_cast<unsigned char *>(a.sync) + SYNC_COEF_BYTES);
    const int vb = (int)blockIdx.x;
    const bool live = vb < a.nb;
    const int64_t p0 = live ? (int64_t)vb * a.chunk : 0;
    const int64_t p1 = live ? (p0 + a.chunk < a.n_paths ? p0 + a.chunk : a.n_paths) : 0;
    Real *cfp = cf_lds;
    const int N = a.n_steps;
    if (vb == 0 && threadIdx.x == 0)   // the last date has no regression
        for (int k = 0; k < NB; ++k) a.coeffs[(int64_t)(N - 1) * NB + k] = 0.0;
    unsigned int k_bar = 0;
    for (int t = N - 1; t >= 0; --t) {
        double C[NB];
#pragma unroll
        for (int k = 0; k < NB; ++k) C[k] = t < N - 1 ? Cs[k] : 0.0;
        const int par = (t - 1) & 1;
        double *part = (t >= 1 && live) ? a.bpart + (int64_t)par * a.nb * ST + (int64_t)vb * ST : nullptr;
        lsm_chunk_date<Real, NB, W, true, G, 8, false>(a.store, cfp, a.n_paths, p0, p1, t, N, a.right, a.K, a.disc, a.itm_eps, C,
                                             part, a.psum, a.pm2, a.pitm, wpart, [](double *) {});
        if (t >= 1) {
            ++k_bar;
            if (coop_arrive(a.sync, k_bar)) {
                double Cn[NB];
                lsm_reduce_solve<NB, G>(a.bpart + (int64_t)par * a.nb * ST, a.nb,
                                        a.factor + (int64_t)(t - 1) * LSM_FACTOR_STRIDE, red, Cn);
                if (threadIdx.x == 0) {
                    for (int k = 0; k < NB; ++k) {
                        Cs[k] = Cn[k];
                        cpub[par * 8 + k] = Cn[k];
                        a.coeffs[(int64_t)(t - 1) * NB + k] = Cn[k];
                    }
                    __hip_atomic_store(&a.sync[SYNC_GEN], k_bar, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                }
            } else {
                if (!coop_wait(a.sync, k_bar)) return;
                if (threadIdx.x < NB) Cs[threadIdx.x] = cpub[par * 8 + threadIdx.x];
            }
            __syncthreads();
        }
    }
    // Final cash flows to device memory (payoff distribution, path-set users).
    for (int64_t i = p0 + threadIdx.x; i < p1; i += G) a.cf[i] = cfp[i - p0];
}


// ---------------------------------------------------------------------------
// Serving: one batch of requests of uneven sizes in one launch.  Work items
// are (request, tile range); each wavefront takes the next item from the
// batch's counter until the list is exhausted, so every wavefront exits once
// the items run out.  A request's tile moments land at its own offset, so its
// result does not depend on the batch it ran in.
// ---------------------------------------------------------------------------
template <typename Real, int MODEL, int SCHEME>
__device__ __forceinline__ void batch_tile(const ReqD<Real> *r, int64_t tile) {
    const int f = r->flags;
    if (f & REQ_STORE) {
        run_tile<Real, MODEL, SCHEME, true, true, 1>(&r->lc, 1, &r->pay, 0, r->n_steps, r->n_paths, 0, r->key0,
                                                     r->key1, r->itm_eps, tile, r->store, nullptr, nullptr, nullptr,
                                                     0, 0, 0);
    } else if (f & REQ_PATHDEP) {
        run_tile<Real, MODEL, SCHEME, true, false, 1>(&r->lc, 1, &r->pay, 1, r->n_steps, r->n_paths, 0, r->key0,
                                                      r->key1, r->itm_eps, tile, nullptr, r->psum, r->pm2, r->pitm,
                                                      tile, 0, 0);
    } else {
        run_tile<Real, MODEL, SCHEME, false, false, 1>(&r->lc, 1, &r->pay, 1, r->n_steps, r->n_paths, 0, r->key0,
                                                       r->key1, r->itm_eps, tile, nullptr, r->psum, r->pm2, r->pitm,
                                                       tile, 0, 0);
    }
}

template <typename Real>
__global__ __launch_bounds__(BLOCK) void k_batch(const ReqD<Real> *__restrict__ reqs, const ItemD *__restrict__ items,
                                                 int32_t n_items, unsigned int *counter, int32_t raised) {
    // The server's express batches (small requests beside a large batch) take
    // issue precedence over the large batch's wavefronts, below the pump's.
    if (raised) __builtin_amdgcn_s_setprio(2);
    const int lid = threadIdx.x & 63;
    for (;;) {
        unsigned int v = 0;
        if (lid == 0) v = atomicAdd(counter, 1u);
        const int it = (int)__builtin_amdgcn_readfirstlane(v);
        if (it >= n_items) return;
        const ItemD item = items[it];
        const ReqD<Real> *r = reqs + item.req;
        const int sel = r->model * 3 + r->scheme;
        for (int k = 0; k < item.n_tiles; ++k) {
            const int64_t tile = item.tile0 + k;
            switch (sel) {
            case 0: batch_tile<Real, MODEL_HESTON, SCHEME_EULER>(r, tile); break;
            case 1: batch_tile<Real, MODEL_HESTON, SCHEME_MILSTEIN>(r, tile); break;
            case 2: batch_tile<Real, MODEL_HESTON, SCHEME_QE>(r, tile); break;
            case 3: batch_tile<Real, MODEL_BATES, SCHEME_EULER>(r, tile); break;
            case 4: batch_tile<Real, MODEL_BATES, SCHEME_MILSTEIN>(r, tile); break;
            default: batch_tile<Real, MODEL_BATES, SCHEME_QE>(r, tile); break;
            }
        }
    }
}

// Batched two-stage merge: stage 1 over a flattened (request, chunk) list,
// stage 2 one block per request; both in fixed order within a request.
// Price, SE and P(ITM) of a request's merged moments (OptionPricer.hpp:204-249).
__device__ __forceinline__ qk_stats bmerge_result(const Mom &a, double factor) {
    const double N = a.n;
    const double var = N > 1.0 ? a.m2 / (N - 1.0) : 0.0;
    qk_stats o;
    o.price = a.mean * factor;
    o.se = sqrt(var / N) * factor;
    o.prob_itm = (double)a.itm / N;
    o.mean = a.mean;
    o.m2 = a.m2;
    o.n = (int64_t)N;
    o.itm = (int64_t)a.itm;
    return o;
}

// Stage 1 also finalises a request of a single chunk: stage 2 would merge
// that chunk into an empty accumulator, which copies it (chan_merge), so the
// result has the same bits, and stage 2 skips the request.
__global__ __launch_bounds__(256) void k_bmerge_chunks(const MergeReqD *__restrict__ mr, const int32_t *__restrict__ chunk_req,
                                                       qk_stats *__restrict__ out) {
    const int req = chunk_req[blockIdx.x];
    const MergeReqD m = mr[req];
    const int64_t chunk = blockIdx.x - m.first_chunk;
    Mom a{0.0, 0.0, 0.0, 0ull};
    const int64_t t0 = chunk * MERGE_CHUNK + (int64_t)threadIdx.x * (MERGE_CHUNK / 256);
#pragma unroll
    for (int k = 0; k < MERGE_CHUNK / 256; ++k) {
        const int64_t t = t0 + k;
        if (t < m.n_tiles) {
            const int64_t rem = m.n_paths - t * TILE;
            const double nt = (double)(rem < TILE ? rem : TILE);
            chan_merge(a.n, a.mean, a.m2, nt, m.psum[t] / nt, m.pm2[t]);
            a.itm += m.pitm[t];
        }
    }
    a = block_tree_merge(a);
    if (threadIdx.x == 0) {
        if (m.n_chunks == 1) out[req] = bmerge_result(a, m.factor);
        else ((Mom *)m.scratch)[chunk] = a;
    }
}

__global__ __launch_bounds__(256) void k_bmerge_final(const MergeReqD *__restrict__ mr, qk_stats *__restrict__ out) {
    const MergeReqD m = mr[blockIdx.x];
    if (m.n_chunks == 1) return;                     // finalised by stage 1 (uniform per block)
    Mom a{0.0, 0.0, 0.0, 0ull};
    for (int64_t c = threadIdx.x; c < m.n_chunks; c += 256) {
        const Mom x = ((const Mom *)m.scratch)[c];
        chan_merge(a.n, a.mean, a.m2, x.n, x.mean, x.m2);
        a.itm += x.itm;
    }
    a = block_tree_merge(a);
    if (threadIdx.x == 0) out[blockIdx.x] = bmerge_result(a, m.factor);
}

// ---------------------------------------------------------------------------
// Launchers.
// ---------------------------------------------------------------------------
template <typename Real, int MODEL, int SCHEME>
static void launch_ms(const PathArgs<Real> &a, bool pathdep, bool store, int L, dim3 grid, hipStream_t s) {
    if (store) {
        hipLaunchKernelGGL((k_paths<Real, MODEL, SCHEME, true, true, 1>), grid, dim3(BLOCK), 0, s, a);
    } else if (L == 8 && sizeof(Real) == 4) {   // FP32 only: eight FP64 lanes exceed the register file
        if constexpr (sizeof(Real) == 4) {
            if (pathdep) hipLaunchKernelGGL((k_paths<Real, MODEL, SCHEME, true, false, 8>), grid, dim3(BLOCK), 0, s, a);
            else hipLaunchKernelGGL((k_paths<Real, MODEL, SCHEME, false, false, 8>), grid, dim3(BLOCK), 0, s, a);
        }
    } else if (L == 4) {
        if (pathdep) hipLaunchKernelGGL((k_paths<Real, MODEL, SCHEME, true, false, 4>), grid, dim3(BLOCK), 0, s, a);
        else hipLaunchKernelGGL((k_paths<Real, MODEL, SCHEME, false, false, 4>), grid, dim3(BLOCK), 0, s, a);
    } else if (L == 2) {
        if (pathdep) hipLaunchKernelGGL((k_paths<Real, MODEL, SCHEME, true, false, 2>), grid, dim3(BLOCK), 0, s, a);
        else hipLaunchKernelGGL((k_paths<Real, MODEL, SCHEME, false, false, 2>), grid, dim3(BLOCK), 0, s, a);
    } else {
        if (pathdep) hipLaunchKernelGGL((k_paths<Real, MODEL, SCHEME, true, false, 1>), grid, dim3(BLOCK), 0, s, a);
        else hipLaunchKernelGGL((k_paths<Real, MODEL, SCHEME, false, false, 1>), grid, dim3(BLOCK), 0, s, a);
    }
}

// One launch per window of at most PATHS_WINDOW_BLOCKS blocks (a single
// launch below 2^30 paths); every window writes its own tiles' moments, so
// the windows change nothing in the results.
template <typename Real>
hipError_t launch_paths(const PathArgs<Real> &a, int model, int scheme, bool pathdep, bool store, int L,
                        hipStream_t s) {
    if (model < 0 || model > 1 || scheme < 0 || scheme > 2) return hipErrorInvalidValue;
    const int64_t blocks = (a.n_tiles + WAVES_PER_BLOCK - 1) / WAVES_PER_BLOCK;
    const unsigned gy = (unsigned)((a.n_lanes + L - 1) / L);
    for (int64_t b0 = 0; b0 < blocks; b0 += PATHS_WINDOW_BLOCKS) {
        PathArgs<Real> w = a;
        w.tile_off = b0 * WAVES_PER_BLOCK;
        const dim3 grid((unsigned)std::min<int64_t>(blocks - b0, PATHS_WINDOW_BLOCKS), gy);
        switch (model * 3 + scheme) {
        case 0: launch_ms<Real, MODEL_HESTON, SCHEME_EULER>(w, pathdep, store, L, grid, s); break;
        case 1: launch_ms<Real, MODEL_HESTON, SCHEME_MILSTEIN>(w, pathdep, store, L, grid, s); break;
        case 2: launch_ms<Real, MODEL_HESTON, SCHEME_QE>(w, pathdep, store, L, grid, s); break;
        case 3: launch_ms<Real, MODEL_BATES, SCHEME_EULER>(w, pathdep, store, L, grid, s); break;
        case 4: launch_ms<Real, MODEL_BATES, SCHEME_MILSTEIN>(w, pathdep, store, L, grid, s); break;
        default: launch_ms<Real, MODEL_BATES, SCHEME_QE>(w, pathdep, store, L, grid, s); break;
        }
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

int64_t merge_chunks(int64_t n_tiles) { return (n_tiles + MERGE_CHUNK - 1) / MERGE_CHUNK; }

hipError_t launch_merge(const double *ps, const double *pm2, const uint32_t *pitm, int64_t n_tiles, int64_t n_paths,
                        int n_lanes, int n_pay, const double *disc, const double *scale, void *scratch, qk_stats *out,
                        hipStream_t s, bool chunks_only) {
    const int64_t nc = merge_chunks(n_tiles);
    const int groups = n_lanes * n_pay;
    hipLaunchKernelGGL(k_merge_chunks, dim3((unsigned)nc, (unsigned)groups), dim3(256), 0, s, ps, pm2, pitm, n_tiles,
                       n_paths, nc, (Mom *)scratch, n_pay, disc, scale, nc == 1 && !chunks_only ? out : nullptr);
    if (nc > 1 && !chunks_only)
        hipLaunchKernelGGL(k_merge_final, dim3(groups), dim3(256), 0, s, (const Mom *)scratch, nc, n_pay, disc, scale,
                           out);
    return hipGetLastError();
}

hipError_t launch_merge_final(const void *chunks, int64_t n_chunks, int groups, int n_pay, const double *disc,
                              const double *scale, qk_stats *out, hipStream_t s) {
    hipLaunchKernelGGL(k_merge_final, dim3(groups), dim3(256), 0, s, (const Mom *)chunks, n_chunks, n_pay, disc, scale,
                       out);
    return hipGetLastError();
}

hipError_t launch_greeks(const qk_stats *st, int n_pay, const GreekMap &m, qk_greeks *out, hipStream_t s) {
    hipLaunchKernelGGL(k_greeks, dim3(1), dim3(64), 0, s, st, n_pay, m, out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Order statistics of per-path values (payoff distribution percentiles) by
// successive 8-bit histograms: each pass counts, for every still-open prefix
// (the key bits fixed so far for one or more wanted ranks), the next 8 bits of
// the keys that carry that prefix.  Counts are integers, so the selected
// values are exact and independent of the launch shape.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t order_key(float x) {
    const uint32_t b = __float_as_uint(x);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ uint64_t order_key(double x) {
    const uint64_t b = (uint64_t)__double_as_longlong(x);
    return (b & 0x8000000000000000ull) ? ~b : (b | 0x8000000000000000ull);
}

template <typename Real>
__global__ __launch_bounds__(256) void k_select_hist(const Real *__restrict__ vals, int64_t n, SelectArgs a,
                                                     uint32_t *__restrict__ hist) {
    __shared__ uint32_t h[SELECT_MAX_GROUPS * 256];
    for (int i = threadIdx.x; i < a.n_groups * 256; i += 256) h[i] = 0u;
    __syncthreads();
    constexpr int BITS = 8 * (int)sizeof(Real);
    const int hs = a.shift + 8;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const uint64_t k = (uint64_t)order_key(vals[i]);
        const uint64_t hi = hs >= BITS ? 0ull : (k >> hs);
        const uint32_t d = (uint32_t)(k >> a.shift) & 255u;
        for (int g = 0; g < a.n_groups; ++g) {
            if (hi == a.prefix[g]) {
                atomicAdd(&h[g * 256 + d], 1u);
                break;
            }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < a.n_groups * 256; i += 256)
        if (h[i]) atomicAdd(&hist[i], h[i]);
}

template <typename Real>
hipError_t launch_select_hist(const Real *vals, int64_t n, const SelectArgs &a, uint32_t *hist, int blocks,
                              hipStream_t s) {
    if (n <= 0 || a.n_groups <= 0 || a.n_groups > SELECT_MAX_GROUPS) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_select_hist<Real>, dim3(blocks), dim3(256), 0, s, vals, n, a, hist);
    return hipGetLastError();
}
template hipError_t launch_select_hist<float>(const float *, int64_t, const SelectArgs &, uint32_t *, int,
                                              hipStream_t);
template hipError_t launch_select_hist<double>(const double *, int64_t, const SelectArgs &, uint32_t *, int,
                                               hipStream_t);

// ---------------------------------------------------------------------------
// Path-store transpose for the ".paths" payload: the store is [step][path]
// (coalesced for the path kernel and the backward pass), the file is
// [path][step].  One 64 x 64 tile per block through LDS (a padded row keeps
// the column reads free of bank conflicts); both the reads and the writes
// are 64 consecutive elements per wavefront.
// ---------------------------------------------------------------------------
template <typename Real>
__global__ __launch_bounds__(256) void k_transpose_paths(const Real *__restrict__ store, int64_t n_paths,
                                                         int32_t n_steps, int64_t p0, int64_t n,
                                                         Real *__restrict__ out) {
    __shared__ Real tile[64][65];
    const int lx = threadIdx.x & 63, ly = threadIdx.x >> 6;
    const int64_t pb = (int64_t)blockIdx.x * 64;     // first path of the tile (relative to p0)
    const int tb = (int)blockIdx.y * 64;              // first step of the tile
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const int t = tb + ly + 4 * k;
        const int64_t i = pb + lx;
        if (t < n_steps && i < n) tile[ly + 4 * k][lx] = store[(int64_t)t * n_paths + p0 + i];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const int64_t i = pb + ly + 4 * k;
        const int t = tb + lx;
        if (t < n_steps && i < n) out[i * n_steps + t] = tile[lx][ly + 4 * k];
    }
}

template <typename Real>
hipError_t launch_transpose_paths(const Real *store, int64_t n_paths, int32_t n_steps, int64_t p0, int64_t n,
                                  Real *out, hipStream_t s) {
    if (n <= 0 || p0 < 0 || p0 + n > n_paths || n_steps < 1) return hipErrorInvalidValue;
    const dim3 grid((unsigned)((n + 63) / 64), (unsigned)((n_steps + 63) / 64));
    hipLaunchKernelGGL(k_transpose_paths<Real>, grid, dim3(256), 0, s, store, n_paths, n_steps, p0, n, out);
    return hipGetLastError();
}
template hipError_t launch_transpose_paths<float>(const float *, int64_t, int32_t, int64_t, int64_t, float *,
                                                  hipStream_t);
template hipError_t launch_transpose_paths<double>(const double *, int64_t, int32_t, int64_t, int64_t, double *,
                                                   hipStream_t);

int lsm_gram_chunks(int64_t n_paths, int64_t *gchunk) {
    // at least 8,192 paths (four passes of the block's 2,048-path rows) and at
    // most 64 chunks per date, at most 2^28 paths (32-bit byte offsets)
    int64_t c = std::max<int64_t>(8192, (n_paths + 63) / 64);
    c = std::min<int64_t>((c + 2047) / 2048 * 2048, (int64_t)1 << 28);
    *gchunk = c;
    return (int)((n_paths + c - 1) / c);
}

int lsm_gram_window(int gchunks) { return std::max(1, 65536 / std::max(gchunks, 1)); }

size_t lsm_prepass_bytes(int64_t n_paths, int32_t n_steps) {
    int64_t gchunk = 0;
    const int gchunks = lsm_gram_chunks(n_paths, &gchunk);
    const int64_t dates = std::max<int64_t>(1, std::min<int64_t>(n_steps - 1, lsm_gram_window(gchunks)));
    return (size_t)dates * gchunks * LsmGram<5>::STRIDE * 8 + (size_t)n_steps * LSM_FACTOR_STRIDE * 8;
}

template <typename Real, int NB, bool W>
static void launch_gram_nb(const LsmGramArgs<Real> &a, hipStream_t s) {
    const int window = lsm_gram_window(a.gchunks);
    for (int d0 = 0; d0 < a.n_steps - 1; d0 += window) {
        const int nd = std::min(window, a.n_steps - 1 - d0);
        hipLaunchKernelGGL((k_lsm_gram<Real, NB, W>), dim3((unsigned)(a.gchunks * nd)), dim3(256), 0, s, a, d0);
        hipLaunchKernelGGL((k_lsm_factor<NB>), dim3((unsigned)nd), dim3(256), 0, s, (const double *)a.gpart,
                           a.gchunks, a.factor, d0);
    }
}

template <typename Real>
hipError_t launch_lsm_gram(const LsmGramArgs<Real> &a, int n_basis, bool weighted, hipStream_t s) {
    if (a.n_steps < 2) return hipSuccess;
    switch (n_basis * 2 + (weighted ? 1 : 0)) {
    case 6: launch_gram_nb<Real, 3, false>(a, s); break;
    case 7: launch_gram_nb<Real, 3, true>(a, s); break;
    case 8: launch_gram_nb<Real, 4, false>(a, s); break;
    case 9: launch_gram_nb<Real, 4, true>(a, s); break;
    case 10: launch_gram_nb<Real, 5, false>(a, s); break;
    case 11: launch_gram_nb<Real, 5, true>(a, s); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
template hipError_t launch_lsm_gram<float>(const LsmGramArgs<float> &, int, bool, hipStream_t);
template hipError_t launch_lsm_gram<double>(const LsmGramArgs<double> &, int, bool, hipStream_t);

int pump_blocks(int64_t n_paths, int group, int max_blocks, int64_t *chunk) {
    const int64_t nb = std::max<int64_t>(1, std::min<int64_t>((n_paths + group - 1) / group, max_blocks));
    const int64_t c = ((n_paths + nb - 1) / nb + 255) / 256 * 256;
    *chunk = c;
    return (int)((n_paths + c - 1) / c);
}

template <typename Real, int NB, bool W>
static void launch_pump_nb(const PumpArgs<Real> &a, hipStream_t s) {
    constexpr int G = lsm_group<Real>();
    hipLaunchKernelGGL((k_lsm_step<Real, NB, W, G>), dim3(a.block_start[a.n_req]), dim3(G), 0, s, a);
}

template <typename Real, int NB, bool W>
static hipError_t launch_coop_nb(const LsmCoopArgs<Real> &a, int blocks, size_t lds, hipStream_t s) {
    void *args[] = {const_cast<LsmCoopArgs<Real> *>(&a)};
    return hipLaunchCooperativeKernel(reinterpret_cast<const void *>(&k_lsm_coop<Real, NB, W>), dim3(blocks),
                                      dim3(lsm_group<Real>()), args, (unsigned int)lds, s);
}

template <typename Real, int NB, bool W>
static int coop_occupancy(size_t lds) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void *>(&k_lsm_coop<Real, NB, W>),
                                                     lsm_group<Real>(), lds) != hipSuccess)
        return 0;
    return per_cu;
}

template <typename Real>
int lsm_coop_occupancy(int n_basis, bool weighted, size_t lds) {
    switch (n_basis * 2 + (weighted ? 1 : 0)) {
    case 6: return coop_occupancy<Real, 3, false>(lds);
    case 7: return coop_occupancy<Real, 3, true>(lds);
    case 8: return coop_occupancy<Real, 4, false>(lds);
    case 9: return coop_occupancy<Real, 4, true>(lds);
    case 10: return coop_occupancy<Real, 5, false>(lds);
    case 11: return coop_occupancy<Real, 5, true>(lds);
    default: return 0;
    }
}

template <typename Real>
hipError_t launch_lsm_coop(const LsmCoopArgs<Real> &a, int n_basis, bool weighted, int blocks, size_t lds,
                           hipStream_t s) {
    switch (n_basis * 2 + (weighted ? 1 : 0)) {
    case 6: return launch_coop_nb<Real, 3, false>(a, blocks, lds, s);
    case 7: return launch_coop_nb<Real, 3, true>(a, blocks, lds, s);
    case 8: return launch_coop_nb<Real, 4, false>(a, blocks, lds, s);
    case 9: return launch_coop_nb<Real, 4, true>(a, blocks, lds, s);
    case 10: return launch_coop_nb<Real, 5, false>(a, blocks, lds, s);
    case 11: return launch_coop_nb<Real, 5, true>(a, blocks, lds, s);
    default: return hipErrorInvalidValue;
    }
}
template int lsm_coop_occupancy<float>(int, bool, size_t);
template int lsm_coop_occupancy<double>(int, bool, size_t);
template hipError_t launch_lsm_coop<float>(const LsmCoopArgs<float> &, int, bool, int, size_t, hipStream_t);
template hipError_t launch_lsm_coop<double>(const LsmCoopArgs<double> &, int, bool, int, size_t, hipStream_t);

template <typename Real, int NB, bool W>
static void launch_pump2_nb(const PumpArgs<Real> &a, hipStream_t s) {
    hipLaunchKernelGGL((k_lsm_solve<Real, NB, LSM_SOLVE_GROUP>), dim3(a.n_req), dim3(LSM_SOLVE_GROUP), 0, s, a);
    hipLaunchKernelGGL((k_lsm_paths<Real, NB, W, LSM_SHARED_GROUP>), dim3(a.block_start[a.n_req]),
                       dim3(LSM_SHARED_GROUP), 0, s, a);
}

template <typename Real>
hipError_t launch_pump_two_phase(const PumpArgs<Real> &a, int n_basis, bool weighted, hipStream_t s) {
    if (a.n_req <= 0) return hipSuccess;
    switch (n_basis * 2 + (weighted ? 1 : 0)) {
    case 6: launch_pump2_nb<Real, 3, false>(a, s); break;
    case 7: launch_pump2_nb<Real, 3, true>(a, s); break;
    case 8: launch_pump2_nb<Real, 4, false>(a, s); break;
    case 9: launch_pump2_nb<Real, 4, true>(a, s); break;
    case 10: launch_pump2_nb<Real, 5, false>(a, s); break;
    case 11: launch_pump2_nb<Real, 5, true>(a, s); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
template hipError_t launch_pump_two_phase<float>(const PumpArgs<float> &, int, bool, hipStream_t);
template hipError_t launch_pump_two_phase<double>(const PumpArgs<double> &, int, bool, hipStream_t);

int server_pump_blocks(int64_t n_paths, int64_t *chunk) {
    // one pass of 8 rows of 256 threads per block, at most SERVER_PUMP_MAX_BLOCKS blocks
    const int64_t rows = (n_paths + SERVER_PUMP_MAX_BLOCKS - 1) / SERVER_PUMP_MAX_BLOCKS;
    const int64_t c = std::max<int64_t>(SERVER_PUMP_CHUNK, (rows + 255) / 256 * 256);
    *chunk = c;
    return (int)((n_paths + c - 1) / c);
}

template <typename Real>
hipError_t launch_pump(const PumpArgs<Real> &a, int n_basis, bool weighted, hipStream_t s) {
    if (a.n_req <= 0) return hipSuccess;
    switch (n_basis * 2 + (weighted ? 1 : 0)) {
    case 6: launch_pump_nb<Real, 3, false>(a, s); break;
    case 7: launch_pump_nb<Real, 3, true>(a, s); break;
    case 8: launch_pump_nb<Real, 4, false>(a, s); break;
    case 9: launch_pump_nb<Real, 4, true>(a, s); break;
    case 10: launch_pump_nb<Real, 5, false>(a, s); break;
    case 11: launch_pump_nb<Real, 5, true>(a, s); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename Real>
hipError_t launch_batch(const ReqD<Real> *reqs, const ItemD *items, int32_t n_items, unsigned int *counter,
                        int blocks, bool raised, hipStream_t s) {
    if (n_items <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_batch<Real>, dim3(blocks), dim3(BLOCK), 0, s, reqs, items, n_items, counter,
                       (int32_t)raised);
    return hipGetLastError();
}

hipError_t launch_batch_merge(const MergeReqD *mr, int n_reqs, const int32_t *chunk_req, int64_t n_chunks,
                              qk_stats *out, hipStream_t s) {
    if (n_reqs <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_bmerge_chunks, dim3((unsigned)n_chunks), dim3(256), 0, s, mr, chunk_req, out);
    // every request has at least one chunk: as many chunks as requests means
    // one each, all finalised by stage 1
    if (n_chunks > n_reqs) hipLaunchKernelGGL(k_bmerge_final, dim3(n_reqs), dim3(256), 0, s, mr, out);
    return hipGetLastError();
}

template hipError_t launch_batch<float>(const ReqD<float> *, const ItemD *, int32_t, unsigned int *, int, bool,
                                        hipStream_t);
template hipError_t launch_batch<double>(const ReqD<double> *, const ItemD *, int32_t, unsigned int *, int, bool,
                                         hipStream_t);
template hipError_t launch_paths<float>(const PathArgs<float> &, int, int, bool, bool, int, hipStream_t);
template hipError_t launch_paths<double>(const PathArgs<double> &, int, int, bool, bool, int, hipStream_t);
template hipError_t launch_pump<float>(const PumpArgs<float> &, int, bool, hipStream_t);
template hipError_t launch_pump<double>(const PumpArgs<double> &, int, bool, hipStream_t);

}  // namespace qk
