/*
 * qk_oracle.h -- FP64 CPU restatement of QuantKos's Monte Carlo hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation in quantkos_amd/.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it.  The product never links or calls it.
 *
 * It restates, in plain C with FP64 arithmetic, the reference's per-path
 * algorithm (schemes, payoff trackers, pricer statistics, Longstaff-Schwartz,
 * finite-difference Greek stencils) and its closed-form pricers.  Every
 * function cites the reference file:line it follows.
 *
 * Random numbers: the reference draws from Kokkos' XorShift64 pool, whose
 * stream depends on the launch shape.  This build replaces it by a
 * counter-based stream that both the oracle and the GPU kernels implement
 * independently, so per-path results can be compared:
 *
 *   key      = (lo32(seed), hi32(seed))
 *   block(path, b, slot) = Philox4x32-10(counter = (lo32(path), hi32(path), b, slot), key)
 *   U(w)     = (w + 0.5) * 2^-32                            (FP64 runs)
 *            = fl32(fl32(w) * 2^-32 + 2^-33), in (0, 1]     (FP32 runs, "u24":
 *              w rounded to a 24-bit significand, nearest even)
 *   words (a0, a1, a2, a3) of step t:
 *     Heston Euler / Milstein: two per step, steps 2i, 2i+1 take words (0, 1)
 *                        and (2, 3) of block(path, i, 0)
 *     Heston QE:         three per step, words 3t, 3t+1, 3t+2 of the sequence
 *                        block(path, 0, 0) w0..w3, block(path, 1, 0) w0..w3, ...
 *     Bates:             the four words of block(path, t, 0)
 *   Z1 = sqrt(-2 ln U(a0)) cos(2 pi U(a1)),  Z2 = sqrt(-2 ln U(a0)) sin(2 pi U(a1))
 *     Euler / Milstein:  Z_v = Z1,  Z_S = rho Z1 + sqrt(1-rho^2) Z2
 *     QE:                Z_V = Z1,  Z_indep = Z2,  exponential-branch uniform = U(a2)
 *     Bates:             first Poisson uniform = U(a3); further uniforms e_j are
 *                        word (j mod 4) of block(path, t, 1 + j/4); the jump
 *                        normal after N jumps is sqrt(-2 ln e_N) cos(2 pi e_{N+1}).
 */
#ifndef QK_ORACLE_H
#define QK_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Enumerations follow the order of the reference's ConfigFileEnums.hpp:24-58. */
enum { QKO_HESTON = 0, QKO_BATES = 1 };
enum { QKO_EULER = 0, QKO_MILSTEIN = 1, QKO_QE = 2 };
enum {
    QKO_EUROPEAN = 0, QKO_ASIAN, QKO_BARRIER_UO, QKO_BARRIER_DO, QKO_BARRIER_UI,
    QKO_BARRIER_DI, QKO_LOOKBACK_FLOAT, QKO_LOOKBACK_FIXED, QKO_BINARY_CASH,
    QKO_BINARY_ASSET, QKO_AMERICAN
};
enum { QKO_CALL = 0, QKO_PUT = 1 };

typedef struct {
    int32_t model;           /* QKO_HESTON / QKO_BATES */
    int32_t scheme;          /* QKO_EULER / QKO_MILSTEIN / QKO_QE */
    int32_t fp32_semantics;  /* 1: FP32 uniforms (u24) and the FP32 eps constants */
    int32_t n_steps;
    uint64_t seed;
    double S0, v0, r, q, T;
    double kappa, theta, sigma, rho;
    double lambda_j, mu_j, sigma_j;
} qko_model;

typedef struct { int32_t type, right; double K, B; } qko_payoff;

typedef struct {
    double price, se, prob_itm, mean, m2;
    int64_t n, itm;
} qko_stats;

typedef struct {
    int32_t delta_gamma, vega_vomma, vanna, rho, theta;
    int32_t theta_central;   /* 0: the reference's one-sided T-h; 1: also price T+h */
    double rel_S, abs_vol, abs_r, abs_T;
} qko_bumps;

typedef struct {
    double delta, gamma, vega, vomma, vanna, rho, theta;
    double base_price, base_se;
    int32_t n_pricings;
} qko_greeks;

/* ---- random numbers ---- */
void qko_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double qko_uniform(uint32_t w, int u24);
/* Normals (Z1, Z2) of step t of path p, laid out z[(p*n_steps + t)*2 + {0,1}]. */
void qko_normals(uint64_t seed, int u24, int64_t n_paths, int32_t n_steps, double *z);

/* ---- paths ---- */
/* Simulates paths [path0, path0+n_paths).  Any output pointer may be NULL.
 * S_store is laid out [step][path] with leading dimension n_paths (S after step t+1). */
int qko_simulate(const qko_model *m, int64_t path0, int64_t n_paths, double *S_store,
                 double *S_T, double *v_T, double *S_sum, double *S_max, double *S_min);

double qko_payoff_eval(const qko_payoff *p, double S_T, double S_sum, double S_max,
                       double S_min, int32_t n_steps);

/* ---- pricing ---- */
int qko_price(const qko_model *m, int64_t n_paths, const qko_payoff *pay, int32_t n_pay,
              double price_scale, qko_stats *out);
/* qko_price over the paths [path0, path0 + n_paths) of the stream. */
int qko_price_window(const qko_model *m, int64_t path0, int64_t n_paths, const qko_payoff *pay, int32_t n_pay,
                     double price_scale, qko_stats *out);
int qko_price_american(const qko_model *m, int64_t n_paths, int32_t right, double K,
                       double price_scale, qko_stats *out, double *coeffs_out, double *cf_out);
/* nb = 3, 4, 5 basis functions (LaguerreP02/P03/P04); weighted: times exp(-X/2).
 * coeffs_out: [n_steps][nb]. */
#define QKO_MAX_BASIS 5
int qko_price_american_ex(const qko_model *m, int64_t n_paths, int32_t right, double K, int32_t nb,
                          int32_t weighted, double price_scale, qko_stats *out, double *coeffs_out,
                          double *cf_out);
int qko_greeks_eval(const qko_model *m, int64_t n_paths, const qko_payoff *pay,
                    const qko_bumps *b, double price_scale, qko_greeks *out);

/* ---- statistics and regression helpers ---- */
void qko_chan_merge(double *n_a, double *mean_a, double *m2_a,
                    double n_b, double mean_b, double m2_b);
/* sums = {count, M[16] row-major, B[4]} ; writes 4 coefficients. Returns 1 on fallback. */
int qko_lsm_solve(const double sums[21], double coeffs[4]);
void qko_laguerre4(double X, double L[4]);
/* sums = {count, M[nb*nb] row-major, B[nb]} ; writes nb coefficients. */
int qko_lsm_solve_n(const double *sums, int32_t nb, double *coeffs);
void qko_basis(double X, int32_t nb, int32_t weighted, double *L);

/* ---- closed forms ---- */
void qko_gl64(double x[64], double w[64]);
double qko_heston_call(double S0, double K, double T, double r, double q, double v0,
                       double kappa, double theta, double sigma, double rho);
double qko_bates_call(double S0, double K, double T, double r, double q, double v0,
                      double kappa, double theta, double sigma, double rho,
                      double lambda_j, double mu_j, double sigma_j);
double qko_bs_call(double S0, double K, double T, double r, double q, double vol);

#ifdef __cplusplus
}
#endif
#endif
