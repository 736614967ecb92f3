/*
 * qk_oracle.c -- FP64 CPU restatement of QuantKos's Monte Carlo hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP implementation.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Reference files restated here (paths relative to the reference root):
 *   include/core/schemes/markovian/MSchemes.hpp   (schemes)
 *   include/core/schemes/markovian/MSolver.hpp    (per-path loop)
 *   include/core/options/Payoff.hpp               (payoff trackers)
 *   include/core/options/OptionPricer.hpp         (statistics)
 *   include/core/options/LSMAmerican.hpp          (Longstaff-Schwartz)
 *   include/core/engine/ForwardBackwardMC.hpp     (LSM driver)
 *   include/core/engine/MCEngine.hpp              (Greek stencils)
 *   include/core/analytical/                       (closed forms)
 * The random stream is this build's counter-based one (see qk_oracle.h).
 */
#include "qk_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as  */
/* easy as 1, 2, 3", SC'11): 10 rounds of the 4x32 S-box with multipliers   */
/* 0xD2511F53 / 0xCD9E8D57 and Weyl key increments 0x9E3779B9 / 0xBB67AE85. */
/* ------------------------------------------------------------------------ */
void qko_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int round = 0; round < 10; ++round) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* FP32 map (u24 != 0): w rounded to float (nearest even), times 2^-32, plus
 * 2^-33, rounded to float, in (0, 1]; the product is exact, so this is the
 * device's fmaf.  FP64 map: (w + 1/2) 2^-32, exact, in (0, 1). */
double qko_uniform(uint32_t w, int u24) {
    if (u24) {
        const float f = (float)w;
        return (double)(float)(f * 0x1.0p-32f + 0x1.0p-33f);
    }
    return ((double)w + 0.5) * 0x1.0p-32;
}

static void block_words(uint64_t seed, uint64_t path, uint32_t step, uint32_t slot, uint32_t w[4]) {
    const uint32_t ctr[4] = {(uint32_t)path, (uint32_t)(path >> 32), step, slot};
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    qko_philox4x32_10(ctr, key, w);
}

/* Box-Muller as the reference's device branch (RandNGenerator.hpp:212-221). */
static void box_muller(double u0, double u1, double *z1, double *z2) {
    const double R = sqrt(-2.0 * log(u0));
    const double th = 2.0 * M_PI * u1;
    *z1 = R * cos(th);
    *z2 = R * sin(th);
}

void qko_normals(uint64_t seed, int u24, int64_t n_paths, int32_t n_steps, double *z) {
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < n_paths; ++p) {
        for (int32_t t = 0; t < n_steps; ++t) {
            uint32_t w[4];
            block_words(seed, (uint64_t)p, (uint32_t)t, 0u, w);
            box_muller(qko_uniform(w[0], u24), qko_uniform(w[1], u24),
                       &z[(p * n_steps + t) * 2], &z[(p * n_steps + t) * 2 + 1]);
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Scheme constants, computed as the reference's constructors do.           */
/* ------------------------------------------------------------------------ */
typedef struct {
    int model, scheme, u24, n_steps;
    uint64_t seed;
    double dt, eps;
    /* Euler / Milstein (MSchemes.hpp:104-122, :199-216) */
    double one_minus_k_dt, k_theta_dt, sigma_sqrt_dt, r_minus_q_dt, half_dt, sqrt_dt;
    double rho, rho_c, quarter_sigma_sq_dt, implicit_den;
    /* QE (MSchemes.hpp:296-352) */
    double exp_minus_k_dt, theta_m_factor, s2f1, s2f2, K1, K2, K3, K4, A, psi_c;
    /* Bates (MSchemes.hpp:485-491, :535-550) */
    double exp_minus_lambda_dt, mu_j, sigma_j;
} scheme_t;

static void scheme_init(const qko_model *m, scheme_t *s) {
    memset(s, 0, sizeof(*s));
    s->model = m->model;
    s->scheme = m->scheme;
    s->u24 = m->fp32_semantics;
    s->n_steps = m->n_steps;
    s->seed = m->seed;
    s->eps = m->fp32_semantics ? 1e-6 : 1e-12;     /* MSchemes.hpp:286-287 */
    s->dt = m->T / (double)m->n_steps;
    const double dt = s->dt;
    double q = m->q;
    if (m->model == QKO_BATES) {
        /* Martingale-compensated drift q' = q + lambda*kappa_J (MSchemes.hpp:540-549). */
        const double kappa_J = exp(m->mu_j + 0.5 * m->sigma_j * m->sigma_j) - 1.0;
        q += m->lambda_j * kappa_J;
        /* lambda = 0: no jump (a threshold above every uniform; the FP32 map
         * reaches 1, the reference's uniforms stay below it). */
        s->exp_minus_lambda_dt = m->lambda_j > 0.0 ? exp(-m->lambda_j * dt) : 2.0;
        s->mu_j = m->mu_j;
        s->sigma_j = m->sigma_j;
    }
    const double k = m->kappa, theta = m->theta, sigma = m->sigma;
    s->rho = m->rho;
    s->sqrt_dt = sqrt(dt);
    s->r_minus_q_dt = (m->r - q) * dt;
    s->half_dt = 0.5 * dt;
    if (m->scheme == QKO_EULER || m->scheme == QKO_MILSTEIN) {
        s->rho_c = sqrt(1.0 - s->rho * s->rho);
        s->one_minus_k_dt = 1.0 - k * dt;
        s->k_theta_dt = k * theta * dt;
        s->sigma_sqrt_dt = sigma * s->sqrt_dt;
        s->quarter_sigma_sq_dt = 0.25 * sigma * sigma * dt;
        s->implicit_den = 1.0 / (1.0 + k * dt);
    } else {
        s->psi_c = 1.5;
        s->exp_minus_k_dt = exp(-k * dt);
        if (k < s->eps) {                                   /* MSchemes.hpp:322-326 */
            s->theta_m_factor = 0.0;
            s->s2f1 = sigma * sigma * dt;
            s->s2f2 = 0.0;
        } else {
            const double ome = 1.0 - s->exp_minus_k_dt;
            s->theta_m_factor = theta * ome;
            const double sig2 = sigma * sigma;
            s->s2f1 = (sig2 * s->exp_minus_k_dt / k) * ome;
            s->s2f2 = (theta * sig2 / (2.0 * k)) * (ome * ome);
        }
        double rho_div_sigma;
        if (sigma < s->eps) {                               /* MSchemes.hpp:336-341 */
            rho_div_sigma = 0.0;
            s->rho = 0.0;
        } else {
            rho_div_sigma = s->rho / sigma;
        }
        const double one_m_rho2 = 1.0 - s->rho * s->rho;
        const double K12 = k * rho_div_sigma - 0.5;
        const double g1 = 0.5, g2 = 0.5;
        s->K1 = dt * g1 * K12 - rho_div_sigma;
        s->K2 = dt * g2 * K12 + rho_div_sigma;
        s->K3 = dt * g1 * one_m_rho2;
        s->K4 = dt * g2 * one_m_rho2;
        s->A = s->K2 + s->K4 * 0.5;
    }
}

/* One step of the Heston core.  w are the words of step t (see simulate_one). */
static void heston_step(const scheme_t *s, const uint32_t w[4], double *S, double *v) {
    double Z1, Z2;
    box_muller(qko_uniform(w[0], s->u24), qko_uniform(w[1], s->u24), &Z1, &Z2);
    const double v_n = *v;
    if (s->scheme == QKO_EULER || s->scheme == QKO_MILSTEIN) {
        /* MSchemes.hpp:133-155 (Euler) and :229-252 (implicit Milstein). */
        const double Z_v = Z1;
        const double Z_S = s->rho * Z1 + s->rho_c * Z2;
        const double sqrt_v = sqrt(v_n);
        const double expo = s->r_minus_q_dt - (v_n * s->half_dt) + (sqrt_v * s->sqrt_dt) * Z_S;
        *S = *S * exp(expo);
        double v_np1;
        if (s->scheme == QKO_EULER) {
            v_np1 = v_n * s->one_minus_k_dt + s->k_theta_dt + (s->sigma_sqrt_dt * sqrt_v) * Z_v;
        } else {
            const double corr = s->quarter_sigma_sq_dt * (Z_v * Z_v - 1.0);
            v_np1 = s->implicit_den * (v_n + s->k_theta_dt + (s->sigma_sqrt_dt * sqrt_v) * Z_v + corr);
        }
        *v = fmax(v_np1, 0.0);
        return;
    }
    /* Andersen QE, MSchemes.hpp:363-441.  The exponential branch draws its
     * uniform from the step's third word w[2] instead of Phi(Z_V) (same
     * distribution). */
    const double Z_V = Z1, Z_indep = Z2;
    const double m = v_n * s->exp_minus_k_dt + s->theta_m_factor;
    const double s2 = v_n * s->s2f1 + s->s2f2;
    double v_np1, M;
    if (m <= s->eps) {
        v_np1 = 0.0;
        M = 1.0;
    } else {
        const double psi = s2 / (m * m);
        if (psi <= s->eps) {
            v_np1 = m;
            M = 1.0;
        } else if (psi <= s->psi_c) {
            const double inv_psi = 1.0 / psi;
            const double b2 = 2.0 * inv_psi - 1.0 + sqrt(2.0 * inv_psi * (2.0 * inv_psi - 1.0));
            const double a = m / (1.0 + b2);
            const double bz = sqrt(b2) + Z_V;
            v_np1 = a * bz * bz;
            const double safe_A = fmin(s->A, (0.5 / a) - s->eps);
            const double tAa = 1.0 - 2.0 * safe_A * a;
            M = exp((safe_A * b2 * a) / tAa) / sqrt(tAa);
        } else {
            const double p = (psi - 1.0) / (psi + 1.0);
            const double beta = (1.0 - p) / m;
            const double U = qko_uniform(w[2], s->u24);
            if (U <= p) {
                v_np1 = 0.0;
            } else {
                const double safe_U = fmin(U, 1.0 - s->eps);
                v_np1 = log((1.0 - p) / (1.0 - safe_U)) / beta;
            }
            const double safe_A = fmin(s->A, beta - s->eps);
            M = p + (beta * (1.0 - p)) / (beta - safe_A);
        }
    }
    const double K0 = -log(M) - (s->K1 + s->K3 * 0.5) * v_n;
    const double isig = s->K3 * v_n + s->K4 * v_np1;
    const double expo = s->r_minus_q_dt + K0 + (s->K1 * v_n + s->K2 * v_np1) + sqrt(isig) * Z_indep;
    *S = *S * exp(expo);
    *v = v_np1;
}

/* Extra Bates uniform e_j of (path, step). */
static double bates_extra(const scheme_t *s, uint64_t path, uint32_t step, uint32_t j) {
    uint32_t w[4];
    block_words(s->seed, path, step, 1u + j / 4u, w);
    return qko_uniform(w[j % 4u], s->u24);
}

/* Compound-Poisson log-normal jumps, MSchemes.hpp:493-527 (Knuth's method). */
static void bates_jump(const scheme_t *s, const uint32_t w[4], uint64_t path, uint32_t step, double *S) {
    int N = 0;
    double p = qko_uniform(w[3], s->u24);
    uint32_t j = 0;
    while (p >= s->exp_minus_lambda_dt) {
        ++N;
        p *= bates_extra(s, path, step, j++);
    }
    if (N > 0) {
        const double e0 = bates_extra(s, path, step, j);
        const double e1 = bates_extra(s, path, step, j + 1);
        const double Z = sqrt(-2.0 * log(e0)) * cos(2.0 * M_PI * e1);
        const double mu = (double)N * s->mu_j;
        const double sg = sqrt((double)N) * s->sigma_j;
        *S *= exp(mu + sg * Z);
    }
}

/* The per-path loop of MarkovianIntegrationKernel (MSolver.hpp:98-130) with the
 * PayoffTracker state of Payoff.hpp:85-140 (sum excludes S0; max/min start at S0). */
static void simulate_one(const scheme_t *s, double S0, double v0, uint64_t path,
                         double *S_store, int64_t ld, double *S_T, double *v_T,
                         double *S_sum, double *S_max, double *S_min) {
    double S = S0, v = v0, sum = 0.0, mx = S0, mn = S0;
    /* Words of step t.  Heston Euler / Milstein: two per step, steps 2i and
     * 2i+1 take words (0, 1) and (2, 3) of block (path, i, 0).  Heston QE:
     * three per step (the third for the exponential branch), words 3t..3t+2
     * of the sequence that lists blocks (path, 0, 0), (path, 1, 0), ... word
     * by word.  Bates: block (path, t, 0), word 2 for QE, word 3 the first
     * Poisson uniform. */
    const int heston = s->model == QKO_HESTON;
    const int paired = heston && (s->scheme == QKO_EULER || s->scheme == QKO_MILSTEIN);
    const int triple = heston && s->scheme == QKO_QE;
    for (int32_t t = 0; t < s->n_steps; ++t) {
        uint32_t w[4];
        if (paired) {
            uint32_t b[4];
            block_words(s->seed, path, (uint32_t)t >> 1, 0u, b);
            w[0] = b[(t & 1) * 2];
            w[1] = b[(t & 1) * 2 + 1];
            w[2] = w[3] = 0u;
        } else if (triple) {
            uint32_t b[8];
            const uint32_t k = 3u * (uint32_t)t;
            block_words(s->seed, path, k / 4u, 0u, b);
            block_words(s->seed, path, k / 4u + 1u, 0u, b + 4);
            for (int j = 0; j < 3; ++j) w[j] = b[k % 4u + (uint32_t)j];
            w[3] = 0u;
        } else {
            block_words(s->seed, path, (uint32_t)t, 0u, w);
        }
        heston_step(s, w, &S, &v);
        if (s->model == QKO_BATES) bates_jump(s, w, path, (uint32_t)t, &S);
        if (S_store) S_store[(int64_t)t * ld] = S;
        sum += S;
        mx = fmax(mx, S);
        mn = fmin(mn, S);
    }
    if (S_T) *S_T = S;
    if (v_T) *v_T = v;
    if (S_sum) *S_sum = sum;
    if (S_max) *S_max = mx;
    if (S_min) *S_min = mn;
}

int qko_simulate(const qko_model *m, int64_t path0, int64_t n_paths, double *S_store,
                 double *S_T, double *v_T, double *S_sum, double *S_max, double *S_min) {
    if (!m || n_paths < 0 || m->n_steps < 1 || m->n_steps > (1 << 30)) return -1;
    scheme_t s;
    scheme_init(m, &s);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n_paths; ++i) {
        simulate_one(&s, m->S0, m->v0, (uint64_t)(path0 + i), S_store ? S_store + i : NULL, n_paths,
                     S_T ? S_T + i : NULL, v_T ? v_T + i : NULL, S_sum ? S_sum + i : NULL,
                     S_max ? S_max + i : NULL, S_min ? S_min + i : NULL);
    }
    return 0;
}

/* Payoff.hpp:41-72 (vanilla) and :157-242 (evaluate_final_payoff). */
static double vanilla(int right, double S, double K) {
    if (right == QKO_CALL) return S > K ? S - K : 0.0;
    return K > S ? K - S : 0.0;
}

double qko_payoff_eval(const qko_payoff *p, double S, double S_sum, double S_max, double S_min,
                       int32_t n_steps) {
    const double K = p->K, B = p->B;
    const int call = p->right == QKO_CALL;
    switch (p->type) {
    case QKO_EUROPEAN: return vanilla(p->right, S, K);
    case QKO_ASIAN: return vanilla(p->right, S_sum / (double)n_steps, K);
    case QKO_BARRIER_UO: return S_max < B ? vanilla(p->right, S, K) : 0.0;
    case QKO_BARRIER_UI: return S_max >= B ? vanilla(p->right, S, K) : 0.0;
    case QKO_BARRIER_DO: return S_min > B ? vanilla(p->right, S, K) : 0.0;
    case QKO_BARRIER_DI: return S_min <= B ? vanilla(p->right, S, K) : 0.0;
    case QKO_LOOKBACK_FIXED: return call ? vanilla(QKO_CALL, S_max, K) : vanilla(QKO_PUT, S_min, K);
    case QKO_LOOKBACK_FLOAT: return call ? vanilla(QKO_CALL, S, S_min) : vanilla(QKO_PUT, S, S_max);
    case QKO_BINARY_CASH: return call ? (S > K ? 1.0 : 0.0) : (S < K ? 1.0 : 0.0);
    case QKO_BINARY_ASSET: return call ? (S > K ? S : 0.0) : (S < K ? S : 0.0);
    default: return NAN;
    }
}

/* Chan et al. pairwise update, as the reference's host merge (OptionPricer.hpp:152-172). */
void qko_chan_merge(double *n_a, double *mean_a, double *m2_a, double n_b, double mean_b, double m2_b) {
    if (n_b <= 0.0) return;
    const double n = *n_a + n_b;
    const double delta = mean_b - *mean_a;
    *mean_a = (*mean_a * *n_a + mean_b * n_b) / n;
    *m2_a += m2_b + delta * delta * (*n_a * n_b / n);
    *n_a = n;
}

#define TILE 64

/* Statistics over a payoff array: two-pass moments per 64-path tile
 * (OptionPricer.hpp:129-146 restated per tile) merged in tile order. */
static void moments(const double *x, int64_t n, double eps, double *mean, double *m2, int64_t *itm) {
    double N = 0.0, M = 0.0, Q = 0.0;
    int64_t cnt = 0;
    for (int64_t t0 = 0; t0 < n; t0 += TILE) {
        const int64_t t1 = t0 + TILE < n ? t0 + TILE : n;
        double s = 0.0;
        for (int64_t i = t0; i < t1; ++i) {
            s += x[i];
            if (x[i] > eps) ++cnt;
        }
        const double nt = (double)(t1 - t0);
        const double mt = s / nt;
        double q = 0.0;
        for (int64_t i = t0; i < t1; ++i) q += (x[i] - mt) * (x[i] - mt);
        qko_chan_merge(&N, &M, &Q, nt, mt, q);
    }
    *mean = M;
    *m2 = Q;
    *itm = cnt;
}

/* OptionPricer::evaluate_option_price (OptionPricer.hpp:204-249) and
 * get_scaling_factor (:323-332). */
static void finish(qko_stats *o, int64_t n, double mean, double m2, int64_t itm, double disc, double scale) {
    const double N = (double)n;
    const double var = N > 1.0 ? m2 / (N - 1.0) : 0.0;
    o->price = mean * disc * scale;
    o->se = sqrt(var / N) * disc * scale;
    o->prob_itm = (double)itm / N;
    o->mean = mean;
    o->m2 = m2;
    o->n = n;
    o->itm = itm;
}

int qko_price(const qko_model *m, int64_t n_paths, const qko_payoff *pay, int32_t n_pay,
              double price_scale, qko_stats *out) {
    return qko_price_window(m, 0, n_paths, pay, n_pay, price_scale, out);
}

int qko_price_window(const qko_model *m, int64_t path0, int64_t n_paths, const qko_payoff *pay, int32_t n_pay,
                     double price_scale, qko_stats *out) {
    if (!m || !pay || !out || path0 < 0 || n_paths < 1 || n_pay < 1 || m->n_steps < 1 || m->n_steps > (1 << 30))
        return -1;
    double *ST = malloc(sizeof(double) * n_paths * 4);
    double *x = malloc(sizeof(double) * n_paths);
    if (!ST || !x) { free(ST); free(x); return -2; }
    double *sum = ST + n_paths, *mx = ST + 2 * n_paths, *mn = ST + 3 * n_paths;
    qko_simulate(m, path0, n_paths, NULL, ST, NULL, sum, mx, mn);
    const double eps = m->fp32_semantics ? 1e-6 : 1e-12;      /* OptionPricer.hpp:316 */
    for (int32_t j = 0; j < n_pay; ++j) {
        for (int64_t i = 0; i < n_paths; ++i)
            x[i] = qko_payoff_eval(&pay[j], ST[i], sum[i], mx[i], mn[i], m->n_steps);
        double mean, m2;
        int64_t itm;
        moments(x, n_paths, eps, &mean, &m2, &itm);
        const double scale = pay[j].type == QKO_BINARY_CASH ? 1.0 : price_scale;
        finish(&out[j], n_paths, mean, m2, itm, exp(-m->r * m->T), scale);
    }
    free(ST);
    free(x);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Longstaff-Schwartz (LSMAmerican.hpp, ForwardBackwardMC.hpp:204-255).     */
/* ------------------------------------------------------------------------ */
void qko_basis(double X, int32_t nb, int32_t weighted, double *L) {
    /* Laguerre polynomials L0..L(nb-1) in Horner form, LSMAmerican.hpp:172-198
     * (nb = 3, 4, 5 for the reference's LaguerreP02/P03/P04).  weighted: times
     * exp(-X/2), the weighted Laguerre functions of Longstaff & Schwartz (2001),
     * which the reference does not offer. */
    L[0] = 1.0;
    if (nb > 1) L[1] = 1.0 - X;
    if (nb > 2) L[2] = ((0.5 * X - 2.0) * X) + 1.0;
    if (nb > 3) L[3] = (((-1.0 / 6.0 * X + 1.5) * X - 3.0) * X) + 1.0;
    if (nb > 4) L[4] = ((((1.0 / 24.0 * X - 2.0 / 3.0) * X + 3.0) * X - 4.0) * X) + 1.0;
    if (weighted) {
        const double w = exp(-0.5 * X);
        for (int k = 0; k < nb; ++k) L[k] *= w;
    }
}

void qko_laguerre4(double X, double L[4]) { qko_basis(X, 4, 0, L); }

/* solve_system_on_host_cholesky (LSMAmerican.hpp:516-581) and
 * compute_fallback_coefficients (:605-613) for nb basis functions;
 * sums = {count, M[nb*nb] row-major, B[nb]}. */
int qko_lsm_solve_n(const double *sums, int32_t nb, double *C) {
    const double count = sums[0];
    const double *Ms = sums + 1, *Bs = sums + 1 + nb * nb;
    for (int i = 0; i < nb; ++i) C[i] = 0.0;
    if (count < (double)nb) {
        if (count > 0.0) C[0] = Bs[0] / count;
        return 1;
    }
    double M[QKO_MAX_BASIS][QKO_MAX_BASIS], L[QKO_MAX_BASIS][QKO_MAX_BASIS] = {{0}};
    for (int i = 0; i < nb; ++i)
        for (int j = 0; j < nb; ++j) M[i][j] = Ms[i * nb + j] + (i == j ? 1e-8 : 0.0);
    for (int i = 0; i < nb; ++i) {
        for (int j = 0; j <= i; ++j) {
            double s = 0.0;
            for (int k = 0; k < j; ++k) s += L[i][k] * L[j][k];
            if (i == j) {
                const double val = M[i][i] - s;
                if (val < 1e-14) {
                    for (int q = 0; q < nb; ++q) C[q] = 0.0;
                    C[0] = Bs[0] / count;
                    return 1;
                }
                L[i][i] = sqrt(val);
            } else {
                L[i][j] = (M[i][j] - s) / L[j][j];
            }
        }
    }
    double y[QKO_MAX_BASIS];
    for (int i = 0; i < nb; ++i) {
        double s = 0.0;
        for (int k = 0; k < i; ++k) s += L[i][k] * y[k];
        y[i] = (Bs[i] - s) / L[i][i];
    }
    for (int i = nb - 1; i >= 0; --i) {
        double s = 0.0;
        for (int k = i + 1; k < nb; ++k) s += L[k][i] * C[k];
        C[i] = (y[i] - s) / L[i][i];
    }
    return 0;
}

int qko_lsm_solve(const double sums[21], double C[4]) { return qko_lsm_solve_n(sums, 4, C); }

int qko_price_american(const qko_model *m, int64_t n_paths, int32_t right, double K,
                       double price_scale, qko_stats *out, double *coeffs_out, double *cf_out) {
    return qko_price_american_ex(m, n_paths, right, K, 4, 0, price_scale, out, coeffs_out, cf_out);
}

int qko_price_american_ex(const qko_model *m, int64_t n_paths, int32_t right, double K, int32_t nb,
                          int32_t weighted, double price_scale, qko_stats *out, double *coeffs_out,
                          double *cf_out) {
    if (!m || !out || n_paths < 1 || m->n_steps < 1 || m->n_steps > (1 << 30) || nb < 3 || nb > QKO_MAX_BASIS) return -1;
    const int32_t N = m->n_steps;
    double *S = malloc(sizeof(double) * n_paths * (size_t)N);
    double *cf = malloc(sizeof(double) * n_paths);
    if (!S || !cf) { free(S); free(cf); return -2; }
    qko_simulate(m, 0, n_paths, S, NULL, NULL, NULL, NULL, NULL);
    const double disc = exp(-m->r * (m->T / (double)N));
    const int ns = 1 + nb * nb + nb;
    for (int64_t i = 0; i < n_paths; ++i) cf[i] = vanilla(right, S[(int64_t)(N - 1) * n_paths + i], K);
    for (int32_t t = N - 2; t >= 0; --t) {
        const double *St = S + (int64_t)t * n_paths;
        double sums[1 + QKO_MAX_BASIS * QKO_MAX_BASIS + QKO_MAX_BASIS] = {0};
        for (int64_t i = 0; i < n_paths; ++i) {          /* LSMAmerican.hpp:443-471 */
            const double intrinsic = vanilla(right, St[i], K);
            if (intrinsic > 0.0) {
                double L[QKO_MAX_BASIS];
                qko_basis(St[i] / K, nb, weighted, L);
                const double Y = cf[i] * disc;
                for (int a = 0; a < nb; ++a) {
                    sums[1 + nb * nb + a] += L[a] * Y;
                    for (int b = 0; b < nb; ++b) sums[1 + a * nb + b] += L[a] * L[b];
                }
                sums[0] += 1.0;
            }
        }
        (void)ns;
        double C[QKO_MAX_BASIS];
        qko_lsm_solve_n(sums, nb, C);
        if (coeffs_out) memcpy(coeffs_out + (int64_t)t * nb, C, sizeof(double) * (size_t)nb);
        for (int64_t i = 0; i < n_paths; ++i) {          /* LSMAmerican.hpp:318-339 */
            const double intrinsic = vanilla(right, St[i], K);
            const double dcf = cf[i] * disc;
            if (intrinsic > 0.0) {
                double L[QKO_MAX_BASIS];
                qko_basis(St[i] / K, nb, weighted, L);
                double cont = 0.0;
                for (int a = 0; a < nb; ++a) cont += C[a] * L[a];
                cf[i] = intrinsic > cont ? intrinsic : dcf;
            } else {
                cf[i] = dcf;
            }
        }
    }
    /* apply_final_discount, LSMAmerican.hpp:393-419. */
    double tot = 0.0;
    for (int64_t i = 0; i < n_paths; ++i) {
        cf[i] *= disc;
        tot += cf[i];
    }
    const double today = vanilla(right, m->S0, K);
    if (today > tot / (double)n_paths)
        for (int64_t i = 0; i < n_paths; ++i) cf[i] = today;
    double mean, m2;
    int64_t itm;
    moments(cf, n_paths, m->fp32_semantics ? 1e-6 : 1e-12, &mean, &m2, &itm);
    finish(out, n_paths, mean, m2, itm, 1.0, price_scale);   /* no e^{-rT}: OptionPricer.hpp:211-216 */
    if (cf_out) memcpy(cf_out, cf, sizeof(double) * (size_t)n_paths);
    free(S);
    free(cf);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Greek stencils, MCEngine.hpp:576-735 (each bump a full repricing under    */
/* the same random stream).                                                  */
/* ------------------------------------------------------------------------ */
static double price_one(const qko_model *m, int64_t n, const qko_payoff *p, double scale) {
    qko_stats s;
    qko_price(m, n, p, 1, scale, &s);
    return s.price;
}

int qko_greeks_eval(const qko_model *m, int64_t n, const qko_payoff *p, const qko_bumps *b,
                    double price_scale, qko_greeks *g) {
    if (!m || !p || !b || !g) return -1;
    memset(g, 0, sizeof(*g));
    qko_stats base;
    qko_price(m, n, p, 1, price_scale, &base);
    g->base_price = base.price;
    g->base_se = base.se;
    int np = 1;
    const double P0 = base.price;
    const double h_S = m->S0 * b->rel_S;
    const double h_S_real = price_scale * m->S0 * b->rel_S;
    const double vol0 = sqrt(m->v0);
    const double vu = (vol0 + b->abs_vol) * (vol0 + b->abs_vol);
    const double vd = (vol0 - b->abs_vol) * (vol0 - b->abs_vol);
    if (b->delta_gamma) {
        qko_model u = *m, d = *m;
        u.S0 += h_S; d.S0 -= h_S;
        const double pu = price_one(&u, n, p, price_scale), pd = price_one(&d, n, p, price_scale);
        g->delta = (pu - pd) / (2.0 * h_S_real);
        g->gamma = (pu - 2.0 * P0 + pd) / (h_S_real * h_S_real);
        np += 2;
    }
    if (b->vega_vomma) {
        qko_model u = *m, d = *m;
        u.v0 = vu; d.v0 = vd;
        const double pu = price_one(&u, n, p, price_scale), pd = price_one(&d, n, p, price_scale);
        g->vega = (pu - pd) / (2.0 * b->abs_vol);
        g->vomma = (pu - 2.0 * P0 + pd) / (b->abs_vol * b->abs_vol);
        np += 2;
    }
    if (b->rho) {
        qko_model u = *m, d = *m;
        u.r += b->abs_r; d.r -= b->abs_r;
        g->rho = (price_one(&u, n, p, price_scale) - price_one(&d, n, p, price_scale)) / (2.0 * b->abs_r);
        np += 2;
    }
    if (b->theta) {
        if (m->T > b->abs_T) {
            qko_model d = *m;
            d.T = m->T - b->abs_T;
            const double pd = price_one(&d, n, p, price_scale);
            if (b->theta_central) {
                qko_model u = *m;
                u.T = m->T + b->abs_T;
                g->theta = (pd - price_one(&u, n, p, price_scale)) / (2.0 * b->abs_T);
                np += 1;
            } else {
                g->theta = (pd - P0) / b->abs_T;
            }
            np += 1;
        }
    }
    if (b->vanna) {
        qko_model uu = *m, du = *m, ud = *m, dd = *m;
        uu.S0 += h_S; uu.v0 = vu;
        du.S0 -= h_S; du.v0 = vu;
        ud.S0 += h_S; ud.v0 = vd;
        dd.S0 -= h_S; dd.v0 = vd;
        const double puu = price_one(&uu, n, p, price_scale), pdu = price_one(&du, n, p, price_scale);
        const double pud = price_one(&ud, n, p, price_scale), pdd = price_one(&dd, n, p, price_scale);
        g->vanna = (puu - pdu - pud + pdd) / (4.0 * h_S_real * b->abs_vol);
        np += 4;
    }
    g->n_pricings = np;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Closed forms (include/core/analytical).                                         */
/* ------------------------------------------------------------------------ */

/* 64-point Gauss-Legendre nodes/weights by Newton iteration on P_64
 * (Abramowitz & Stegun 25.4.29); the reference tabulates the same rule
 * (Integration.hpp:39-84). */
void qko_gl64(double x[64], double w[64]) {
    const int n = 64;
    for (int i = 0; i < n / 2; ++i) {
        double z = cos(M_PI * (i + 0.75) / (n + 0.5)), pp = 0.0;
        for (int it = 0; it < 100; ++it) {
            double p1 = 1.0, p2 = 0.0;
            for (int j = 1; j <= n; ++j) {
                const double p3 = p2;
                p2 = p1;
                p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
            }
            pp = n * (z * p1 - p2) / (z * z - 1.0);
            const double dz = p1 / pp;
            z -= dz;
            if (fabs(dz) < 1e-16) break;
        }
        {
            double p1 = 1.0, p2 = 0.0;
            for (int j = 1; j <= n; ++j) {
                const double p3 = p2;
                p2 = p1;
                p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
            }
            pp = n * (z * p1 - p2) / (z * z - 1.0);
        }
        x[i] = -z;
        x[n - 1 - i] = z;
        w[i] = w[n - 1 - i] = 2.0 / ((1.0 - z * z) * pp * pp);
    }
}

/* Exponent of the Heston characteristic function in the Albrecher form
 * (HestonExact.hpp:56-102; Rouah eqs. 1.48, 2.14-2.17, 2.54). */
static double complex heston_exponent(double phi, int j, double S0, double T, double r, double q,
                                      double v0, double k, double theta, double sigma, double rho) {
    const double complex Iu = _Complex_I;
    const double u = j == 1 ? 0.5 : -0.5;
    const double b = j == 1 ? k - rho * sigma : k;
    const double complex t1 = (rho * sigma * Iu * phi - b) * (rho * sigma * Iu * phi - b);
    const double complex t2 = sigma * sigma * (2.0 * u * Iu * phi - phi * phi);
    const double complex d = csqrt(t1 - t2);
    const double complex c = (b - rho * sigma * Iu * phi - d) / (b - rho * sigma * Iu * phi + d);
    const double complex edT = cexp(-d * T);
    const double complex C = (r - q) * Iu * phi * T +
        (k * theta) / (sigma * sigma) * ((b - rho * sigma * Iu * phi - d) * T - 2.0 * clog((1.0 - c * edT) / (1.0 - c)));
    const double complex D = (b - rho * sigma * Iu * phi - d) / (sigma * sigma) * ((1.0 - edT) / (1.0 - c * edT));
    return C + D * v0 + Iu * phi * log(S0);
}

typedef struct {
    double S0, K, T, r, q, v0, k, theta, sigma, rho, lam, muJ, sigJ;
    int bates;
} cf_args;

static double integrand(double phi, const cf_args *a, int j) {
    const double complex Iu = _Complex_I;
    double complex e = heston_exponent(phi, j, a->S0, a->T, a->r, a->q, a->v0, a->k, a->theta, a->sigma, a->rho);
    if (a->bates) {
        /* Merton jump exponent E_j, BatesExact.hpp:60-88. */
        const double kbar = exp(a->muJ + 0.5 * a->sigJ * a->sigJ) - 1.0;
        const double uj = j == 1 ? 1.0 : 0.0;
        const double muj = j == 1 ? a->muJ + a->sigJ * a->sigJ : a->muJ;
        const double complex inner = cexp(Iu * phi * muj - 0.5 * a->sigJ * a->sigJ * phi * phi) - 1.0;
        e += a->lam * a->T * (-Iu * phi * kbar + pow(1.0 + kbar, uj) * inner);
    }
    const double complex f = cexp(e);
    return creal(cexp(-Iu * phi * log(a->K)) * f / (Iu * phi));
}

static double probability(const cf_args *a, int j, double phi_max, double panel) {
    double x[64], w[64];
    qko_gl64(x, w);
    double total = 0.0;
    const int n_panels = (int)(phi_max / panel);
    for (int pi = 0; pi < n_panels; ++pi) {
        const double lo = pi * panel, hi = (pi + 1) * panel;
        const double sc = 0.5 * (hi - lo), sh = 0.5 * (hi + lo);
        double s = 0.0;
        for (int i = 0; i < 64; ++i) s += w[i] * integrand(sc * x[i] + sh, a, j);
        total += s * sc;
    }
    return 0.5 + total / M_PI;
}

double qko_heston_call(double S0, double K, double T, double r, double q, double v0,
                       double kappa, double theta, double sigma, double rho) {
    /* get_exact_eu_call_option_price: one 64-point panel on [0, 100] (HestonExact.hpp:174-195). */
    cf_args a = {S0, K, T, r, q, v0, kappa, theta, sigma, rho, 0, 0, 0, 0};
    const double P1 = probability(&a, 1, 100.0, 100.0), P2 = probability(&a, 2, 100.0, 100.0);
    return S0 * exp(-q * T) * P1 - K * exp(-r * T) * P2;
}

double qko_bates_call(double S0, double K, double T, double r, double q, double v0,
                      double kappa, double theta, double sigma, double rho,
                      double lambda_j, double mu_j, double sigma_j) {
    /* Panels of width 50 up to phi = 8000 (BatesExact.hpp:96-114, :130-145). */
    cf_args a = {S0, K, T, r, q, v0, kappa, theta, sigma, rho, lambda_j, mu_j, sigma_j, 1};
    const double P1 = probability(&a, 1, 8000.0, 50.0), P2 = probability(&a, 2, 8000.0, 50.0);
    return S0 * exp(-q * T) * P1 - K * exp(-r * T) * P2;
}

double qko_bs_call(double S0, double K, double T, double r, double q, double vol) {
    /* BlackScholesExact.hpp:46-65. */
    const double d1 = (log(S0 / K) + (r - q + 0.5 * vol * vol) * T) / (vol * sqrt(T));
    const double d2 = d1 - vol * sqrt(T);
    const double N1 = 0.5 * (1.0 + erf(d1 / sqrt(2.0))), N2 = 0.5 * (1.0 + erf(d2 / sqrt(2.0)));
    return S0 * exp(-q * T) * N1 - K * exp(-r * T) * N2;
}
