// qk_analytic.cpp -- semi-analytic European prices (see qk_analytic.hpp).
#include "qk_analytic.hpp"

#include <cmath>
#include <complex>
#include <mutex>
#include <stdexcept>

#include "../../include/qk.h"

namespace quantkos::Engine::Analytical {

namespace {
using cplx = std::complex<double>;
constexpr double PI = 3.14159265358979323846;

struct Rule {
    double x[64], w[64];
    Rule() {
        // Roots of P_64 by Newton's iteration from Tricomi's estimate, weights
        // 2 / ((1 - x^2) P'_64(x)^2); the rule is symmetric.
        constexpr int n = 64;
        for (int i = 0; i < n / 2; ++i) {
            double z = std::cos(PI * (i + 0.75) / (n + 0.5)), dp = 0.0;
            for (int it = 0; it < 100; ++it) {
                double p0 = 1.0, p1 = z;
                for (int k = 2; k <= n; ++k) {
                    const double pk = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
                    p0 = p1;
                    p1 = pk;
                }
                dp = n * (z * p1 - p0) / (z * z - 1.0);
                const double dz = p1 / dp;
                z -= dz;
                if (std::fabs(dz) < 1e-16) break;
            }
            // recompute the derivative at the converged root
            double p0 = 1.0, p1 = z;
            for (int k = 2; k <= n; ++k) {
                const double pk = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = pk;
            }
            dp = n * (z * p1 - p0) / (z * z - 1.0);
            const double wt = 2.0 / ((1.0 - z * z) * dp * dp);
            x[i] = -z;
            x[n - 1 - i] = z;
            w[i] = w[n - 1 - i] = wt;
        }
    }
};

const Rule &rule() {
    static const Rule r;
    return r;
}

// log f_j(phi): the Heston characteristic exponent in the form of Albrecher,
// Mayer, Schoutens and Tistaert (2007), which stays on the principal branch
// of the logarithm, plus the Bates compound-Poisson term.  Under measure j = 1
// (share numeraire) u = 1/2, b = kappa - rho sigma, the jump intensity is
// lambda (1 + kbar) and the log-jump mean mu_j + sigma_j^2; under j = 2
// u = -1/2, b = kappa.
cplx log_cf(const EuropeanInputs &in, int j, double phi) {
    const cplx I(0.0, 1.0);
    const double u = j == 1 ? 0.5 : -0.5;
    const double b = j == 1 ? in.kappa - in.rho * in.sigma : in.kappa;
    const double s2 = in.sigma * in.sigma;
    const cplx beta = b - in.rho * in.sigma * phi * I;                // b - rho sigma i phi
    const cplx d = std::sqrt(beta * beta - s2 * (2.0 * u * phi * I - phi * phi));
    const cplx g = (beta - d) / (beta + d);
    const cplx e = std::exp(-d * in.T);
    const cplx C = (in.r - in.q) * phi * I * in.T +
                   (in.kappa * in.theta / s2) * ((beta - d) * in.T - 2.0 * std::log((1.0 - g * e) / (1.0 - g)));
    const cplx D = (beta - d) / s2 * ((1.0 - e) / (1.0 - g * e));
    cplx lf = C + D * in.v0 + phi * I * std::log(in.S0);
    if (in.lambda_j > 0.0) {
        const double sj2 = in.sigma_j * in.sigma_j;
        const double kbar = std::exp(in.mu_j + 0.5 * sj2) - 1.0;
        const double mu = j == 1 ? in.mu_j + sj2 : in.mu_j;
        const double amp = j == 1 ? 1.0 + kbar : 1.0;
        lf += in.lambda_j * in.T * (-phi * I * kbar + amp * (std::exp(phi * I * mu - 0.5 * sj2 * phi * phi) - 1.0));
    }
    return lf;
}

double integrand(const EuropeanInputs &in, int j, double phi) {
    const cplx I(0.0, 1.0);
    return std::real(std::exp(log_cf(in, j, phi) - phi * I * std::log(in.K)) / (phi * I));
}

void require_european_call(const Config::UInputs &conf, const char *what) {
    if (conf.options.opt_type != Implemented::OptType::European || conf.options.opt_right != Implemented::OptRight::Call)
        throw std::runtime_error(std::string("Must consider a European Call option for the ") + what + " !");
}

EuropeanInputs inputs_of(const Config::UInputs &conf, bool bates) {
    const auto &h = bates ? static_cast<const Config::MathModelConfig::Heston &>(conf.model.bates) : conf.model.heston;
    EuropeanInputs in{conf.market.S0, conf.options.StrikePrice, conf.time.t_end, conf.market.r, conf.market.q,
                      conf.market.v0, h.k, h.theta, h.sigma, h.rho};
    if (bates) {
        in.lambda_j = conf.model.bates.lambda_J;
        in.mu_j = conf.model.bates.mu_J;
        in.sigma_j = conf.model.bates.sigma_J;
    }
    return in;
}

double norm_cdf(double x) { return 0.5 * std::erfc(-x / std::sqrt(2.0)); }
}  // namespace

void gl64_rule(double x[64], double w[64]) {
    for (int i = 0; i < 64; ++i) {
        x[i] = rule().x[i];
        w[i] = rule().w[i];
    }
}

double integrate_gl64(const std::function<double(double)> &f, double a, double b) {
    const Rule &R = rule();
    const double h = 0.5 * (b - a), m = 0.5 * (b + a);
    double s = 0.0;
    for (int i = 0; i < 64; ++i) s += R.w[i] * f(m + h * R.x[i]);
    return s * h;
}

double black76(int right, double F, double K, double w, double disc) {
    if (w <= 0.0) return disc * (right == 0 ? std::max(F - K, 0.0) : std::max(K - F, 0.0));
    const double sd = std::sqrt(w);
    const double d1 = (std::log(F / K) + 0.5 * w) / sd, d2 = d1 - sd;
    return right == 0 ? disc * (F * norm_cdf(d1) - K * norm_cdf(d2)) : disc * (K * norm_cdf(-d2) - F * norm_cdf(-d1));
}

double european_price(const EuropeanInputs &in, int right, double phi_max, double panel, bool early_stop) {
    const double dq = std::exp(-in.q * in.T), dr = std::exp(-in.r * in.T);
    if (!(in.sigma > 0.0)) {
        // No vol-of-vol: v(t) = theta + (v0 - theta) e^{-kappa t}, so the total
        // variance is w = theta T + (v0 - theta)(1 - e^{-kappa T}) / kappa.
        const double w = in.kappa > 0.0
                             ? in.theta * in.T + (in.v0 - in.theta) * (-std::expm1(-in.kappa * in.T)) / in.kappa
                             : in.v0 * in.T;
        if (!(in.lambda_j > 0.0)) return black76(right, in.S0 * dq / dr, in.K, w, dr);
        // Merton (1976): given n jumps, ln S_T is normal with variance
        // w + n sigma_j^2 and the forward is S0 e^{(r - q - lambda kbar) T} (1 + kbar)^n.
        const double sj2 = in.sigma_j * in.sigma_j;
        const double kbar = std::exp(in.mu_j + 0.5 * sj2) - 1.0;
        const double lt = in.lambda_j * in.T;
        double pn = std::exp(-lt), acc = 0.0, mass = 0.0;
        const double F0 = in.S0 * dq / dr * std::exp(-in.lambda_j * kbar * in.T);
        for (int n = 0; n < 400; ++n) {
            acc += pn * black76(right, F0 * std::pow(1.0 + kbar, n), in.K, w + n * sj2, dr);
            mass += pn;
            if (n > lt && 1.0 - mass < 1e-17) break;
            pn *= lt / (n + 1);
        }
        return acc;
    }
    // Characteristic function, integrated in panels; with early_stop the
    // integration ends at the first panel end where the integrand's envelope
    // |f_j(phi)| / phi is below 1e-15 S0.
    const Rule &R = rule();
    const int n_pan = panel >= phi_max ? 1 : (int)(phi_max / panel);
    const double width = panel >= phi_max ? phi_max : panel;
    double P[2];
    for (int j = 1; j <= 2; ++j) {
        double integral = 0.0;
        for (int p = 0; p < n_pan; ++p) {
            const double a = width * p, h = 0.5 * width, m = a + h;
            double s = 0.0;
            for (int i = 0; i < 64; ++i) s += R.w[i] * integrand(in, j, m + h * R.x[i]);
            integral += s * h;
            if (early_stop && std::abs(std::exp(log_cf(in, j, a + width))) / (a + width) < 1e-15 * in.S0) break;
        }
        P[j - 1] = 0.5 + integral / PI;
    }
    const double call = in.S0 * dq * P[0] - in.K * dr * P[1];
    return right == 0 ? call : call - in.S0 * dq + in.K * dr;
}

namespace Heston {
double get_exact_eu_call_option_price(const Config::UInputs &conf, double upper_bound) {
    require_european_call(conf, "Heston weak convergence test");
    const EuropeanInputs in = inputs_of(conf, false);
    return european_price(in, 0, upper_bound, upper_bound, false);
}
}  // namespace Heston

namespace Bates {
double get_exact_eu_call_option_price(const Config::UInputs &conf, double upper_bound) {
    require_european_call(conf, "Bates analytical test");
    const EuropeanInputs in = inputs_of(conf, true);
    return european_price(in, 0, upper_bound, 50.0, false);
}
}  // namespace Bates

namespace BlackScholes {
double get_exact_eu_call_option_price(const Config::UInputs &conf) {
    require_european_call(conf, "Black Scholes weak convergence test");
    const double S = conf.market.S0, K = conf.options.StrikePrice, T = conf.time.t_end;
    const double dr = std::exp(-conf.market.r * T), dq = std::exp(-conf.market.q * T);
    return black76(0, S * dq / dr, K, conf.market.v0 * T, dr);
}
}  // namespace BlackScholes

}  // namespace quantkos::Engine::Analytical

// ---- C ABI ----
extern "C" qk_status qk_analytic_european(const qk_contract *c, int32_t right, double K, const qk_cf_quadrature *quad,
                                          double *price) {
    if (!c || !price) return QK_EINVAL;
    if (quad && !(quad->phi_max > 0.0 && quad->panel > 0.0 && quad->phi_max / quad->panel <= 1e5)) return QK_EINVAL;
    if (right != QK_CALL && right != QK_PUT) return QK_EINVAL;
    if (!(K > 0.0) || !(c->S0 > 0.0) || !(c->T > 0.0) || !(c->v0 >= 0.0) || !(c->kappa >= 0.0) ||
        !(c->theta >= 0.0) || !(c->sigma >= 0.0) || !(c->rho >= -1.0 && c->rho <= 1.0) || !(c->price_scale > 0.0))
        return QK_EINVAL;
    using namespace quantkos::Engine::Analytical;
    EuropeanInputs in{c->S0, K, c->T, c->r, c->q, c->v0, c->kappa, c->theta, c->sigma, c->rho};
    if (c->model == QK_MODEL_BATES) {
        if (!(c->lambda_j >= 0.0) || !(c->sigma_j >= 0.0)) return QK_EINVAL;
        in.lambda_j = c->lambda_j;
        in.mu_j = c->mu_j;
        in.sigma_j = c->sigma_j;
    } else if (c->model != QK_MODEL_HESTON) {
        return QK_EINVAL;
    }
    const double v = quad ? european_price(in, right, quad->phi_max, quad->panel, false) : european_price(in, right);
    if (!std::isfinite(v)) return QK_EINVAL;
    *price = v * c->price_scale;
    return QK_OK;
}
