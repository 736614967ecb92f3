// qk_analytic.hpp -- semi-analytic European prices on the host: Heston and
// Bates by characteristic-function inversion, Black-Scholes in closed form.
//
// The reference ships these as header-only test oracles
// (include/core/analytical/HestonExact.hpp:56-195, BatesExact.hpp:49-145,
// BlackScholesExact.hpp:36-65, Integration.hpp:39-84).  Here they are part of
// the product: the engine reports every European Monte Carlo price against
// them.  The functions in the Heston / Bates / BlackScholes namespaces keep the
// reference's signatures and quadrature (one 64-point Gauss-Legendre panel on
// [0, 100] for Heston; panels of width 50 up to 8000 for Bates), so a caller of
// the reference gets the same numbers; european_price() is the engine's own
// entry, with panels of width 50 until the integrand has decayed.
#pragma once

#include <functional>

#include "qk_config.hpp"

namespace quantkos::Engine::Analytical {

// Gauss-Legendre rule of 64 points on [a, b] (Integration.hpp:39-84); nodes
// and weights by Newton's iteration on P_64, computed once.
double integrate_gl64(const std::function<double(double)> &f, double a, double b);
void gl64_rule(double x[64], double w[64]);   // nodes and weights on [-1, 1]

// Model inputs of one European contract, unnormalised or normalised alike.
struct EuropeanInputs {
    double S0, K, T, r, q, v0;
    double kappa, theta, sigma, rho;
    double lambda_j = 0.0, mu_j = 0.0, sigma_j = 0.0;   // Bates jumps (lambda_j = 0: Heston)
};

// Black-76 price of a call (right 0) or put (right 1) on a forward F with total
// variance w = vol^2 T, discounted at r.
double black76(int right, double F, double K, double w, double disc);

// European price (right: 0 call, 1 put; puts by put-call parity).  Heston /
// Bates by characteristic function: P_j = 1/2 + (1/pi) * integral over
// (0, phi_max) of Re[exp(-i phi ln K) f_j(phi) / (i phi)], by the 64-point
// rule on panels of `panel` (one panel when panel >= phi_max); by default
// panels of 50 that stop once the integrand's envelope |f_j| / phi is below
// 1e-15 S0 (at most phi = 8000).  With no vol-of-vol (sigma = 0, variance
// deterministic): the Black-Scholes price at the integrated variance
// (Heston) or Merton's jump series over it (Bates).
double european_price(const EuropeanInputs &in, int right, double phi_max = 8000.0, double panel = 50.0,
                      bool early_stop = true);

namespace Heston {
// HestonExact.hpp:174-195 (European call; conf already validated).
double get_exact_eu_call_option_price(const Config::UInputs &conf, double upper_bound = 100.0);
}
namespace Bates {
// BatesExact.hpp:130-145.
double get_exact_eu_call_option_price(const Config::UInputs &conf, double upper_bound = 8000.0);
}
namespace BlackScholes {
// BlackScholesExact.hpp:46-65: volatility sqrt(v0).
double get_exact_eu_call_option_price(const Config::UInputs &conf);
}

}  // namespace quantkos::Engine::Analytical
