"""Coefficients of the FP64 device polynomials in qk_device.hpp (Math<double>).

Each polynomial is the Chebyshev near-minimax fit (mpmath.chebyfit, 50 digits)
of a smooth function on a reduced interval, rounded to the nearest double:

  log:  P(z) = (2 atanh(s)/s - 2) / z,  z = s^2 in [0, (3 - 2 sqrt 2)^2]
        (log(1+f) = 2 atanh(s), s = f/(2+f), f = m - 1, m in [sqrt(1/2), sqrt(2)));
        7 coefficients, z P(z) absolute error 9e-18.
  sin:  S(w) = sin(pi/2 sqrt w) / sqrt w,  w = r^2 in [0, 1/4];  7 coefficients.
  cos:  C(w) = cos(pi/2 sqrt w),            w in [0, 1/4];       8 coefficients.
  exp:  E(r) = e^r on r in [-0.35, 0.35] (|r| <= ln2/2 after the reduction);
        12 coefficients, relative error 5e-18.

    python quantkos_amd/csrc/fit_fp64_poly.py      # prints the C initialisers

tests/test_fp64_poly.py evaluates the coefficients written in qk_device.hpp
against these functions at 50 digits.
"""
import mpmath as mp

mp.mp.dps = 50
Z_MAX = (3 - 2 * mp.sqrt(2)) ** 2
W_MAX = mp.mpf(1) / 4


def log_p(z):
    if z == 0:
        return mp.mpf(2) / 3
    s = mp.sqrt(z)
    return (2 * mp.atanh(s) / s - 2) / z


def sin_s(w):
    if w == 0:
        return mp.pi / 2
    t = mp.sqrt(w)
    return mp.sin(mp.pi / 2 * t) / t


def cos_c(w):
    return mp.cos(mp.pi / 2 * mp.sqrt(w))


R_EXP = mp.mpf("0.35")
FITS = {"LOG_P": (log_p, 0, Z_MAX, 7), "SIN_S": (sin_s, 0, W_MAX, 7), "COS_C": (cos_c, 0, W_MAX, 8),
        "EXP_P": (mp.exp, -R_EXP, R_EXP, 12)}


def coefficients(name):
    """Coefficients, lowest order first, rounded to double."""
    f, lo, hi, n = FITS[name]
    c, _ = mp.chebyfit(f, [lo, hi], n, error=True)
    return [float(x) for x in reversed(c)]


def main():
    for name in FITS:
        cs = coefficients(name)
        print("%s[%d] = {%s};" % (name, len(cs), ", ".join(x.hex() for x in cs)))


if __name__ == "__main__":
    main()
