"""Coefficients of the FP64 device polynomials in qk_device.hpp (Math<double>).

Each polynomial is the Chebyshev near-minimax fit (mpmath.chebyfit, 50 digits)
of a smooth function on a reduced interval, rounded to the nearest double:

  sin/cos(2 pi u) = table entry j (angle 2 pi j/64) combined with the angle
        2 pi r/64, 64 u = j + r, |r| <= 1/2:
        SIN64_S(w) = sin(pi/32 sqrt w) / sqrt w,       w = r^2 in [0, 1/4];  4 coefficients,
        COS64_M(w) = (cos(pi/32 sqrt w) - 1) / w;                              4 coefficients,
        SINCOS_TAB[j] = {sin, cos}(2 pi j / 64) as double + double remainder.
  exp:  E(r) = e^r on r in [-0.35, 0.35] (|r| <= ln2/2 after the reduction);
        12 coefficients, relative error 5e-18.
  log1p for the table-driven log:  Q(r) = (log1p(r) - r) / r^2 on |r| <= R_LOG
        (6 coefficients), and LOG_TAB: for m in [sqrt(1/2), sqrt(2)) split into 128
        intervals by the high word of m (steps of 2^13 from 0x3fe6a09e), invc = 1/c
        at the interval's centre c rounded to 7 significant bits (1 on the intervals
        within R_LOG of 1), so r = m invc - 1 is exact in one fma, and -log(invc) as
        a double plus a float remainder.

    python quantkos_amd/csrc/fit_fp64_poly.py      # prints the C initialisers

tests/test_fp64_poly.py evaluates the coefficients written in qk_device.hpp
against these functions at 50 digits.
"""
import mpmath as mp

mp.mp.dps = 50
W_MAX = mp.mpf(1) / 4


def sin64_s(w):
    if w == 0:
        return mp.pi / 32
    t = mp.sqrt(w)
    return mp.sin(mp.pi / 32 * t) / t


def cos64_m(w):
    if w == 0:
        return -(mp.pi / 32) ** 2 / 2
    return (mp.cos(mp.pi / 32 * mp.sqrt(w)) - 1) / w


def sincos_table():
    """[(sin_hi, cos_hi, sin_lo, cos_lo)] of the angles 2 pi j / 64, j = 0..63."""
    out = []
    for j in range(64):
        sv, cv = mp.sin(2 * mp.pi * j / 64), mp.cos(2 * mp.pi * j / 64)
        sh, ch = float(sv), float(cv)
        out.append((sh, ch, float(sv - sh), float(cv - ch)))
    return out


def log1p_q(r):
    if abs(r) < mp.mpf(10) ** -20:
        return -mp.mpf(1) / 2 + r / 3 - r ** 2 / 4
    return (mp.log1p(r) - r) / r ** 2


LOG_TAB_N, LOG_TAB_BASE, LOG_TAB_STEP = 128, 0x3fe6a09e, 1 << 13


def _hi_word_double(h):
    import struct
    return struct.unpack("<d", struct.pack("<Q", h << 32))[0]


def log_table():
    """[(lhi, invc, llo)] per interval, and the largest |r| = |m invc - 1|."""
    import numpy as np
    out, rmax = [], mp.mpf(0)
    for j in range(LOG_TAB_N):
        lo = _hi_word_double(LOG_TAB_BASE + j * LOG_TAB_STEP)
        hi = _hi_word_double(LOG_TAB_BASE + (j + 1) * LOG_TAB_STEP)
        if max(abs(lo - 1.0), abs(hi - 1.0)) <= R_LOG:
            invc = 1.0      # near 1 the result is log1p(m - 1) alone: no cancellation
        else:
            m, e = mp.frexp(2 / (mp.mpf(lo) + mp.mpf(hi)))
            invc = float(mp.ldexp(mp.nint(m * 2 ** 7), e - 7))
        rmax = max(rmax, abs(mp.mpf(lo) * invc - 1), abs(mp.mpf(hi) * invc - 1))
        L = -mp.log(mp.mpf(invc))
        lhi = float(L)
        llo = float(np.float32(float(L - mp.mpf(lhi))))
        out.append((lhi, invc, llo))
    return out, rmax


R_LOG = mp.mpf("0.009")
R_EXP = mp.mpf("0.35")
FITS = {"SIN64_S": (sin64_s, 0, W_MAX, 4), "COS64_M": (cos64_m, 0, W_MAX, 4),
        "EXP_P": (mp.exp, -R_EXP, R_EXP, 12), "LOG1P_Q": (log1p_q, -R_LOG, R_LOG, 6)}


def coefficients(name):
    """Coefficients, lowest order first, rounded to double."""
    f, lo, hi, n = FITS[name]
    c, _ = mp.chebyfit(f, [lo, hi], n, error=True)
    return [float(x) for x in reversed(c)]


def main():
    for name in FITS:
        cs = coefficients(name)
        print("%s[%d] = {%s};" % (name, len(cs), ", ".join(x.hex() for x in cs)))
    print("SINCOS_TAB[64] = {")
    for e in sincos_table():
        print("    {%s}," % ", ".join(x.hex() for x in e))
    print("};")
    tab, rmax = log_table()
    print("// LOG_TAB: max |r| = %s" % mp.nstr(rmax, 6))
    print("LOG_TAB[%d] = {" % len(tab))
    for lhi, invc, llo in tab:
        print("    {%s, %sf, %sf}," % (lhi.hex(), float(invc).hex(), float(llo).hex()))
    print("};")


if __name__ == "__main__":
    main()
