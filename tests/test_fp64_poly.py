"""The FP64 device polynomials (qk_device.hpp, Math<double>) against the
functions they approximate, at 50 digits.

The coefficients and the log table are read from the header as written, so a
mistyped digit, a dropped term or a reversed order fails here.  The bounds are
the near-minimax fit error plus the rounding of the coefficients to double:
  log1p: r + r^2 Q(r) vs log1p(r) on |r| <= 0.009                   < 2e-17 relative
  sin:   S(w) vs sin(pi/32 sqrt w)/sqrt w on w in [0, 1/4]          < 6e-17 relative
  cos:   w M(w) vs cos(pi/32 sqrt w) - 1                            < 2e-19 absolute
  exp:   E(r) vs e^r on r in [-0.35, 0.35]                          < 1e-17 relative
the sin/cos table (hi + lo = sin, cos of 2 pi j/64 to 2^-100), the log
table (each entry's interval of m gives |m invc - 1| <= 0.009, invc
has at most 7 significant bits so that m invc - 1 is exact in one fma, and
lhi + llo = -log(invc) to 2^-76), and the split of ln 2 into LN2_HI (32
significant bits) + LN2_LO.
"""
import os
import re

import pytest

mp = pytest.importorskip("mpmath")

HEADER = os.path.join(os.path.dirname(__file__), "..", "quantkos_amd", "csrc", "qk_device.hpp")
FIT = os.path.join(os.path.dirname(__file__), "..", "quantkos_amd", "csrc", "fit_fp64_poly.py")


def header_array(name):
    text = open(HEADER).read()
    m = re.search(r"\b%s\[(\d+)\]\s*=\s*\{([^}]*)\}" % name, text)
    assert m, name
    vals = [float.fromhex(t.strip()) for t in m.group(2).split(",") if t.strip()]
    assert len(vals) == int(m.group(1))
    return vals


def header_scalar(name):
    m = re.search(r"\b%s\s*=\s*(-?0x[0-9a-fA-Fp.+-]+);" % name, open(HEADER).read())
    assert m, name
    return float.fromhex(m.group(1))


def poly(c, x):
    p = mp.mpf(c[-1])
    for ci in reversed(c[:-1]):
        p = p * x + mp.mpf(ci)
    return p


def grid(hi, n=1500):
    return [hi * mp.mpf(i) / n for i in range(n + 1)]


def test_log1p_polynomial():
    mp.mp.dps = 50
    c = header_array("LOG1P_Q")
    worst = 0
    for i in range(-1500, 1501):
        if i == 0:
            continue
        r = mp.mpf("0.009") * i / 1500
        worst = max(worst, abs((r + r * r * poly(c, r)) / mp.log1p(r) - 1))
    assert worst < 2e-17, worst


def header_log_table():
    text = open(HEADER).read()
    m = re.search(r"\bLOG_TAB\[(\d+)\]\s*=\s*\{(.*?)\n\};", text, re.S)
    assert m
    rows = re.findall(r"\{\s*([^,}]+),\s*([^,}]+?)f,\s*([^,}]+?)f\s*\}", m.group(2))
    assert len(rows) == int(m.group(1)) == 128
    return [tuple(float.fromhex(x.strip()) for x in row) for row in rows]


def test_log_table():
    import struct
    mp.mp.dps = 50
    hi_word = lambda h: struct.unpack("<d", struct.pack("<Q", h << 32))[0]
    for j, (lhi, invc, llo) in enumerate(header_log_table()):
        mant, _ = mp.frexp(mp.mpf(invc))
        assert mant * 2 ** 7 == int(mant * 2 ** 7), (j, invc)           # <= 7 significant bits
        assert abs(mp.mpf(lhi) + mp.mpf(llo) + mp.log(invc)) < mp.mpf(2) ** -76, j
        lo, hi = hi_word(0x3fe6a09e + j * 2 ** 13), hi_word(0x3fe6a09e + (j + 1) * 2 ** 13)
        assert max(abs(mp.mpf(lo) * invc - 1), abs(mp.mpf(hi) * invc - 1)) <= mp.mpf("0.009"), j
        if lo <= 1.0 < hi:
            assert invc == 1.0 and lhi == 0.0 and llo == 0.0


def test_sin_cos_polynomials():
    mp.mp.dps = 50
    cs, cm = header_array("SIN64_S"), header_array("COS64_M")
    ws, wc = 0, 0
    for w in grid(mp.mpf(1) / 4)[1:]:
        r = mp.sqrt(w)
        ws = max(ws, abs(poly(cs, w) * r / mp.sin(mp.pi / 32 * r) - 1))
        wc = max(wc, abs(w * poly(cm, w) - (mp.cos(mp.pi / 32 * r) - 1)))
    assert ws < 6e-17, ws        # the rounding of pi/32 to double (<= 2^-54) dominates
    assert wc < 2e-19, wc


def header_sincos_table():
    text = open(HEADER).read()
    m = re.search(r"\bSINCOS_TAB\[(\d+)\]\s*=\s*\{(.*?)\n\};", text, re.S)
    assert m
    rows = re.findall(r"\{([^}]*)\}", m.group(2))
    assert len(rows) == int(m.group(1)) == 64
    return [tuple(float.fromhex(x.strip()) for x in row.split(",")) for row in rows]


def test_sincos_table():
    mp.mp.dps = 50
    for j, (sh, ch, sl, cl) in enumerate(header_sincos_table()):
        a = 2 * mp.pi * j / 64
        assert abs(mp.mpf(sh) + mp.mpf(sl) - mp.sin(a)) < mp.mpf(2) ** -100, j
        assert abs(mp.mpf(ch) + mp.mpf(cl) - mp.cos(a)) < mp.mpf(2) ** -100, j


def test_exp_polynomial():
    mp.mp.dps = 50
    c = header_array("EXP_P")
    worst = 0
    for i in range(-1500, 1501):
        r = mp.mpf("0.35") * i / 1500
        worst = max(worst, abs(poly(c, r) / mp.exp(r) - 1))
    assert worst < 1e-17, worst


def test_ln2_split():
    mp.mp.dps = 50
    hi, lo = header_scalar("LN2_HI"), header_scalar("LN2_LO")
    assert abs(mp.mpf(hi) + mp.mpf(lo) - mp.log(2)) < mp.mpf(2) ** -85
    # k * LN2_HI exact for |k| <= 1100: the significand of LN2_HI fits in 32 bits
    m, e = mp.frexp(mp.mpf(hi))
    assert (m * 2 ** 32) == int(m * 2 ** 32)


def test_fit_script_reproduces_header():
    import importlib.util
    spec = importlib.util.spec_from_file_location("fit_fp64_poly", FIT)
    fit = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(fit)
    for name in ("SIN64_S", "COS64_M", "EXP_P", "LOG1P_Q"):
        assert fit.coefficients(name) == header_array(name), name
    assert fit.sincos_table() == header_sincos_table()
    tab, _ = fit.log_table()
    assert [(a, float(b), c) for a, b, c in tab] == header_log_table()
