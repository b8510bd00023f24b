"""The rational approximation behind the FP64 erfc of the Ewald option's real-space passes
(kErfcxP / kErfcxQ in tdp_ewald.hip): e^{-x^2} P(x) / Q(x), evaluated in FP64 as the kernels do
(Horner with fused multiply-adds, one division), against SciPy's erfc on [0, 6]."""
import os
import re

import numpy as np
import pytest

SRC = os.path.join(os.path.dirname(__file__), "..", "tholedipoleplugin_amd", "csrc", "tdp_ewald.hip")


def _coeffs(name):
    text = open(SRC).read()
    m = re.search(r"constexpr double %s\[10\] = \{([^}]*)\}" % name, text)
    assert m, name
    return [float(v) for v in m.group(1).replace("\n", " ").split(",")]


def _fma(a, b, c):
    # FP64 fused multiply-add, exactly rounded: a*b + c in extended precision, then rounded once
    return np.float64(np.longdouble(a) * np.longdouble(b) + np.longdouble(c))


def test_erfc_rational_matches_library():
    scipy_special = pytest.importorskip("scipy.special")
    P, Q = _coeffs("kErfcxP"), _coeffs("kErfcxQ")
    assert P[0] == 1.0 and Q[0] == 1.0 and len(P) == len(Q) == 10
    x = np.concatenate([np.linspace(0.0, 6.0, 60001), np.array([1e-300, 1e-8, 0.5, 2.9184, 6.0])])
    p = np.full_like(x, P[9])
    q = np.full_like(x, Q[9])
    for k in range(8, -1, -1):
        p = _fma(p, x, P[k])
        q = _fma(q, x, Q[k])
    e = np.exp(-x * x)
    got = e * p / q
    want = scipy_special.erfc(x)
    rel = np.abs(got - want) / want
    # e^{-x^2} and erfc carry a few ulps each; the rational part adds below 1e-15
    assert rel.max() < 4e-15, rel.max()
    assert np.all(q > 0.0)


def test_erfc_rational_is_erfcx():
    """Closed forms: erfcx(0) = 1, erfcx'(0) = -2/sqrt(pi) (P_1 - Q_1), and erfcx(x) ~ 1/(x sqrt(pi))
    is monotone decreasing on [0, 6]."""
    P, Q = _coeffs("kErfcxP"), _coeffs("kErfcxQ")
    assert abs((P[1] - Q[1]) + 2.0 / np.sqrt(np.pi)) < 1e-12
    x = np.linspace(0.0, 6.0, 2001)
    r = np.polyval(P[::-1], x) / np.polyval(Q[::-1], x)
    assert np.all(np.diff(r) < 0.0)


def test_erfc_asymptotic_branch():
    """Beyond x = 6 the kernels take e^{-x^2} / (x sqrt(pi)) (1 - t/2 + 3t^2/4 - 15t^3/8), t = 1/x^2:
    relative error below 105 / (16 x^8) against SciPy's erfc."""
    scipy_special = pytest.importorskip("scipy.special")
    x = np.linspace(6.0, 26.0, 20001)
    t = 1.0 / (x * x)
    got = np.exp(-x * x) * (0.56418958354775628 / x) * (1.0 - t * (0.5 - t * (0.75 - 1.875 * t)))
    want = scipy_special.erfc(x)
    rel = np.abs(got - want) / want
    assert np.all(rel < 105.0 / (16.0 * x ** 8) + 1e-13), rel.max()
    assert rel[0] > 1e-7   # the bound is near-tight at x = 6 (a dropped term would show)
    src = open(SRC).read()
    assert "e * (kInvSqrtPi / x) * (1.0 - t * (0.5 - t * (0.75 - 1.875 * t)))" in src
    assert "constexpr double kInvSqrtPi = 0.56418958354775628;" in src
