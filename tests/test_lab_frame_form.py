"""The lab-frame closed form of the energy-pass pair terms (CPU only).

Without permanent dipoles the FP32 energy/force kernels (elec_far32_lab, elec_near32_lab in
tdp_kernels.hip) do not build the reference's QI frame: they evaluate a closed form in the
frame's X axis alone, derived from calculateElectrostaticPairIxn
(ReferenceTholeDipoleForce.cpp:129-366) with its index pattern (:150-159, :161-185).  This test
states that closed form in numpy, in the kernels' lane view (the lane's own site m, the other
site o, s = +1 when the lane's particle has the lower caller index), and checks it against the
oracle's pair function for both orientations, Mutual and Direct polarization, undamped and
Thole-damped pairs and separations close to the x axis.  The GPU parity tests then hold the
kernels to the FP64 path and the oracle.
"""
import numpy as np
import pytest

import oracle


def lab_form(qm, qo, m, o, d, am, ao, s, mutual, damped):
    """Energy and force on the lane's particle; d = x_other - x_me (minimum image)."""
    dx, dy, dz = d
    rho2 = dy * dy + dz * dz
    r2 = dx * dx + rho2
    rinv = 1.0 / np.sqrt(r2)
    rI2 = rinv * rinv
    rI3 = rI2 * rinv
    rI4 = rI2 * rI2
    if damped:
        u = r2 * rinv / (am * ao)
        au3 = min(min(am, ao) * u ** 3, 100.0)
        ev = np.exp(-au3)
    else:
        au3, ev = 0.0, 0.0
    thc = 1.0 - ev                              # thole_c = thole_d1
    th0 = 1.0 - ev * (1.0 + 1.5 * au3)          # thole_d0 = dthole_c
    dd1 = 1.0 - ev * (1.0 + au3)                # dthole_d1
    dd0 = dd1 - ev * 1.5 * au3 * au3            # dthole_d0
    kx = rinv / np.sqrt(rho2)
    X = np.array([rho2 * kx, -dx * dy * kx, -dx * dz * kx])
    Xm = X @ m
    Xo = X @ o
    qq = qm * qo
    xw = qm * Xo - qo * Xm
    t3 = rI3 * thc
    if mutual:
        c6 = 4.0 * th0 + 2.0 * thc
        a = s * c6 * rI4
        am_c = a * Xo - t3 * qo
        ao_c = a * Xm + t3 * qm
        inner = rI2 * qq + rI4 * (6.0 * dd1 * (m @ o) - (12.0 * dd0 + 6.0 * dd1 + 2.0 * c6) * Xm * Xo)
    else:
        am_c = -t3 * qo
        ao_c = t3 * qm
        inner = rI2 * qq
    lam = s * inner - rI3 * (2.0 * th0 + thc) * xw
    G = m * am_c + o * ao_c + X * lam
    e = rinv * qq - 0.5 * s * rI2 * thc * xw
    return e, -np.array([G[1], G[2], G[0]])


def _pairs(rng, n, rmin, rmax):
    for _ in range(n):
        d = rng.normal(size=3)
        d *= rng.uniform(rmin, rmax) / np.linalg.norm(d)
        yield rng.uniform(-1, 1, size=3), d


@pytest.mark.parametrize("mutual", [True, False])
@pytest.mark.parametrize("regime", ["undamped", "damped", "near_x_axis"])
def test_lab_form_matches_oracle_pair(mutual, regime):
    rng = np.random.default_rng(7 if mutual else 8)
    pol = 1 if mutual else 0   # TDPO_MUTUAL, TDPO_DIRECT
    zero = np.zeros(3)
    worst = 0.0
    if regime == "undamped":
        pairs = list(_pairs(rng, 300, 1.2, 3.0))   # a u^3 >= 50 for a <= 0.45: no damping, as on far tiles
    elif regime == "damped":
        pairs = list(_pairs(rng, 300, 0.12, 0.7))
    else:   # nearly along x: 1 - Z_x^2 down to 1e-12
        pairs = []
        for k in range(60):
            eps = 10.0 ** (-rng.uniform(0, 6))
            d = np.array([rng.choice([-1, 1]) * rng.uniform(0.5, 2.0), eps * rng.normal(), eps * rng.normal()])
            pairs.append((rng.uniform(-1, 1, size=3), d))
    for p0, d in pairs:
        qi, qk = rng.uniform(-1, 1, size=2)
        ui, uk = rng.normal(size=3) * 0.02, rng.normal(size=3) * 0.02
        ai, ak = rng.uniform(0.3, 0.45, size=2)
        pi, pk = p0, p0 + d
        e_ref, fk, ti, tk = oracle.pair_electrostatic(pi, pk, qi, qk, zero, zero, ui, uk, ai, ak, 1.0, 1.0, pol)
        assert np.all(ti == 0) and np.all(tk == 0)
        damped = regime != "undamped"
        # lane = I (lower index, s = +1): force on I = -f_K
        e1, f1 = lab_form(qi, qk, ui, uk, d, ai, ak, 1.0, mutual, damped)
        # lane = K (s = -1): force on K = f_K, separation seen from K
        e2, f2 = lab_form(qk, qi, uk, ui, -d, ak, ai, -1.0, mutual, damped)
        fs = np.abs(fk).max() + 1e-300
        es = abs(e_ref) + 1e-300
        worst = max(worst, abs(e1 - e_ref) / es, abs(e2 - e_ref) / es,
                    np.abs(f1 + fk).max() / fs, np.abs(f2 - fk).max() / fs)
    # measured: <= 3e-11 (rounding of the two FP64 evaluations); a wrong sign, factor or index is O(1)
    assert worst < 1e-9, worst
