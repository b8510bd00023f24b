"""Pins of the oracle's Ewald sums (oracle/tdp_ewald.c), CPU only.

The reference has no Ewald summation (its PME method is the minimum image over all pairs), so the
oracle of the library's Ewald option restates published methods; these checks fix what a
plausible restatement mistake would break:

* the FFT against numpy.fft, and the B-splines against the closed form of the cardinal B-spline
  and its derivatives;
* the charge-charge sum against the Madelung constant of rock salt (1.747564594633182);
* the dipole-dipole sum, self-field included, against the Lorentz field of a cubic lattice of
  parallel dipoles in tin-foil boundary conditions, E = 4 pi mu / (3 V);
* real and reciprocal space together by the independence of the converged energy, forces and
  induced dipoles from the splitting parameter (exact reciprocal sums);
* smooth PME against the exact reciprocal sum as the grid is refined;
* the forces against central differences of the energy (induced dipoles re-solved at each
  displaced geometry), for Direct and Mutual polarization, with the water topology's scaled pairs.
"""
import math
from fractions import Fraction

import numpy as np
import pytest

import oracle
from tholedipoleplugin_amd import systems
from tholedipoleplugin_amd.force import TholeDipoleForce

T = TholeDipoleForce


def test_fft_matches_numpy():
    rng = np.random.default_rng(0)
    for shape in [(6, 7, 10), (24, 4, 9), (11, 8, 5), (16, 16, 3)]:
        a = rng.normal(size=shape) + 1j * rng.normal(size=shape)
        assert np.allclose(oracle.fft3(a, -1), np.fft.fftn(a), rtol=0, atol=1e-11 * np.abs(a).sum() / a.size * 10)
        assert np.allclose(oracle.fft3(a, +1), np.fft.ifftn(a) * a.size, rtol=0,
                           atol=1e-11 * np.abs(a).sum() / a.size * 10)


def _cardinal_bspline(x, p):
    """M_p(x) = 1/(p-1)! sum_k (-1)^k C(p,k) (x-k)_+^(p-1), in exact rational arithmetic (the
    alternating sum cancels badly in floating point)."""
    x = Fraction(x)
    tot = sum((-1) ** k * math.comb(p, k) * max(x - k, Fraction(0)) ** (p - 1) for k in range(p + 1))
    return float(tot / math.factorial(p - 1))


@pytest.mark.parametrize("p", [4, 5, 6])
def test_bspline_closed_form_and_derivatives(p):
    for w in (0.0, 0.13, 0.5, 0.77, 0.999):
        th, d1, d2 = oracle.bspline(w, p)
        expect = [_cardinal_bspline(w + p - 1 - j, p) for j in range(p)]
        assert np.allclose(th, expect, rtol=0, atol=1e-14)
        assert abs(th.sum() - 1.0) < 1e-14 and abs(d1.sum()) < 1e-14 and abs(d2.sum()) < 1e-13
        if w == 0.0:
            continue   # knots: the second derivative of M_(p-1) jumps there, central differences do not apply
        h = 1e-5
        up = [_cardinal_bspline(w + h + p - 1 - j, p) for j in range(p)]
        dn = [_cardinal_bspline(w - h + p - 1 - j, p) for j in range(p)]
        assert np.allclose(d1, (np.array(up) - np.array(dn)) / (2 * h), rtol=0, atol=1e-8)
        up1 = [_cardinal_bspline(w + h + p - 1 - j, p - 1) - _cardinal_bspline(w + h + p - 2 - j, p - 1)
               for j in range(p)]
        dn1 = [_cardinal_bspline(w - h + p - 1 - j, p - 1) - _cardinal_bspline(w - h + p - 2 - j, p - 1)
               for j in range(p)]
        assert np.allclose(d2, (np.array(up1) - np.array(dn1)) / (2 * h), rtol=0, atol=1e-7)


def _force_of(charges, pol, thole=0.0, pbc_box=1.0, polarization=T.Direct):
    f = T()
    f.setNonbondedMethod(T.PME)
    f.setPolarizationType(polarization)
    for q, a in zip(charges, pol):
        f.addParticle(float(q), [0.0, 0.0, 0.0], float(a), float(thole))
    return f, np.diag([pbc_box] * 3)


@pytest.mark.parametrize("recip", ["exact", "spme"])
def test_madelung_rock_salt(recip):
    a = 1.0
    na = [(0, 0, 0), (0, .5, .5), (.5, 0, .5), (.5, .5, 0)]
    cl = [(.5, 0, 0), (0, .5, 0), (0, 0, .5), (.5, .5, .5)]
    pos = np.array(na + cl, float) * a + 0.0123
    f, box = _force_of([1] * 4 + [-1] * 4, [0.0] * 8, pbc_box=a)
    if recip == "exact":
        r = oracle.ewald_evaluate(f, pos, box, 0.49, 12.0, kmax=(24, 24, 24), threads=2)
        tol = 1e-10
    else:
        r = oracle.ewald_evaluate(f, pos, box, 0.49, 10.0, grid=(96, 96, 96), order=5, threads=2)
        tol = 1e-7
    expect = -8 * 1.747564594633182 / a
    assert abs(r["energy"] - expect) <= tol * abs(expect), (r["energy"], expect)
    assert np.abs(r["forces"]).max() <= (1e-10 if recip == "exact" else 1e-4)   # centres of symmetry


@pytest.mark.parametrize("n_side", [1, 2])
def test_dipole_lattice_lorentz_field(n_side):
    """Identical dipoles on a simple cubic lattice (one per cell, the box holding n_side^3 cells):
    with tin-foil boundaries the field at every dipole is 4 pi mu / (3 v), v the cell volume.  The
    exact reciprocal sum gives it to 1e-10; smooth PME converges to it as the grid is refined (its
    error falls by more than ten when the grid spacing halves)."""
    cell = 0.7
    L = cell * n_side
    g = (np.arange(n_side) + 0.25) * cell
    pos = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    n = pos.shape[0]
    f, box = _force_of([0.0] * n, [1e-3] * n, pbc_box=L)
    mu = np.tile([0.011, -0.023, 0.017], (n, 1))
    expect = 4 * math.pi / (3 * cell ** 3) * mu
    a = 6.0 / (0.499 * L)   # erfc(a r_c) ~ 2e-17
    e0, tmu = oracle.ewald_fields(f, pos, mu, box, cutoff=0.499 * L, alpha=a, kmax=(21, 21, 21), threads=2)
    assert np.allclose(tmu, expect, rtol=0, atol=1e-10 * np.abs(expect).max()), (tmu[0], expect[0])
    assert np.abs(e0).max() == 0.0
    errs = []   # the grid error is relative to the self-field 4 a^3 mu / (3 sqrt(pi)): a L = 8 here
    for K in (32 * n_side, 64 * n_side):
        _, tmu = oracle.ewald_fields(f, pos, mu, box, cutoff=0.499 * L, alpha=8.0 / L, grid=(K, K, K), order=5,
                                     threads=2)
        errs.append(np.abs(tmu - expect).max() / np.abs(expect).max())
    assert errs[1] < 2e-3 and errs[1] < errs[0] / 8, errs


def _small_water(polarization, seed=11):
    s = systems.water_box(3, 8, 1.2, seed=seed, cutoff=0.59)
    s.force.setPolarizationType(polarization)
    s.force.setMutualInducedTargetEpsilon(1e-14)
    s.force.setMutualInducedMaxIterations(500)
    return s


@pytest.mark.parametrize("polarization", [T.Direct, T.Mutual])
def test_energy_independent_of_splitting(polarization):
    """Exact reciprocal sums at two splitting parameters (real-space terms negligible at the cutoff
    for both): energy, forces and induced dipoles agree, with the water topology's scaled pairs and
    Thole damping."""
    s = _small_water(polarization)
    r1 = oracle.ewald_evaluate(s.force, s.positions, s.box, 0.59, 9.0, kmax=(18, 18, 18), threads=4)
    r2 = oracle.ewald_evaluate(s.force, s.positions, s.box, 0.59, 11.5, kmax=(23, 23, 23), threads=4)
    assert r1["converged"] == 1 and r2["converged"] == 1
    assert abs(r1["energy"] - r2["energy"]) <= 1e-10 * abs(r1["energy"])
    assert np.allclose(r1["forces"], r2["forces"], rtol=0, atol=1e-9 * np.abs(r1["forces"]).max())
    assert np.allclose(r1["induced"], r2["induced"], rtol=0, atol=1e-10 * np.abs(r1["induced"]).max())
    assert np.abs(r1["induced"]).max() > 1e-4   # the solve did something


def test_spme_converges_to_exact_sum():
    s = _small_water(T.Mutual)
    ex = oracle.ewald_evaluate(s.force, s.positions, s.box, 0.59, 6.0, kmax=(16, 16, 16), threads=4)
    errs = []
    for K in (16, 32, 64):
        sp = oracle.ewald_evaluate(s.force, s.positions, s.box, 0.59, 6.0, grid=(K, K, K), order=5, threads=4)
        de = abs(sp["energy"] - ex["energy"]) / abs(ex["energy"])
        df = np.sqrt(np.mean((sp["forces"] - ex["forces"]) ** 2)) / np.sqrt(np.mean(ex["forces"] ** 2))
        errs.append((de, df))
    assert errs[2][0] < 1e-7 and errs[2][1] < 1e-5, errs
    for k in (1, 2):   # order-5 convergence: at least 8x per halving of the grid spacing
        assert errs[k][0] < errs[k - 1][0] / 8 and errs[k][1] < errs[k - 1][1] / 8, errs


@pytest.mark.parametrize("polarization", [T.Direct, T.Mutual])
def test_forces_are_minus_energy_gradient(polarization):
    s = _small_water(polarization, seed=12)
    kw = dict(cutoff=0.59, alpha=5.0, grid=(20, 20, 20), order=5, threads=4)
    base = oracle.ewald_evaluate(s.force, s.positions, s.box, **kw)
    rng = np.random.default_rng(3)
    h = 2e-6
    checked = 0
    for i in rng.choice(s.n, 6, replace=False):
        for c in range(3):
            p1 = s.positions.copy()
            p2 = s.positions.copy()
            p1[i, c] += h
            p2[i, c] -= h
            e1 = oracle.ewald_evaluate(s.force, p1, s.box, **kw)["energy"]
            e2 = oracle.ewald_evaluate(s.force, p2, s.box, **kw)["energy"]
            fd = -(e1 - e2) / (2 * h)
            assert abs(fd - base["forces"][i, c]) <= 2e-6 * np.abs(base["forces"]).max(), (i, c, fd,
                                                                                         base["forces"][i, c])
            checked += 1
    assert checked == 18
