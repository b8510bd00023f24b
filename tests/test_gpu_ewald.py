"""The Ewald option (TDP_FLAG_EWALD) against the oracle's Ewald sums, through the C ABI (GPU).

The oracle (oracle/tdp_ewald.c) is pinned on the CPU by tests/test_ewald_oracle.py.  Here the GPU
runs the same model with the same Ewald parameter, grid and cutoff:

* the pairs within the cutoff are the oracle's brute-force set exactly (bit-exact FP64 criterion),
  including pairs placed one rounding step inside and outside the cutoff and half a box apart;
* the GPU's induced dipoles satisfy mu = alpha (E0 + T mu) with E0 and T mu from the oracle, and the
  energy and forces at those dipoles match the oracle's;
* full evaluations match the oracle's (solve included) at the parity bar: energy relative 1e-6,
  force relative RMS 1e-5, induced dipoles 1e-6 D; FP64 is held far tighter where it can be.
"""
import os

import numpy as np
import pytest

import oracle
from tholedipoleplugin_amd import _lib, systems
from tholedipoleplugin_amd.context import Context, System
from tholedipoleplugin_amd.force import OpenMMException, TholeDipoleForce

pytestmark = pytest.mark.gpu

T = TholeDipoleForce
DEBYE = 0.020819434
E_TOL, F_TOL, MU_TOL = 1e-6, 1e-5, 1e-6 * DEBYE
THREADS = max(1, min(16, os.cpu_count() or 1))


def _say(msg):
    print(msg, flush=True)


def rel_rms(a, b):
    a = np.asarray(a).reshape(-1)
    b = np.asarray(b).reshape(-1)
    return float(np.sqrt(np.mean((a - b) ** 2)) / max(np.sqrt(np.mean(b ** 2)), 1e-300))


def _context(s, precision="double", solver="pcg"):
    sysm = System()
    for m in s.masses:
        sysm.addParticle(m)
    sysm.setDefaultPeriodicBoxVectors(*s.box)
    sysm.addForce(s.force)
    ctx = Context(sysm, precision=precision, solver=solver, ewald=True)
    ctx.setPeriodicBoxVectors(*s.box)
    ctx.setPositions(s.positions)
    return ctx


def _params(s, ctx):
    a, nx, ny, nz = s.force.getPMEParametersInContext(ctx)
    return dict(cutoff=s.force.getCutoffDistance(), alpha=a, grid=(nx, ny, nz), order=5)


def _box648(polarization=T.Mutual, eps=1e-11):
    s = systems.config0()
    s.force.setEwaldErrorTolerance(5e-4)
    s.force.setPMEParameters(np.sqrt(-np.log(2 * 5e-4)) / 0.9, 24, 24, 24)   # alpha from the tolerance
    s.force.setPolarizationType(polarization)
    s.force.setMutualInducedTargetEpsilon(eps)
    s.force.setMutualInducedMaxIterations(500)
    return s


def _edge_system():
    """Pairs at the cutoff's edge: each second particle sits one rounding step inside or outside
    the cutoff from its partner along one axis, or exactly half a box away."""
    L, rc = 2.0, 0.8
    f = T()
    f.setNonbondedMethod(T.PME)
    f.setCutoffDistance(rc)
    pos = []
    rng = np.random.default_rng(5)
    for k in range(24):
        base = rng.uniform(0.0, L, 3)
        d = [np.nextafter(rc, 0.0), rc, np.nextafter(rc, 1.0), 0.5 * L][k % 4]
        other = base.copy()
        other[k % 3] += d if k % 2 == 0 else -d
        pos += [base, other]
    for _ in range(len(pos)):
        f.addParticle(0.1, [0.0, 0.0, 0.0], 1e-3, 0.39)
    return f, np.array(pos), np.diag([L, L, L])


def test_neighbor_pairs_bit_exact():
    cases = []
    s = _box648()
    cases.append(("box648", s.force, s.positions, s.box))
    f, p, b = _edge_system()
    cases.append(("edges", f, p, b))
    s1 = systems.config1()
    cases.append(("box36864", s1.force, s1.positions, s1.box))
    for name, f, pos, box in cases:
        h = _lib.HipContextHandle(f, box, precision="double", solver="pcg", flags=_lib.FLAG_EWALD)
        got = h.neighbor_pairs(pos, box)
        h.close()
        ref = oracle.cutoff_pairs(f, pos, box, f.getCutoffDistance())
        _say("  %s: %d pairs within the cutoff (oracle %d)" % (name, len(got), len(ref)))
        assert got.shape == ref.shape and np.array_equal(got, ref), name


@pytest.mark.parametrize("case", ["box648-double", "box648-mixed", "box36864-mixed"])
def test_stored_lists_hold_every_pair(case):
    """The stored real-space blocks of T: one entry per direction of every pair within the cutoff
    whose iScale is nonzero (the oracle's scale maps).  In mixed precision the pairs closer than the
    FP64 radius (by the neighbour criterion's own FP64 r^2) are FP64 entries; the farther ones are
    FP32 entries when their tile pair takes the image of the tile origins, which on the 36,864-site
    box (4.95 nm, cutoff 1.0 nm) every neighbour tile pair does, so the split is exact there; on the
    648-site box (1.86 nm, cutoff 0.9 nm) most tile pairs fail it and keep FP64 entries.  The padded
    lists hold at least as many slots."""
    name, precision = case.split("-")
    _say("case ewald stored lists %s" % case)
    import ctypes as C
    if name == "box648":
        s = _box648()
    else:
        s = systems.config1()
    h = _lib.HipContextHandle(s.force, s.box, precision=precision, solver="pcg", flags=_lib.FLAG_EWALD)
    h.execute(s.positions, s.box)
    info = h.ewald_info()
    h.close()
    pairs = oracle.cutoff_pairs(s.force, s.positions, s.box, s.force.getCutoffDistance())
    os_ = oracle.OracleSystem(s.force)
    L = oracle.lib()
    m, sc = C.c_double(), C.c_double()

    def scale(i, j):
        assert L.tdpo_pair_scales(C.byref(os_.s), int(i), int(j), C.byref(m), C.byref(sc)) == 0
        return sc.value

    # water boxes: covalent lists stay within a molecule (sites 3k, 3k+1, 3k+2), so pairs of
    # different molecules have iScale 1 (checked on a sample)
    keep = np.ones(len(pairs), bool)
    same = np.flatnonzero(pairs[:, 0] // 3 == pairs[:, 1] // 3)
    for k in same:
        keep[k] = scale(*pairs[k]) != 0.0
    rng = np.random.default_rng(7)
    other = np.flatnonzero(pairs[:, 0] // 3 != pairs[:, 1] // 3)
    for k in rng.choice(other, size=min(2000, len(other)), replace=False):
        assert scale(*pairs[k]) == 1.0
    box = np.diag(s.box)
    d = s.positions[pairs[:, 0]] - s.positions[pairs[:, 1]]
    d -= box * np.floor(d / box + 0.5)
    r2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
    near = r2 < info["r_fp64"] * info["r_fp64"] if precision == "mixed" else np.ones(len(pairs), bool)
    want64 = 2 * int(np.count_nonzero(keep & near))
    want32 = 2 * int(np.count_nonzero(keep & ~near))
    _say("  %s: %d FP32 + %d FP64 entries (FP64 radius split: %d + %d), slots %d + %d, %d pairs with iScale 0"
         % (case, info["entries32"], info["entries64"], want32, want64, info["slots32"], info["slots64"],
            int(np.count_nonzero(~keep))))
    assert info["stored"] == 1
    assert info["entries32"] + info["entries64"] == want32 + want64
    assert info["entries64"] >= want64
    if name == "box36864" or precision == "double":
        assert info["entries32"] == want32 and info["entries64"] == want64
    assert info["slots32"] >= info["entries32"] and info["slots64"] >= info["entries64"]
    assert np.count_nonzero(~keep) > 0 and (precision == "double" or want32 > want64 > 0)


@pytest.mark.parametrize("polarization", [T.Mutual, T.Direct])
def test_box648_fp64_matches_oracle(polarization):
    _say("case ewald box648 fp64 %d" % polarization)
    s = _box648(polarization)
    ctx = _context(s)
    st = ctx.getState(getForces=True, getEnergy=True)
    mu = s.force.getInducedDipoles(ctx)
    prm = _params(s, ctx)
    assert prm["grid"] == (24, 24, 24) and abs(prm["alpha"] - np.sqrt(-np.log(1e-3)) / 0.9) < 1e-12
    ref = oracle.ewald_evaluate(s.force, s.positions, s.box, threads=THREADS, **prm)
    de = abs(st.getPotentialEnergy() - ref["energy"]) / abs(ref["energy"])
    df = rel_rms(st.getForces(), ref["forces"])
    dmu = np.max(np.abs(mu - ref["induced"]))
    _say("  energy rel %.3e, force rel RMS %.3e, mu max %.3e e*nm" % (de, df, dmu))
    assert de <= 1e-9 and df <= 1e-8 and dmu <= 1e-3 * MU_TOL
    # the operator itself: the GPU's dipoles are the fixed point of the oracle's fields, and the
    # energy and forces at those dipoles are the oracle's
    e0, tmu = oracle.ewald_fields(s.force, s.positions, mu, s.box, threads=THREADS, **prm)
    alpha = s.force.packed_arrays()["polarity"][:, None]
    if polarization == T.Mutual:
        assert np.max(np.abs(mu - alpha * (e0 + tmu))) <= 1e-3 * MU_TOL
    else:
        assert np.max(np.abs(mu - alpha * e0)) <= 1e-12
    at = oracle.ewald_energy_at(s.force, s.positions, mu, s.box, threads=THREADS, **prm)
    assert abs(st.getPotentialEnergy() - at["energy"]) <= 1e-10 * abs(at["energy"])
    assert rel_rms(st.getForces(), at["forces"]) <= 1e-9


def test_box648_mixed_matches_oracle():
    _say("case ewald box648 mixed")
    s = _box648(T.Mutual)
    ctx = _context(s, precision="mixed")
    st = ctx.getState(getForces=True, getEnergy=True)
    mu = s.force.getInducedDipoles(ctx)
    ref = oracle.ewald_evaluate(s.force, s.positions, s.box, threads=THREADS, **_params(s, ctx))
    de = abs(st.getPotentialEnergy() - ref["energy"]) / abs(ref["energy"])
    df = rel_rms(st.getForces(), ref["forces"])
    dmu = np.max(np.abs(mu - ref["induced"]))
    _say("  energy rel %.3e, force rel RMS %.3e, mu max %.3e e*nm" % (de, df, dmu))
    assert de <= E_TOL and df <= F_TOL and dmu <= MU_TOL


def _scaled_far_system():
    """Random charges with covalent maps that also link particles farther apart than the cutoff,
    and a net charge (the background term)."""
    rng = np.random.default_rng(9)
    L, rc, n = 2.2, 0.75, 120
    f = T()
    f.setNonbondedMethod(T.PME)
    f.setCutoffDistance(rc)
    f.setPolarizationType(T.Mutual)
    f.setMutualInducedTargetEpsilon(1e-12)
    f.setMutualInducedMaxIterations(500)
    pos = []
    while len(pos) < n:
        x = rng.uniform(0.0, L, 3)
        if all(np.linalg.norm((x - y) - L * np.round((x - y) / L)) > 0.15 for y in pos):
            pos.append(x)
    for i in range(n):
        f.addParticle(float(rng.uniform(-0.6, 0.7)), [0.0, 0.0, 0.0], float(rng.uniform(0.4e-3, 1.2e-3)),
                      float(rng.uniform(0.3, 0.45)))
    for i in range(n):
        f.setCovalentMap(i, T.Covalent12, [int(x) for x in rng.integers(0, n, size=3)])
        f.setCovalentMap(i, T.Covalent13, [int(x) for x in rng.integers(0, n, size=2)])
    return systems.SyntheticSystem("scaled_far", f, np.array(pos), np.diag([L, L, L]), np.ones(n))


def test_scaled_pairs_beyond_cutoff_and_net_charge():
    _say("case ewald scaled pairs beyond the cutoff")
    s = _scaled_far_system()
    s.force.setPMEParameters(3.6, 30, 30, 30)
    assert abs(s.force.packed_arrays()["charge"].sum()) > 0.5
    for precision, tol in (("double", 1e-9), ("mixed", None)):
        ctx = _context(s, precision=precision)
        st = ctx.getState(getForces=True, getEnergy=True)
        mu = s.force.getInducedDipoles(ctx)
        ref = oracle.ewald_evaluate(s.force, s.positions, s.box, threads=THREADS, **_params(s, ctx))
        de = abs(st.getPotentialEnergy() - ref["energy"]) / abs(ref["energy"])
        df = rel_rms(st.getForces(), ref["forces"])
        dmu = np.max(np.abs(mu - ref["induced"]))
        _say("  %s: energy rel %.3e, force rel RMS %.3e, mu max %.3e" % (precision, de, df, dmu))
        if tol:
            assert de <= tol and df <= 10 * tol and dmu <= 1e-3 * MU_TOL
        else:
            assert de <= E_TOL and df <= F_TOL and dmu <= MU_TOL


def test_ewald_rejects_unsupported_input():
    s = _box648()
    f = s.force
    f.setParticleParameters(0, -0.834, [0.0, 0.01, 0.0], 1.44e-3, 0.39, T.NoAxisType, -1, -1, -1)
    with pytest.raises(OpenMMException, match="molecular dipoles must be zero"):
        _context(s)
    s = _box648(T.Extrapolated)
    with pytest.raises(OpenMMException, match="Direct and Mutual"):
        _context(s)
    s = _box648()
    ctx = _context(s)
    with pytest.raises(OpenMMException, match="not available with the Ewald option"):
        s.force.getElectrostaticPotential(np.zeros((2, 3)), ctx)


def test_box36864_every_site():
    """36,864 sites, cutoff 1.0 nm, 64^3 grid: the GPU's dipoles are the fixed point of the oracle's
    fields at every site and the energy and forces at those dipoles are the oracle's (FP64 solve to
    1e-11 e*nm), and the mixed-precision solve (1e-9 e*nm) agrees with the FP64 one at every site."""
    _say("case ewald box36864")
    s = systems.config1()
    s.force.setCutoffDistance(1.0)
    s.force.setPMEParameters(np.sqrt(-np.log(2 * 1e-4)) / 1.0, 64, 64, 64)   # tolerance 1e-4
    s.force.setMutualInducedTargetEpsilon(1e-11)
    s.force.setMutualInducedMaxIterations(200)
    ctx = _context(s, precision="double")
    st = ctx.getState(getForces=True, getEnergy=True)
    stats = ctx._kernel_for(s.force).solverStats()
    _say("  double stats %s" % stats)
    assert stats["converged"] == 1
    mu = s.force.getInducedDipoles(ctx)
    prm = _params(s, ctx)
    e0, tmu = oracle.ewald_fields(s.force, s.positions, mu, s.box, threads=THREADS, **prm)
    alpha = s.force.packed_arrays()["polarity"][:, None]
    resid = np.max(np.abs(mu - alpha * (e0 + tmu)))
    at = oracle.ewald_energy_at(s.force, s.positions, mu, s.box, threads=THREADS, **prm)
    de = abs(st.getPotentialEnergy() - at["energy"]) / abs(at["energy"])
    df = rel_rms(st.getForces(), at["forces"])
    _say("  every site: max |mu - alpha (E0 + T mu)| %.3e e*nm, energy rel %.3e, force rel RMS %.3e" % (resid, de, df))
    assert resid <= 1e-2 * MU_TOL and de <= 1e-9 and df <= 1e-8
    s.force.setMutualInducedTargetEpsilon(1e-9)
    cm = _context(s, precision="mixed")
    sm = cm.getState(getForces=True, getEnergy=True)
    mum = s.force.getInducedDipoles(cm)
    dmu = np.max(np.abs(mum - mu))
    dem = abs(sm.getPotentialEnergy() - st.getPotentialEnergy()) / abs(st.getPotentialEnergy())
    dfm = rel_rms(sm.getForces(), st.getForces())
    _say("  mixed vs double: mu max %.3e e*nm, energy rel %.3e, force rel RMS %.3e" % (dmu, dem, dfm))
    assert dmu <= MU_TOL and dem <= E_TOL and dfm <= F_TOL


def test_device_entry_point_and_late_list_check():
    """tdp_execute_device with the Ewald option (list counts examined at the solver's first host
    check, forces enqueued ahead of it) against tdp_execute, and a compressed box whose stored lists
    overflow on a context sized for the first box (the evaluation runs again with grown lists)."""
    import subprocess
    import sys
    _say("case ewald device entry point, late list check")
    script = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ewald_device_check.py")
    r = subprocess.run([sys.executable, "-u", script], capture_output=True, text=True, timeout=600)
    print(r.stdout[-4000:])
    print(r.stderr[-2000:])
    assert r.returncode == 0 and "OK" in r.stdout
