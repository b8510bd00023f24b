"""HIP path vs the CPU oracle, through the C ABI.

Tolerances (from the path's parity bar): energy relative error <= 1e-6, force relative RMS
error <= 1e-5, induced dipoles max error <= 1e-6 D (2.08e-8 e*nm).  The FP64 path is expected
to agree far tighter; the assertions use the bar.  Solves compared here run to a tight target
(1e-12 e*nm) unless a test checks the reference's own stopping rule, so that the comparison is
not dominated by solver truncation.
"""
import os

import numpy as np
import pytest

import oracle
import ref_systems
from tholedipoleplugin_amd import systems
from tholedipoleplugin_amd.context import Context
from tholedipoleplugin_amd.force import OpenMMException, TholeDipoleForce

pytestmark = pytest.mark.gpu

T = TholeDipoleForce
DEBYE = 0.020819434
E_TOL = 1e-6
F_TOL = 1e-5
MU_TOL = 1e-6 * DEBYE


def _say(msg):
    print(msg, flush=True)


def rel_rms(a, b):
    a = np.asarray(a).reshape(-1)
    b = np.asarray(b).reshape(-1)
    return float(np.sqrt(np.mean((a - b) ** 2)) / max(np.sqrt(np.mean(b ** 2)), 1e-300))


def gpu_eval(rs, precision="double", solver="jacobi"):
    ctx = Context(rs.system, precision=precision, solver=solver)
    if rs.box is not None:
        ctx.setPeriodicBoxVectors(*rs.box)
    ctx.setPositions(rs.positions)
    st = ctx.getState(getForces=True, getEnergy=True)
    k = ctx._kernel_for(rs.force)
    return ctx, st, k


def tighten(rs, eps=1e-12, maxit=1000):
    rs.force.setMutualInducedTargetEpsilon(eps)
    rs.force.setMutualInducedMaxIterations(maxit)
    return rs


CASES = [("ammonia", T.Direct), ("ammonia", T.Mutual), ("ammonia", T.Extrapolated), ("water4", T.Direct),
         ("water4", T.Mutual), ("water4", T.Extrapolated), ("triclinic", T.Mutual), ("zbisect", T.Mutual),
         ("zbisect", T.Direct), ("zonly", T.Mutual), ("plasma", T.Mutual)]


@pytest.mark.parametrize("name,pol", CASES)
def test_reference_systems_fp64(name, pol):
    _say("case %s pol=%d" % (name, pol))
    rs = tighten(ref_systems.ALL[name](pol))
    ctx, st, k = gpu_eval(rs)
    ref = oracle.force_energy(rs.force, rs.positions, rs.box)
    e, f = st.getPotentialEnergy(), st.getForces()
    assert abs(e - ref["energy"]) <= E_TOL * max(abs(ref["energy"]), 1e-12)
    assert rel_rms(f, ref["forces"]) <= F_TOL
    ind = rs.force.getInducedDipoles(ctx)
    assert np.max(np.abs(ind - ref["induced"])) <= MU_TOL
    lab = rs.force.getLabFramePermanentDipoles(ctx)
    assert np.max(np.abs(lab - ref["labdip"])) <= 1e-12
    tot = rs.force.getTotalDipoles(ctx)
    assert np.max(np.abs(tot - (ref["labdip"] + ref["induced"]))) <= MU_TOL
    stats = k.solverStats()
    if pol == T.Mutual:
        assert stats["converged"] == 1


@pytest.mark.parametrize("name", ["ammonia", "water4", "triclinic"])
def test_reference_stopping_rule(name):
    """With the system's own target and the Jacobi solver, the iteration count and the final
    epsilon follow convergeInducedDipolesByDIIS (ReferenceTholeDipoleForce.cpp:829-850)."""
    _say("case stopping %s" % name)
    rs = ref_systems.ALL[name](T.Mutual)
    ctx, st, k = gpu_eval(rs)
    ref = oracle.force_energy(rs.force, rs.positions, rs.box)
    stats = k.solverStats()
    assert stats["iterations"] == ref["iterations"]
    assert abs(stats["epsilon"] - ref["epsilon"]) <= 1e-6 * ref["epsilon"]
    assert rel_rms(st.getForces(), ref["forces"]) <= F_TOL


@pytest.mark.parametrize("solver,precision", [("pcg", "double"), ("pcg", "mixed"), ("jacobi", "mixed")])
@pytest.mark.parametrize("name", ["ammonia", "water4", "triclinic", "zbisect"])
def test_solvers_and_precision(name, solver, precision):
    _say("case %s %s %s" % (name, solver, precision))
    rs = tighten(ref_systems.ALL[name](T.Mutual), eps=1e-11)
    ctx, st, k = gpu_eval(rs, precision=precision, solver=solver)
    ref = oracle.force_energy(rs.force, rs.positions, rs.box, target_eps=1e-13, max_iter=5000)
    assert abs(st.getPotentialEnergy() - ref["energy"]) <= E_TOL * abs(ref["energy"])
    assert rel_rms(st.getForces(), ref["forces"]) <= F_TOL
    assert np.max(np.abs(rs.force.getInducedDipoles(ctx) - ref["induced"])) <= MU_TOL


def test_water216_fp64_jacobi_matches_oracle():
    """216 waters, 648 sites, reference PME semantics, Mutual to 1e-6 D with the reference iteration."""
    _say("case water216 jacobi fp64")
    s = systems.config0()
    rs = ref_systems.RefSystem("water216", None, s.force, s.positions, s.box)
    from tholedipoleplugin_amd.context import System
    sysm = System()
    for m in s.masses:
        sysm.addParticle(m)
    sysm.setDefaultPeriodicBoxVectors(*s.box)
    sysm.addForce(s.force)
    rs.system = sysm
    ctx, st, k = gpu_eval(rs)
    ref = oracle.force_energy(s.force, s.positions, s.box, threads=8)
    stats = k.solverStats()
    assert stats["iterations"] == ref["iterations"]
    assert abs(st.getPotentialEnergy() - ref["energy"]) <= E_TOL * abs(ref["energy"])
    assert rel_rms(st.getForces(), ref["forces"]) <= F_TOL
    assert np.max(np.abs(s.force.getInducedDipoles(ctx) - ref["induced"])) <= MU_TOL
    # mixed precision + PCG against the same oracle
    ctx2, st2, k2 = gpu_eval(rs, precision="mixed", solver="pcg")
    assert rel_rms(st2.getForces(), ref["forces"]) <= F_TOL
    assert abs(st2.getPotentialEnergy() - ref["energy"]) <= 1e-5 * abs(ref["energy"])


def test_potential_and_moments():
    _say("case potential/moments")
    for name in ("ammonia", "water4"):
        rs = tighten(ref_systems.ALL[name](T.Mutual))
        ctx, st, k = gpu_eval(rs)
        rng = np.random.default_rng(5)
        grid = rng.uniform(-0.8, 0.8, (100, 3))
        pot = rs.force.getElectrostaticPotential(grid, ctx)
        ref = oracle.potential(rs.force, rs.positions, grid, rs.box)
        assert np.max(np.abs(pot - ref)) <= 1e-9 * max(1.0, np.max(np.abs(ref)))
        mom = rs.force.getSystemMultipoleMoments(ctx)
        masses = [rs.system.getParticleMass(i) for i in range(rs.system.getNumParticles())]
        refm = oracle.multipole_moments(rs.force, rs.positions, masses, rs.box)
        assert np.max(np.abs(mom - refm)) <= 1e-9 * max(1.0, np.max(np.abs(refm)))


def test_update_parameters_in_context():
    """testTholeDipoleAmmoniaMutualPolarization live assertions (TestReferenceTholeDipoleForce.cpp:285-313)."""
    _say("case updateParametersInContext")
    rs = ref_systems.ammonia(T.Mutual)
    ctx, st1, _ = gpu_eval(rs)
    f = rs.force
    for i in range(f.getNumParticles()):
        q, d, a, th, ax, z, x, y = f.getParticleParameters(i)
        d[0] *= 0.7
        f.setParticleParameters(i, 1.1 * q, d, 1.5 * a, 1.3 * th, ax, z, x, y)
    ctx2 = Context(rs.system)
    ctx2.setPositions(rs.positions)
    s1 = ctx.getState(getForces=True, getEnergy=True)
    s2 = ctx2.getState(getForces=True, getEnergy=True)
    assert abs(s1.getPotentialEnergy() - s2.getPotentialEnergy()) > 1e-4 * abs(s2.getPotentialEnergy())
    f.updateParametersInContext(ctx)
    s1 = ctx.getState(getForces=True, getEnergy=True)
    assert abs(s1.getPotentialEnergy() - s2.getPotentialEnergy()) <= 1e-10 * abs(s2.getPotentialEnergy())
    assert rel_rms(s1.getForces(), s2.getForces()) <= 1e-10


def test_plasma_energy_independent_of_cutoff():
    """testNeutralizingPlasmaCorrection (TestReferenceTholeDipoleForce.cpp:761-801)."""
    _say("case plasma")
    rs = ref_systems.plasma()
    ctx, st, _ = gpu_eval(rs)
    e1 = st.getPotentialEnergy()
    rs.force.setCutoffDistance(0.7)
    ctx.reinitialize(True)
    e2 = ctx.getState(getEnergy=True).getPotentialEnergy()
    assert abs(e1 - e2) <= 1e-4 * abs(e1)
    rs.force.setParticleParameters(0, 2.0, [0, 0, 0], 0.001, 0.39, T.NoAxisType, 0, 0, 0)
    rs.force.updateParametersInContext(ctx)
    e3 = ctx.getState(getEnergy=True).getPotentialEnergy()
    rs.force.setCutoffDistance(1.0)
    ctx.reinitialize(True)
    e4 = ctx.getState(getEnergy=True).getPotentialEnergy()
    assert abs(e3 - e4) <= 1e-4 * abs(e3)
    assert abs(e3 - 2.0 * e1) <= 1e-9 * abs(e3)  # q0 doubled: charge-charge energy doubles


def test_error_paths():
    _say("case errors")
    rs = ref_systems.water4()
    ctx, _, k = gpu_eval(rs)
    # box shrunk below twice the cutoff (ReferenceTholeDipoleKernels.cpp:170-173)
    ctx.setPeriodicBoxVectors([1.3, 0, 0], [0, 1.3, 0], [0, 0, 1.3])
    with pytest.raises(OpenMMException, match="less than twice the nonbonded cutoff"):
        ctx.getState(getEnergy=True)
    # particle count change (ReferenceTholeDipoleKernels.cpp:318-319)
    rs.force.addParticle(0.0, [0, 0, 0], 0.001, 0.39)
    with pytest.raises(OpenMMException, match="number of particles has changed"):
        k.copyParametersToContext(ctx, rs.force)
    # PME parameters on a NoCutoff context (ReferenceTholeDipoleKernels.cpp:341-343)
    am = ref_systems.ammonia()
    actx, _, ak = gpu_eval(am)
    with pytest.raises(OpenMMException, match="not using PME"):
        am.force.getPMEParametersInContext(actx)
    # explicit PME parameters are reported as given; alpha = 0 selects from the tolerance
    w = ref_systems.water4()
    wctx, _, _ = gpu_eval(w)
    a, nx, ny, nz = w.force.getPMEParametersInContext(wctx)
    assert (a, nx, ny, nz) == (5.4459052, 20, 20, 20)
    w2 = ref_systems.water4()
    w2.force.setPMEParameters(0.0, 0, 0, 0)
    w2ctx, _, _ = gpu_eval(w2)
    a, nx, ny, nz = w2.force.getPMEParametersInContext(w2ctx)
    tol, rc, L = 1e-4, 0.7, 1.8643
    assert abs(a - np.sqrt(-np.log(2 * tol)) / rc) < 1e-12

    def legal(n):   # smallest size >= n with prime factors 2, 3, 5, 7 (OpenMM's FFT sizes)
        while True:
            m = n
            for p in (2, 3, 5, 7):
                while m % p == 0:
                    m //= p
            if m == 1:
                return n
            n += 1
    n0 = int(np.ceil(2 * a * L / (3 * tol ** 0.2)))
    assert nx == legal(max(6, n0)) and (n0, nx) == (33, 35)


def test_device_entry_point_matches_host():
    """tdp_execute_device with torch-allocated device buffers on a torch stream equals tdp_execute.
    Runs in a fresh process in which torch initialises the HIP runtime first."""
    import os
    import subprocess
    import sys
    _say("case device entry")
    script = os.path.join(os.path.dirname(os.path.abspath(__file__)), "device_entry_check.py")
    r = subprocess.run([sys.executable, script], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr[-2000:])
    assert r.returncode == 0 and "OK" in r.stdout


def _random_system(n, periodic, seed):
    """n polarizable charges (no exclusions) spread over a 1.6 nm box, at least 0.12 nm apart."""
    from tholedipoleplugin_amd.context import System
    rng = np.random.default_rng(seed)
    L = 1.6
    pos = []
    while len(pos) < n:
        x = rng.uniform(0.0, L, 3)
        if all(np.linalg.norm((x - y) - L * np.round((x - y) / L)) > 0.12 for y in pos):
            pos.append(x)
    system = System()
    f = TholeDipoleForce()
    f.setNonbondedMethod(T.PME if periodic else T.NoCutoff)
    f.setCutoffDistance(0.7)
    f.setPolarizationType(T.Mutual)
    for i in range(n):
        system.addParticle(1.0)
        f.addParticle(float(rng.uniform(-0.8, 0.8)), [0.0, 0.0, 0.0], float(rng.uniform(0.3e-3, 1.5e-3)),
                      float(rng.uniform(0.3, 0.45)), T.NoAxisType, -1, -1, -1)
    box = np.diag([L, L, L]) if periodic else None
    if periodic:
        system.setDefaultPeriodicBoxVectors(*box)
    system.addForce(f)
    return ref_systems.RefSystem("random%d" % n, system, f, np.array(pos).reshape(-1, 3), box)


@pytest.mark.parametrize("periodic", [False, True])
@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 129])
def test_ragged_sizes(n, periodic):
    """Systems of 1 and 2 particles and around the 64-particle tile boundary (a partial last tile,
    exactly full tiles, one particle past), against the oracle at the bar, both precisions."""
    _say("case ragged n=%d periodic=%d" % (n, periodic))
    rs = tighten(_random_system(n, periodic, 40 + n), eps=1e-12)
    ref = oracle.force_energy(rs.force, rs.positions, rs.box, target_eps=1e-13, max_iter=5000)
    for precision in ("double", "mixed"):
        ctx, st, k = gpu_eval(rs, precision=precision, solver="pcg")
        e = st.getPotentialEnergy()
        assert abs(e - ref["energy"]) <= E_TOL * abs(ref["energy"]) + 1e-12
        if n > 1:
            assert rel_rms(st.getForces(), ref["forces"]) <= F_TOL
        assert np.max(np.abs(rs.force.getInducedDipoles(ctx) - ref["induced"])) <= MU_TOL



def _grouped_system(ngroup, periodic, seed):
    """Groups of a center and five satellites 0.1-0.2 nm from it, the centers at least 0.45 nm apart
    in a 2.2 nm box.  The center's 1-2 list holds its satellites, so by list position the pairs take
    mScale 0, 0, 0, 0.5 and 1; each satellite's 1-2 list holds the center and its 1-3 list the other
    satellites, so the satellite pairs take iScale 0 (first in the list) or 1."""
    from tholedipoleplugin_amd.context import System
    rng = np.random.default_rng(seed)
    L = 2.2
    centers = []
    while len(centers) < ngroup:
        x = rng.uniform(0.0, L, 3)
        if all(np.linalg.norm((x - y) - L * np.round((x - y) / L)) > 0.45 for y in centers):
            centers.append(x)
    pos = []
    for c in centers:
        sats = []
        while len(sats) < 5:
            u = rng.normal(size=3)
            x = c + u / np.linalg.norm(u) * rng.uniform(0.1, 0.2)
            if all(np.linalg.norm(x - y) > 0.1 for y in sats):
                sats.append(x)
        pos += [c] + sats
    system = System()
    f = TholeDipoleForce()
    f.setNonbondedMethod(T.PME if periodic else T.NoCutoff)
    f.setCutoffDistance(0.7)
    f.setPolarizationType(T.Mutual)
    for i in range(6 * ngroup):
        system.addParticle(1.0)
        f.addParticle(float(rng.uniform(-0.8, 0.8)), [0.0, 0.0, 0.0], float(rng.uniform(0.3e-3, 1.0e-3)),
                      float(rng.uniform(0.3, 0.45)), T.NoAxisType, -1, -1, -1)
    for g in range(ngroup):
        c = 6 * g
        sat = [c + 1 + k for k in range(5)]
        f.setCovalentMap(c, T.Covalent12, sat)
        for k, sk in enumerate(sat):
            f.setCovalentMap(sk, T.Covalent12, [c])
            f.setCovalentMap(sk, T.Covalent13, [x for x in sat[k + 1:] + sat[:k]])
    box = np.diag([L, L, L]) if periodic else None
    if periodic:
        system.setDefaultPeriodicBoxVectors(*box)
    system.addForce(f)
    return ref_systems.RefSystem("grouped%d" % ngroup, system, f, np.array(pos).reshape(-1, 3), box)


@pytest.mark.parametrize("periodic", [False, True])
def test_scaled_pairs_both_precisions(periodic):
    """Excluded and half-scaled pairs (mScale 0 and 0.5, iScale 0) inside near tile pairs: the
    mixed paths defer every step with such a pair to FP64, which applies the scales; both
    precisions against the oracle at the bar."""
    _say("case scaled pairs periodic=%d" % periodic)
    rs = tighten(_grouped_system(50, periodic, 61 + periodic), eps=1e-12)
    assert tuple(oracle.pair_scales(rs.force, 0, 4)) == (0.5, 1.0)   # center, fourth satellite
    assert tuple(oracle.pair_scales(rs.force, 1, 2)) == (1.0, 0.0)   # satellites, first in the 1-3 list
    ref = oracle.force_energy(rs.force, rs.positions, rs.box, target_eps=1e-13, max_iter=5000)
    for precision in ("double", "mixed"):
        ctx, st, k = gpu_eval(rs, precision=precision, solver="pcg")
        e = st.getPotentialEnergy()
        assert abs(e - ref["energy"]) <= E_TOL * abs(ref["energy"]), (precision, e, ref["energy"])
        assert rel_rms(st.getForces(), ref["forces"]) <= F_TOL, precision
        assert np.max(np.abs(rs.force.getInducedDipoles(ctx) - ref["induced"])) <= MU_TOL, precision

@pytest.mark.parametrize("solver", ["pcg", "jacobi"])
def test_warm_start_mutual(solver):
    """A warm-started Mutual solve (previous evaluation's dipoles as the initial guess) reaches the
    cold-start dipoles within the bar in fewer iterations, over a short sequence of perturbed
    configurations (the next configuration of an MD loop)."""
    from tholedipoleplugin_amd import _lib
    _say("case warm start %s" % solver)
    s = systems.config0()
    s.force.setMutualInducedTargetEpsilon(1e-9)
    s.force.setMutualInducedMaxIterations(500)
    rng = np.random.default_rng(8)
    p = np.asarray(s.positions, np.float64)
    seq = [p]
    for _ in range(3):
        seq.append(seq[-1] + rng.normal(0.0, 0.0005, p.shape))
    hw = _lib.HipContextHandle(s.force, s.box, precision="double", solver=solver)
    hw.set_warm_start(True)
    hc = _lib.HipContextHandle(s.force, s.box, precision="double", solver=solver)
    for k, pos in enumerate(seq):
        ew, fw = hw.execute(pos, s.box)
        ec, fc = hc.execute(pos, s.box)
        nw, nc = hw.solver_stats()["iterations"], hc.solver_stats()["iterations"]
        muw = hw.dipoles(pos, s.box, _lib.DIPOLES_INDUCED)
        muc = hc.dipoles(pos, s.box, _lib.DIPOLES_INDUCED)
        assert hw.solver_stats()["converged"] == 1
        assert np.max(np.abs(muw - muc)) <= MU_TOL
        assert abs(ew - ec) <= E_TOL * abs(ec)
        assert rel_rms(fw, fc) <= F_TOL
        if k > 0:
            assert nw < nc, (k, nw, nc)
    hw.close()
    hc.close()


def test_autotune_keeps_results():
    """tdp_autotune picks one of its candidate work-item lengths and changes results only at the
    rounding level; repeated evaluations after it are identical."""
    from tholedipoleplugin_amd import _lib
    _say("case autotune")
    s = systems.config0()
    s.force.setMutualInducedTargetEpsilon(1e-10)
    p0 = np.asarray(s.positions, np.float64)
    h0 = _lib.HipContextHandle(s.force, s.box, precision="double", solver="pcg")
    e0, f0 = h0.execute(p0, s.box)
    h0.close()
    h = _lib.HipContextHandle(s.force, s.box, precision="double", solver="pcg")
    C = h.autotune(p0, s.box)
    assert C in (1, 2, 3, 4, 6, 8, 12, 16)
    e1, f1 = h.execute(p0, s.box)
    e2, f2 = h.execute(p0, s.box)
    h.close()
    assert e1 == e2 and np.array_equal(f1, f2)
    assert abs(e1 - e0) <= 1e-12 * abs(e0)
    assert rel_rms(f1, f0) <= 1e-10


@pytest.mark.parametrize("solver", ["pcg", "jacobi"])
def test_iteration_graph_replays(solver, monkeypatch):
    """Solver iterations replayed from captured HIP graphs give the same results as iterations
    enqueued launch by launch (TDP_ITERATION_GRAPHS=0), over a sequence of evaluations that reuses
    the graphs with new positions, a box change (new capture) and a repeat of the first
    configuration."""
    from tholedipoleplugin_amd import _lib
    _say("case iteration graphs %s" % solver)
    s = systems.config0()
    s.force.setMutualInducedTargetEpsilon(1e-9)
    s.force.setMutualInducedMaxIterations(500)
    rng = np.random.default_rng(5)
    p0 = np.asarray(s.positions, np.float64)
    p1 = p0 + rng.uniform(-0.002, 0.002, p0.shape)
    box0 = np.asarray(s.box, np.float64)
    box1 = box0 * 1.001
    seq = [(p0, box0), (p0, box0), (p1, box0), (p1, box1), (p0, box0)]
    out = {}
    for graphs in ("1", "0"):
        monkeypatch.setenv("TDP_ITERATION_GRAPHS", graphs)
        h = _lib.HipContextHandle(s.force, box0, precision="mixed", solver=solver)
        res = []
        for pos, box in seq:
            e, f = h.execute(pos, box)
            res.append((e, f.copy(), h.solver_stats()["iterations"]))
        h.close()
        out[graphs] = res
    for (e1, f1, n1), (e0, f0, n0) in zip(out["1"], out["0"]):
        assert n1 == n0
        assert e1 == e0
        assert np.array_equal(f1, f0)


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_water12288_every_site_full_size(precision):
    """36,864 sites (12,288 waters, 4.95 nm box, reference PME semantics), PCG solved to an RMS
    dipole update of 1e-9 e*nm (well below the 1e-6 D bar, so that solver truncation does not
    hide the operator's error).  The oracle evaluates the reference's three pair passes in FP64
    over all 6.8e8 pairs with the GPU's induced dipoles (tdpo_evaluate_at, on the host's threads);
    at EVERY site the fixed point mu_i = alpha_i (E0_i + (T mu)_i) holds to 1e-6 D, the forces match
    to the force tolerance, and the total energy matches to the energy tolerance."""
    _say("case water12288 every site %s" % precision)
    s = systems.config1()
    s.force.setMutualInducedTargetEpsilon(1e-9)
    from tholedipoleplugin_amd.context import System
    sysm = System()
    for m in s.masses:
        sysm.addParticle(m)
    sysm.setDefaultPeriodicBoxVectors(*s.box)
    sysm.addForce(s.force)
    ctx = Context(sysm, precision=precision, solver="pcg")
    ctx.setPeriodicBoxVectors(*s.box)
    ctx.setPositions(s.positions)
    st = ctx.getState(getForces=True, getEnergy=True)
    k = ctx._kernel_for(s.force)
    stats = k.solverStats()
    _say("  stats %s" % stats)
    assert stats["converged"] == 1
    mu = s.force.getInducedDipoles(ctx)
    threads = max(1, min(16, os.cpu_count() or 1))
    _say("  oracle passes over all pairs on %d threads" % threads)
    ref = oracle.evaluate_at(s.force, s.positions, mu, s.box, threads=threads)
    alpha = s.force.packed_arrays()["polarity"][:, None]
    resid = mu - alpha * (ref["e0"] + ref["eind"])
    de = abs(st.getPotentialEnergy() - ref["energy"]) / abs(ref["energy"])
    df = rel_rms(st.getForces(), ref["forces"])
    _say("  every site: max |mu - alpha (E0 + T mu)| %.3e e*nm, force rel RMS %.3e, energy rel %.3e"
         % (np.max(np.abs(resid)), df, de))
    assert np.max(np.abs(resid)) <= MU_TOL
    assert df <= F_TOL
    assert de <= E_TOL


def test_saline96000_mixed_matches_double_and_rows():
    """96,000 sites (31,978 waters + 33 Na+/Cl- pairs, 7.1 nm box, Thole a ~ U[0.3, 0.45]):
    mixed and FP64 PCG solves (1e-10 e*nm) agree at every site to the bar, and the FP64 solution
    satisfies mu = alpha (E0 + T mu) on sampled rows evaluated by the oracle over all partners."""
    _say("case saline96000")
    s = systems.config2()
    s.force.setMutualInducedTargetEpsilon(1e-10)
    from tholedipoleplugin_amd.context import System
    out = {}
    for precision in ("double", "mixed"):
        sysm = System()
        for m in s.masses:
            sysm.addParticle(m)
        sysm.setDefaultPeriodicBoxVectors(*s.box)
        sysm.addForce(s.force)
        ctx = Context(sysm, precision=precision, solver="pcg")
        ctx.setPeriodicBoxVectors(*s.box)
        ctx.setPositions(s.positions)
        st = ctx.getState(getForces=True, getEnergy=True)
        stats = ctx._kernel_for(s.force).solverStats()
        _say("  %s stats %s" % (precision, stats))
        assert stats["converged"] == 1
        out[precision] = (st.getPotentialEnergy(), st.getForces(), s.force.getInducedDipoles(ctx))
    (e64, f64, mu64), (em, fm, mum) = out["double"], out["mixed"]
    _say("  max |mu_mixed - mu_double| = %.3e D" % (np.max(np.abs(mum - mu64)) / DEBYE))
    assert np.max(np.abs(mum - mu64)) <= MU_TOL
    assert rel_rms(fm, f64) <= F_TOL
    assert abs(em - e64) <= E_TOL * abs(e64)
    rows = np.random.default_rng(8).choice(s.n, 24, replace=False)
    ref = oracle.rows(s.force, s.positions, mu64, rows, s.box, threads=8)
    alpha = s.force.packed_arrays()["polarity"][rows][:, None]
    resid = mu64[rows] - alpha * (ref["e0"] + ref["eind"])
    _say("  max |mu - alpha (E0 + T mu)| on sampled rows: %.3e e*nm" % np.max(np.abs(resid)))
    assert np.max(np.abs(resid)) <= MU_TOL
    assert rel_rms(f64[rows], ref["force"]) <= F_TOL


def half_box_system():
    """128 polarizable charges in a 4 nm periodic box: cluster A (64 sites on a 4x4x4 grid, 0.2 nm
    spacing, in the box octant nearest the origin, so that A and B fill one 64-site tile each
    and the pair of tiles lies beyond the Thole range) and cluster B, each site of B displaced
    from its partner in A by
    (L/2 - delta_k, 0.3, 0.2) nm with delta_k from +-1e-7 down to +-1e-11 nm and 0.  Those pairs
    sit closer to half a box than FP32 coordinates can resolve, and the 0.3/0.2 nm transverse
    offset makes the two images differ; the minimum image of each must follow the FP64 rule
    (d - L floor(d/L + 0.5), ReferenceTholeDipoleForce.cpp:1414-1423)."""
    from tholedipoleplugin_amd.context import System
    L = 4.0
    g = np.arange(4) * 0.2
    a = np.array([[x, y, z] for x in g for y in g for z in g]) + 0.3
    deltas = np.array([1e-7, -1e-7, 1e-8, -1e-8, 1e-9, -1e-9, 1e-10, -1e-10, 1e-11, -1e-11, 0.0])
    d = np.resize(deltas, 64)
    b = a + np.stack([L / 2 - d, np.full(64, 0.3), np.full(64, 0.2)], axis=1)
    pos = np.concatenate([a, b])
    rng = np.random.default_rng(11)
    q = rng.choice([-0.5, 0.5], 128)
    system = System()
    f = TholeDipoleForce()
    f.setNonbondedMethod(T.PME)
    f.setPolarizationType(T.Mutual)
    for i in range(128):
        system.addParticle(1.0)
        f.addParticle(float(q[i]), [0.0, 0.0, 0.0], 1.0e-3, 0.39, T.NoAxisType, -1, -1, -1)
    box = np.diag([L, L, L])
    system.setDefaultPeriodicBoxVectors(*box)
    system.addForce(f)
    return ref_systems.RefSystem("halfbox", system, f, pos, box)


def collinear_grid_system(periodic):
    """192 polarizable charges in three 64-site clusters on one exact grid (4x4x4, 0.2 nm spacing):
    A, C = A + (0.9, 0, 0) and B = A + (1.8, 0, 0) nm.  Every site of C and B has partners in the
    other clusters with bit-identical raw y and z, so those pairs are separated exactly along x: the
    reference then builds the QI frame from e_y instead of e_x (ReferenceTholeDipoleForce.cpp:584-588),
    and the closed form of the FP32 energy pass has no frame (rho = 0).  A and B are one tile each
    and beyond the Thole range of each other (far tile pair); C is within it of both (near tile
    pairs)."""
    from tholedipoleplugin_amd.context import System
    L = 4.0
    g = np.arange(4) * 0.2
    a = np.array([[x, y, z] for x in g for y in g for z in g]) + 0.3
    pos = np.concatenate([a, a + np.array([0.9, 0.0, 0.0]), a + np.array([1.8, 0.0, 0.0])])
    rng = np.random.default_rng(12)
    q = rng.choice([-0.5, 0.5], len(pos))
    system = System()
    f = TholeDipoleForce()
    f.setNonbondedMethod(T.PME if periodic else T.NoCutoff)
    f.setPolarizationType(T.Mutual)
    for i in range(len(pos)):
        system.addParticle(1.0)
        f.addParticle(float(q[i]), [0.0, 0.0, 0.0], 1.0e-3, 0.39, T.NoAxisType, -1, -1, -1)
    box = np.diag([L, L, L]) if periodic else None
    if periodic:
        system.setDefaultPeriodicBoxVectors(*box)
    system.addForce(f)
    return ref_systems.RefSystem("collinear", system, f, pos, box)


@pytest.mark.parametrize("precision", ["double", "mixed"])
@pytest.mark.parametrize("periodic", [False, True])
def test_collinear_pairs_equal_yz(periodic, precision):
    """Pairs with identical raw y and z (the reference's e_y frame) and separations exactly along x,
    on far and near tile pairs, against the oracle at the bar: the FP32 closed form of the mixed
    energy pass must hand these pairs to the explicit frame."""
    _say("case collinear periodic=%d %s" % (periodic, precision))
    rs = tighten(collinear_grid_system(periodic), eps=1e-12)
    ctx, st, k = gpu_eval(rs, precision=precision, solver="pcg")
    ref = oracle.force_energy(rs.force, rs.positions, rs.box, target_eps=1e-13, max_iter=5000)
    assert k.solverStats()["converged"] == 1
    assert abs(st.getPotentialEnergy() - ref["energy"]) <= E_TOL * abs(ref["energy"])
    assert rel_rms(st.getForces(), ref["forces"]) <= F_TOL
    assert np.max(np.abs(rs.force.getInducedDipoles(ctx) - ref["induced"])) <= MU_TOL


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_half_box_images(precision):
    _say("case half box %s" % precision)
    rs = tighten(half_box_system(), eps=1e-12)
    ctx, st, k = gpu_eval(rs, precision=precision, solver="pcg")
    ref = oracle.force_energy(rs.force, rs.positions, rs.box, target_eps=1e-13, max_iter=5000)
    assert k.solverStats()["converged"] == 1
    assert abs(st.getPotentialEnergy() - ref["energy"]) <= E_TOL * abs(ref["energy"])
    assert rel_rms(st.getForces(), ref["forces"]) <= F_TOL
    assert np.max(np.abs(rs.force.getInducedDipoles(ctx) - ref["induced"])) <= MU_TOL


def dipolar_water(kind):
    """Waters with molecular dipoles and axis frames, large enough for tile pairs beyond the Thole
    range: a periodic box of 512 waters at liquid density (2.48 nm) or a 200-molecule cluster.  O:
    Bisector (Z = H1, X = H2), H: ZThenX (Z = O, X = the other H); dipoles drawn per site type."""
    from tholedipoleplugin_amd.context import System
    if kind == "box":
        s = systems.water_box(8, 512, 2.48, 31, method=T.PME)
    else:
        s = systems.water_cluster(200, seed=32)
    f = s.force
    rng = np.random.default_rng(33)
    dO = rng.uniform(-0.01, 0.01, 3)
    dH = rng.uniform(-0.005, 0.005, 3)
    for m in range(f.getNumParticles() // 3):
        o, h1, h2 = 3 * m, 3 * m + 1, 3 * m + 2
        for i, d, ax, z, x in ((o, dO, T.Bisector, h1, h2), (h1, dH, T.ZThenX, o, h2), (h2, dH, T.ZThenX, o, h1)):
            q, _, a, th = f.getParticleParameters(i)[:4]
            f.setParticleParameters(i, q, list(d), a, th, ax, z, x, -1)
    system = System()
    for mass in s.masses:
        system.addParticle(mass)
    if s.box is not None:
        system.setDefaultPeriodicBoxVectors(*s.box)
    system.addForce(f)
    return ref_systems.RefSystem("dipolar_" + kind, system, f, np.asarray(s.positions, np.float64), s.box)


@pytest.mark.parametrize("precision", ["double", "mixed"])
@pytest.mark.parametrize("kind", ["box", "cluster"])
def test_permanent_dipoles_far_tiles(kind, precision):
    """Permanent dipoles, axis frames and torques on systems with far tile pairs (the FP32 far path
    of the mixed energy pass and T*mu, FP64 otherwise) against the oracle at the parity bar."""
    _say("case dipolar %s %s" % (kind, precision))
    rs = tighten(dipolar_water(kind), eps=1e-12)
    ctx, st, k = gpu_eval(rs, precision=precision, solver="pcg")
    ref = oracle.force_energy(rs.force, rs.positions, rs.box, target_eps=1e-13, max_iter=5000)
    assert k.solverStats()["converged"] == 1
    assert abs(st.getPotentialEnergy() - ref["energy"]) <= E_TOL * abs(ref["energy"])
    assert rel_rms(st.getForces(), ref["forces"]) <= F_TOL
    assert np.max(np.abs(rs.force.getInducedDipoles(ctx) - ref["induced"])) <= MU_TOL


def test_water12288_mixed_matches_double():
    """36,864 sites: the mixed-precision solve agrees with the FP64 one at every site (induced
    dipoles within 1e-6 D, forces within the force tolerance, energy within 1e-6), both solved
    to an RMS dipole update of 1e-10 e*nm so that the difference is the operator's."""
    _say("case water12288 mixed vs double")
    s = systems.config1()
    s.force.setMutualInducedTargetEpsilon(1e-10)
    from tholedipoleplugin_amd.context import System
    out = {}
    for precision in ("double", "mixed"):
        sysm = System()
        for m in s.masses:
            sysm.addParticle(m)
        sysm.setDefaultPeriodicBoxVectors(*s.box)
        sysm.addForce(s.force)
        ctx = Context(sysm, precision=precision, solver="pcg")
        ctx.setPeriodicBoxVectors(*s.box)
        ctx.setPositions(s.positions)
        st = ctx.getState(getForces=True, getEnergy=True)
        assert ctx._kernel_for(s.force).solverStats()["converged"] == 1
        out[precision] = (st.getPotentialEnergy(), st.getForces(), s.force.getInducedDipoles(ctx))
    (e64, f64, mu64), (em, fm, mum) = out["double"], out["mixed"]
    dmu = np.max(np.abs(mum - mu64))
    _say("  max |mu_mixed - mu_double| = %.3e D, force rel RMS %.3e, energy rel %.3e"
         % (dmu / DEBYE, rel_rms(fm, f64), abs(em - e64) / abs(e64)))
    assert dmu <= MU_TOL
    assert rel_rms(fm, f64) <= F_TOL
    assert abs(em - e64) <= E_TOL * abs(e64)


def test_cpp_host_kernel():
    """The C++ host-side kernel mirror (tholedipoleplugin_amd/host) run through its own driver."""
    import os
    import subprocess
    _say("case C++ host kernel")
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "test_host_kernel")
    assert os.path.exists(exe), "build() has not built tests/cpp/test_host_kernel"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0 and "Done" in r.stdout
