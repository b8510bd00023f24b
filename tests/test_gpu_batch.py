"""Batched evaluation of independent NoCutoff systems (tdp_batch_* entry points) through the C ABI.

Each system of a batch must get exactly what its own single-system context gives (same pair terms,
scale maps, solver and stopping rule): energies, forces and induced dipoles are compared with
single contexts (FP64, to 1e-10 relative: only the summation order of the tile partials differs)
and with the CPU oracle at the parity bar.  Tolerances of the bar: energy relative 1e-6, force
relative RMS 1e-5, induced dipoles 1e-6 D.
"""
import numpy as np
import pytest

import oracle
import ref_systems
from tholedipoleplugin_amd import _lib, systems
from tholedipoleplugin_amd.force import TholeDipoleForce

pytestmark = pytest.mark.gpu

T = TholeDipoleForce
DEBYE = 0.020819434
MU_TOL = 1e-6 * DEBYE


def _say(msg):
    print(msg, flush=True)


def rel_rms(a, b):
    a = np.asarray(a).reshape(-1)
    b = np.asarray(b).reshape(-1)
    return float(np.sqrt(np.mean((a - b) ** 2)) / max(np.sqrt(np.mean(b ** 2)), 1e-300))


def batch_members(eps=1e-11, polarization=T.Mutual):
    """Reference test systems with axis frames and exclusions (ammonia, ZOnly) and water clusters
    of 7, 20, 45 and 100 molecules (one to five tiles), all NoCutoff."""
    members = []
    for rs in (ref_systems.ammonia(polarization), ref_systems.zonly(polarization)):
        members.append((rs.force, np.asarray(rs.positions, np.float64)))
    for k, m in enumerate((7, 20, 45, 100)):
        c = systems.water_cluster(m, seed=500 + k)
        c.force.setPolarizationType(polarization)
        members.append((c.force, np.asarray(c.positions, np.float64)))
    for f, _ in members:
        f.setMutualInducedTargetEpsilon(eps)
        f.setMutualInducedMaxIterations(1000)
    return members


@pytest.mark.parametrize("solver", ["pcg", "jacobi"])
def test_batch_matches_single_contexts(solver):
    _say("case batch vs single contexts, %s" % solver)
    members = batch_members()
    pos = np.concatenate([p for _, p in members])
    b = _lib.HipBatchHandle([f for f, _ in members], precision="double", solver=solver)
    e_b, f_b = b.execute(pos)
    mu_b = b.dipoles(pos, None, _lib.DIPOLES_INDUCED)
    off = 0
    for g, (f, p) in enumerate(members):
        n = f.getNumParticles()
        h = _lib.HipContextHandle(f, np.zeros(9), precision="double", solver=solver)
        e1, f1 = h.execute(p, None)
        mu1 = h.dipoles(p, None, _lib.DIPOLES_INDUCED)
        st1 = h.solver_stats()
        h.close()
        stb = b.system_stats(g)
        assert stb["converged"] == 1 and st1["converged"] == 1
        assert abs(stb["iterations"] - st1["iterations"]) <= 1
        assert abs(e_b[g] - e1) <= 1e-10 * abs(e1)
        assert rel_rms(f_b[off:off + n], f1) <= 1e-10
        assert np.max(np.abs(mu_b[off:off + n] - mu1)) <= 1e-12
        off += n
    b.close()


def test_batch_matches_oracle():
    _say("case batch vs oracle")
    members = batch_members()
    pos = np.concatenate([p for _, p in members])
    b = _lib.HipBatchHandle([f for f, _ in members], precision="double", solver="pcg")
    e_b, f_b = b.execute(pos)
    mu_b = b.dipoles(pos, None, _lib.DIPOLES_INDUCED)
    b.close()
    off = 0
    for g, (f, p) in enumerate(members):
        n = f.getNumParticles()
        ref = oracle.force_energy(f, p, None, target_eps=1e-13, max_iter=5000)
        assert abs(e_b[g] - ref["energy"]) <= 1e-6 * abs(ref["energy"])
        assert rel_rms(f_b[off:off + n], ref["forces"]) <= 1e-5
        assert np.max(np.abs(mu_b[off:off + n] - ref["induced"])) <= MU_TOL
        off += n


def test_batch_reference_stopping_rule():
    """With each system's own target and the Jacobi solver, every system stops after the iteration
    count and at the epsilon of convergeInducedDipolesByDIIS (ReferenceTholeDipoleForce.cpp:829-850)."""
    _say("case batch stopping rule")
    members = batch_members()
    targets = [1e-5, 1e-6, 1e-7, 3e-6, 1e-8, 2e-7]
    for (f, _), t in zip(members, targets):
        f.setMutualInducedTargetEpsilon(t)
    pos = np.concatenate([p for _, p in members])
    b = _lib.HipBatchHandle([f for f, _ in members], precision="double", solver="jacobi")
    b.execute(pos)
    for g, (f, p) in enumerate(members):
        ref = oracle.force_energy(f, p, None)
        st = b.system_stats(g)
        assert st["iterations"] == ref["iterations"], g
        assert abs(st["epsilon"] - ref["epsilon"]) <= 1e-6 * ref["epsilon"], g
    b.close()


@pytest.mark.parametrize("pol", [T.Direct, T.Extrapolated])
def test_batch_other_polarization_types(pol):
    _say("case batch polarization %d" % pol)
    members = batch_members(polarization=pol)
    pos = np.concatenate([p for _, p in members])
    b = _lib.HipBatchHandle([f for f, _ in members], precision="double", solver="pcg")
    e_b, f_b = b.execute(pos)
    b.close()
    off = 0
    for g, (f, p) in enumerate(members):
        n = f.getNumParticles()
        ref = oracle.force_energy(f, p, None)
        assert abs(e_b[g] - ref["energy"]) <= 1e-6 * abs(ref["energy"])
        assert rel_rms(f_b[off:off + n], ref["forces"]) <= 1e-5
        off += n


def test_batch_mixed_precision():
    _say("case batch mixed")
    members = batch_members(eps=1e-10)
    pos = np.concatenate([p for _, p in members])
    bd = _lib.HipBatchHandle([f for f, _ in members], precision="double", solver="pcg")
    bm = _lib.HipBatchHandle([f for f, _ in members], precision="mixed", solver="pcg")
    e_d, f_d = bd.execute(pos)
    e_m, f_m = bm.execute(pos)
    mu_d = bd.dipoles(pos, None, _lib.DIPOLES_INDUCED)
    mu_m = bm.dipoles(pos, None, _lib.DIPOLES_INDUCED)
    bd.close()
    bm.close()
    assert np.max(np.abs(mu_m - mu_d)) <= MU_TOL
    assert rel_rms(f_m, f_d) <= 1e-5
    assert np.all(np.abs(e_m - e_d) <= 1e-6 * np.abs(e_d))


def test_batch_of_config3_clusters_size_range():
    """Twelve clusters spanning the 20-400 molecule range of the cluster workload, solved to
    1e-7 D, against single FP64 contexts."""
    _say("case batch config3 sizes")
    sizes = [20, 400, 33, 250, 57, 140, 21, 380, 75, 99, 160, 310]
    members = []
    for k, m in enumerate(sizes):
        c = systems.water_cluster(m, seed=900 + k)
        c.force.setMutualInducedTargetEpsilon(1e-7 * DEBYE)
        c.force.setMutualInducedMaxIterations(1000)
        members.append((c.force, np.asarray(c.positions, np.float64)))
    pos = np.concatenate([p for _, p in members])
    b = _lib.HipBatchHandle([f for f, _ in members], precision="mixed", solver="pcg")
    e_b, f_b = b.execute(pos)
    mu_b = b.dipoles(pos, None, _lib.DIPOLES_INDUCED)
    b.close()
    off = 0
    for g, (f, p) in enumerate(members):
        n = f.getNumParticles()
        f.setMutualInducedTargetEpsilon(1e-11)
        h = _lib.HipContextHandle(f, np.zeros(9), precision="double", solver="pcg")
        e1, f1 = h.execute(p, None)
        mu1 = h.dipoles(p, None, _lib.DIPOLES_INDUCED)
        h.close()
        assert np.max(np.abs(mu_b[off:off + n] - mu1)) <= MU_TOL, g
        assert rel_rms(f_b[off:off + n], f1) <= 1e-5, g
        assert abs(e_b[g] - e1) <= 1e-6 * abs(e1), g
        off += n


def test_batch_size_bucket_schedules_keep_results():
    """Work-item lengths set per size bucket (tdp_set_work_schedule) and chosen per bucket by
    tdp_autotune change the results of a batch whose clusters fall into several buckets only at the
    rounding level; evaluations after the tuner are deterministic."""
    _say("case batch size-bucket schedules")
    sizes = [20, 45, 90, 170, 330, 400, 25, 130]   # 1 to 19 tiles: all five buckets
    members = []
    for k, m in enumerate(sizes):
        c = systems.water_cluster(m, seed=1300 + k)
        c.force.setMutualInducedTargetEpsilon(1e-11)
        c.force.setMutualInducedMaxIterations(1000)
        members.append((c.force, np.asarray(c.positions, np.float64)))
    pos = np.concatenate([p for _, p in members])
    b = _lib.HipBatchHandle([f for f, _ in members], precision="double", solver="pcg")
    per0, bound = b.work_schedule()
    assert bound[:4] == [2, 4, 8, 16] and bound[4] == 2**31 - 1
    assert len(set(per0)) == 1
    e0, f0 = b.execute(pos)
    b.set_work_schedule([1, 2, 3, 6, 16])
    assert b.work_schedule()[0] == [1, 2, 3, 6, 16]
    e1, f1 = b.execute(pos)
    assert np.all(np.abs(e1 - e0) <= 1e-12 * np.abs(e0))
    assert rel_rms(f1, f0) <= 1e-10
    C = b.autotune(pos)
    per, _ = b.work_schedule()
    assert all(v in (1, 2, 3, 4, 6, 8, 12, 16) for v in per) and C == per[4]
    e2, f2 = b.execute(pos)
    e3, f3 = b.execute(pos)
    assert np.array_equal(e2, e3) and np.array_equal(f2, f3)
    assert np.all(np.abs(e2 - e0) <= 1e-12 * np.abs(e0))
    assert rel_rms(f2, f0) <= 1e-10
    with pytest.raises(_lib.TdpError, match="tiles per work item"):
        b.set_work_schedule([0, 1, 1, 1, 1])
    b.close()


def test_batch_rejects_bad_input():
    members = batch_members()
    rs = ref_systems.water4()
    with pytest.raises(_lib.TdpError, match="NoCutoff"):
        _lib.HipBatchHandle([members[0][0], rs.force])
    f = members[2][0]
    f2 = systems.water_cluster(5, seed=1).force
    f2.setPolarizationType(T.Direct)
    with pytest.raises(_lib.TdpError, match="polarization"):
        _lib.HipBatchHandle([f, f2])


def _dense_members(sizes, seed0):
    members = []
    for k, m in enumerate(sizes):
        c = systems.water_cluster(m, seed=seed0 + k)
        members.append((c.force, np.asarray(c.positions, np.float64)))
    return members


@pytest.mark.parametrize("method", ["fp16", "bf16"])
def test_dense_responses_match_cg_and_mutual(method):
    """Dense 16-bit Cholesky + FP64 refinement: the E0 response equals the Mutual induced dipoles of
    the batch PCG, and all four responses equal FP64 CG solves of the same fields."""
    _say("case dense %s" % method)
    members = _dense_members([5, 17, 40, 120, 333], 700)
    for f, _ in members:
        f.setMutualInducedTargetEpsilon(1e-12)
        f.setMutualInducedMaxIterations(1000)
    pos = np.concatenate([p for _, p in members])
    b = _lib.HipBatchHandle([f for f, _ in members], precision="double", solver="pcg")
    X, st = b.responses(pos, method=method, tolerance=1e-10)
    _say("  dense stats %s" % st)
    assert st["factor_failed"] == 0 and st["max_relative_residual"] <= 1e-10
    Y, stc = b.responses(pos, method="cg", tolerance=1e-12, max_steps=1000)
    _say("  cg stats %s" % stc)
    b.execute(pos)
    mu = b.dipoles(pos, None, _lib.DIPOLES_INDUCED)
    b.close()
    assert np.max(np.abs(X[0] - mu)) <= MU_TOL
    for q in range(4):
        assert np.max(np.abs(X[q] - Y[q])) <= MU_TOL, q



@pytest.mark.parametrize("method", ["fp16", "bf16"])
def test_dense_zero_field_and_size_limit(method):
    """A system without charges has E0 = 0: its E0 response is exactly zero (the solve scales each
    field by a power of two from its largest entry, and a zero field keeps scale 1), and its unit
    field responses equal FP64 CG.  A system past 1,024 particles is refused with an error before
    any work."""
    _say("case dense zero field %s" % method)
    members = _dense_members([6, 25], 900)
    f0 = members[0][0]
    for k in range(f0.getNumParticles()):
        q, d, a, th, ax, z, x, y = f0.getParticleParameters(k)
        f0.setParticleParameters(k, 0.0, d, a, th, ax, z, x, y)
    pos = np.concatenate([p for _, p in members])
    b = _lib.HipBatchHandle([f for f, _ in members], precision="double", solver="pcg")
    X, st = b.responses(pos, method=method, tolerance=1e-10)
    Y, _ = b.responses(pos, method="cg", tolerance=1e-12, max_steps=1000)
    b.close()
    _say("  dense stats %s" % st)
    n0 = f0.getNumParticles()
    assert st["factor_failed"] == 0 and st["max_relative_residual"] <= 1e-10
    assert np.all(X[0][:n0] == 0.0)
    for q in range(4):
        assert np.max(np.abs(X[q] - Y[q])) <= MU_TOL, q
    big = _dense_members([342], 950)   # 1,026 particles
    bb = _lib.HipBatchHandle([big[0][0]], precision="double", solver="pcg")
    with pytest.raises(_lib.TdpError, match="more than 1024 particles"):
        bb.responses(big[0][1], method=method, tolerance=1e-10)
    bb.close()

def test_dense_responses_residual_by_oracle():
    """The dense responses satisfy (alpha^-1 - T) mu = b with T applied by the CPU oracle
    (rows of the induced field over all partners), for E0 and the three unit fields."""
    _say("case dense residual via oracle")
    members = _dense_members([9, 31], 800)
    pos = np.concatenate([p for _, p in members])
    b = _lib.HipBatchHandle([f for f, _ in members], precision="double", solver="pcg")
    X, st = b.responses(pos, method="fp16", tolerance=1e-10)
    b.close()
    off = 0
    for g, (f, p) in enumerate(members):
        n = f.getNumParticles()
        alpha = f.packed_arrays()["polarity"][:, None]
        rows = np.arange(n)
        fields = [None, np.array([1.0, 0, 0]), np.array([0, 1.0, 0]), np.array([0, 0, 1.0])]
        for q in range(4):
            x = X[q][off:off + n]
            ref = oracle.rows(f, p, x, rows, None)
            bq = ref["e0"] if q == 0 else np.tile(fields[q], (n, 1))
            r = np.sqrt(alpha) * (bq - x / alpha + ref["eind"])
            rel = np.linalg.norm(r) / np.linalg.norm(np.sqrt(alpha) * bq)
            assert rel <= 2e-10, (g, q, rel)
        off += n
