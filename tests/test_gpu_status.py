"""Error statuses, include flags and solver-state isolation of the C ABI (GPU).

- A batch that alternates precisions (mixed PCG evaluations and FP64 CG responses) must not replay
  the captured iteration chunks of one precision in the other: the responses equal those of an
  FP64 batch, and a later mixed evaluation equals the first one.
- A system whose operator (alpha^-1 - T) is indefinite (undamped pairs with large polarizabilities)
  makes the 16-bit factorization fail; the dense responses then fall back to FP64 CG.  Responses that
  no method brings below the tolerance (non-finite pair terms, too few CG iterations) return
  TDP_ERR_NOT_CONVERGED with the statistics filled, and every kernel still exits normally.
- include_forces / include_energy (TholeDipoleKernels.h:64-68) skip the torque mapping and the
  energy reduction without changing what is returned.
- TDP_FLAG_REQUIRE_CONVERGENCE turns a Mutual solve that stops at its iteration limit into an error.
"""
import numpy as np
import pytest

import ref_systems
from tholedipoleplugin_amd import _lib, systems
from tholedipoleplugin_amd.force import TholeDipoleForce

pytestmark = pytest.mark.gpu

T = TholeDipoleForce
DEBYE = 0.020819434
MU_TOL = 1e-6 * DEBYE


def _say(msg):
    print(msg, flush=True)


def _clusters(sizes, seed0):
    members = []
    for k, m in enumerate(sizes):
        c = systems.water_cluster(m, seed=seed0 + k)
        c.force.setMutualInducedTargetEpsilon(1e-10)
        c.force.setMutualInducedMaxIterations(500)
        members.append((c.force, np.asarray(c.positions, np.float64)))
    return members


def test_cg_responses_between_mixed_evaluations():
    _say("case mixed batch: execute, cg responses, execute")
    members = _clusters([12, 40, 90], 1200)
    pos = np.concatenate([p for _, p in members])
    forces = [f for f, _ in members]
    bm = _lib.HipBatchHandle(forces, precision="mixed", solver="pcg")
    bd = _lib.HipBatchHandle(forces, precision="double", solver="pcg")
    e1, f1 = bm.execute(pos)
    mu1 = bm.dipoles(pos, None, _lib.DIPOLES_INDUCED)
    bd.execute(pos)
    Xm, stm = bm.responses(pos, method="cg", tolerance=1e-12, max_steps=1000)
    Xd, std = bd.responses(pos, method="cg", tolerance=1e-12, max_steps=1000)
    _say("  cg stats mixed ctx %s, double ctx %s" % (stm, std))
    assert stm["max_relative_residual"] <= 1e-12 and std["max_relative_residual"] <= 1e-12
    assert np.max(np.abs(Xm - Xd)) <= 1e-3 * MU_TOL
    e2, f2 = bm.execute(pos)
    mu2 = bm.dipoles(pos, None, _lib.DIPOLES_INDUCED)
    bm.close()
    bd.close()
    assert np.max(np.abs(mu2 - mu1)) <= 1e-3 * MU_TOL
    assert np.max(np.abs(e2 - e1)) <= 1e-10 * np.max(np.abs(e1))
    assert np.max(np.abs(f2 - f1)) <= 1e-10 * np.max(np.abs(f1))


def _indefinite_pair(separation=0.1):
    """Two undamped particles 0.1 nm apart with alpha = 1 nm^3: alpha^-1 - T has the eigenvalues
    1 -+ 2/r^3 along the axis and 1 -+ 1/r^3 across it (four distinct values, two negative), so the
    operator is indefinite and the 16-bit Cholesky meets a negative pivot.  separation 0 puts both
    particles at one point: the pair terms are not finite."""
    f = T()
    f.addParticle(1.0, [0.0, 0.0, 0.0], 1.0, 0.0, T.NoAxisType, -1, -1, -1)
    f.addParticle(-1.0, [0.0, 0.0, 0.0], 1.0, 0.0, T.NoAxisType, -1, -1, -1)
    f.setMutualInducedTargetEpsilon(1e-8)
    f.setMutualInducedMaxIterations(50)
    return f, np.array([[0.0, 0.0, 0.0], [separation, 0.0, 0.0]])


def _indefinite_triple():
    """Three undamped particles with unequal charges and polarizabilities, not collinear: alpha^-1 - T
    has nine distinct eigenvalues (five negative) and Jacobi-preconditioned CG needs six iterations
    for E0 (checked in FP64 with a dense assembly of the same operator)."""
    f = T()
    for q, a in ((1.0, 1.0), (-0.5, 2.0), (0.3, 0.5)):
        f.addParticle(q, [0.0, 0.0, 0.0], a, 0.0, T.NoAxisType, -1, -1, -1)
    f.setMutualInducedTargetEpsilon(1e-8)
    f.setMutualInducedMaxIterations(50)
    return f, np.array([[0.0, 0.0, 0.0], [0.1, 0.0, 0.0], [0.03, 0.12, 0.05]])


def _batch_with(f_bad, p_bad):
    members = _clusters([7], 1300)
    forces = [members[0][0], f_bad]
    pos = np.concatenate([members[0][1], p_bad])
    return _lib.HipBatchHandle(forces, precision="double", solver="pcg"), pos


@pytest.mark.parametrize("method", ["fp16", "bf16"])
def test_indefinite_operator_dense_falls_back_to_cg(method):
    """The factorization reports the non-positive pivot; the refinement cannot reach the tolerance
    with that factor, so the batch is solved by FP64 CG, which converges here (four distinct
    eigenvalues): the call succeeds and its responses equal a direct CG call."""
    _say("case indefinite operator, responses %s" % method)
    b, pos = _batch_with(*_indefinite_pair())
    X, st, rc = b.responses(pos, method=method, tolerance=1e-10, max_steps=20, check=False)
    _say("  status %d, stats %s" % (rc, st))
    Y, stc = b.responses(pos, method="cg", tolerance=1e-10, max_steps=200)
    b.close()
    assert st["factor_failed"] != 0
    assert rc == _lib.TDP_OK and st["max_relative_residual"] <= 1e-10
    assert np.max(np.abs(X - Y)) <= 1e-9 * np.max(np.abs(Y))


@pytest.mark.parametrize("method", ["fp16", "cg"])
def test_unsolvable_responses_report_not_converged(method):
    """Coincident particles (non-finite pair terms), and CG limited to two iterations on a system
    that needs six, end above the tolerance: TDP_ERR_NOT_CONVERGED, kernels exit normally."""
    _say("case unsolvable responses %s" % method)
    b, pos = _batch_with(*_indefinite_pair(0.0))
    _, st, rc = b.responses(pos, method=method, tolerance=1e-10, max_steps=20, check=False)
    _say("  coincident: status %d, stats %s" % (rc, st))
    assert rc == _lib.TDP_ERR_NOT_CONVERGED
    assert not (st["max_relative_residual"] <= 1e-10)
    b.close()
    if method == "cg":
        b, pos = _batch_with(*_indefinite_triple())
        _, st, rc = b.responses(pos, method="cg", tolerance=1e-10, max_steps=2, check=False)
        _say("  two CG iterations: status %d, stats %s" % (rc, st))
        assert rc == _lib.TDP_ERR_NOT_CONVERGED
        b.close()


def test_include_flags():
    _say("case include flags")
    rs = ref_systems.water4(T.Mutual)
    rs.force.setMutualInducedTargetEpsilon(1e-12)
    rs.force.setMutualInducedMaxIterations(500)
    h = _lib.HipContextHandle(rs.force, rs.box, precision="double", solver="pcg")
    pos = np.asarray(rs.positions, np.float64)
    e_all, f_all = h.execute(pos, rs.box)
    seed = np.full(pos.size, 7.0)
    e_nf, f_nf = h.execute(pos, rs.box, forces=seed.copy(), include_forces=False)
    assert np.all(f_nf == 7.0)
    assert e_nf == e_all
    e_ne, f_ne = h.execute(pos, rs.box, include_energy=False)
    assert e_ne == 0.0
    assert np.array_equal(f_ne, f_all)
    h.close()


def test_require_convergence_flag():
    _say("case require-convergence flag")
    members = _clusters([20], 1400)
    f, pos = members[0]
    f.setMutualInducedMaxIterations(2)
    f.setMutualInducedTargetEpsilon(1e-12)
    plain = _lib.HipContextHandle(f, np.zeros(9), precision="double", solver="pcg")
    e0, f0 = plain.execute(pos, None)
    st = plain.solver_stats()
    plain.close()
    assert st["converged"] == 0 and st["iterations"] == 2
    strict = _lib.HipContextHandle(f, np.zeros(9), precision="double", solver="pcg",
                                   flags=_lib.FLAG_REQUIRE_CONVERGENCE)
    with pytest.raises(_lib.TdpError) as ei:
        strict.execute(pos, None)
    assert ei.value.code == _lib.TDP_ERR_NOT_CONVERGED
    assert "did not converge" in str(ei.value)
    f.setMutualInducedMaxIterations(500)
    strict.close()
    ok = _lib.HipContextHandle(f, np.zeros(9), precision="double", solver="pcg",
                               flags=_lib.FLAG_REQUIRE_CONVERGENCE)
    ok.execute(pos, None)
    assert ok.solver_stats()["converged"] == 1
    ok.close()
