"""ctypes wrapper of the CPU oracle (libtdp_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg as the checker.  The product never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libtdp_oracle.so")
_lib = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class _Sys(C.Structure):
    _fields_ = [("n", C.c_int), ("charge", _dp), ("dipole", _dp), ("polarity", _dp), ("thole", _dp),
                ("axis_type", _ip), ("atom_z", _ip), ("atom_x", _ip), ("atom_y", _ip),
                ("cov_offsets", _ip), ("cov_atoms", _ip), ("method", C.c_int), ("box", C.c_double * 3),
                ("polarization", C.c_int), ("max_iter", C.c_int), ("target_eps", C.c_double),
                ("extrap", _dp), ("n_extrap", C.c_int)]


class _Ewald(C.Structure):
    _fields_ = [("cutoff", C.c_double), ("alpha", C.c_double), ("grid", C.c_int * 3), ("order", C.c_int),
                ("kmax", C.c_int * 3)]


def _ewald(cutoff, alpha, grid=(0, 0, 0), order=5, kmax=None):
    e = _Ewald()
    e.cutoff = float(cutoff)
    e.alpha = float(alpha)
    for k in range(3):
        e.grid[k] = int(grid[k])
        e.kmax[k] = int(kmax[k]) if kmax is not None else 0
    e.order = int(order)
    return e


class _Stats(C.Structure):
    _fields_ = [("energy", C.c_double), ("iterations", C.c_int), ("epsilon", C.c_double), ("converged", C.c_int)]


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        L.tdpo_force_energy.argtypes = [C.POINTER(_Sys), _dp, _dp, C.POINTER(_Stats), _dp, _dp, _dp, _dp]
        L.tdpo_induced_dipoles.argtypes = [C.POINTER(_Sys), _dp, _dp, _dp, _dp, C.POINTER(_Stats)]
        L.tdpo_potential.argtypes = [C.POINTER(_Sys), _dp, _dp, C.c_int, _dp]
        L.tdpo_multipole_moments.argtypes = [C.POINTER(_Sys), _dp, _dp, _dp]
        L.tdpo_pair_scales.argtypes = [C.POINTER(_Sys), C.c_int, C.c_int, _dp, _dp]
        L.tdpo_damped_rr.argtypes = [C.c_double, C.c_double, C.c_double, _dp, _dp]
        L.tdpo_qi_frame.argtypes = [_dp, _dp, _dp, C.c_double, _dp]
        L.tdpo_pair_electrostatic.restype = C.c_double
        L.tdpo_pair_electrostatic.argtypes = [_dp, _dp, C.c_double, C.c_double, _dp, _dp, _dp, _dp,
                                              C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, _dp,
                                              _dp, _dp, _dp]
        L.tdpo_lab_dipoles.argtypes = [C.POINTER(_Sys), _dp, _dp]
        L.tdpo_map_torques.argtypes = [C.POINTER(_Sys), _dp, _dp, _dp]
        L.tdpo_set_threads.argtypes = [C.c_int]
        L.tdpo_pass_slice.restype = C.c_double
        L.tdpo_pass_slice.argtypes = [C.POINTER(_Sys), _dp, _dp, C.c_int, C.c_int, C.c_int, _dp]
        L.tdpo_rows.argtypes = [C.POINTER(_Sys), _dp, _dp, _ip, C.c_int, _dp, _dp, _dp, _dp, _dp]
        L.tdpo_evaluate_at.argtypes = [C.POINTER(_Sys), _dp, _dp, _dp, _dp, _dp, _dp]
        L.tdpo_ewald_fields.argtypes = [C.POINTER(_Sys), C.POINTER(_Ewald), _dp, _dp, _dp, _dp]
        L.tdpo_ewald_energy_at.argtypes = [C.POINTER(_Sys), C.POINTER(_Ewald), _dp, _dp, _dp, _dp, _dp]
        L.tdpo_ewald_evaluate.argtypes = [C.POINTER(_Sys), C.POINTER(_Ewald), _dp, _dp, _dp, _dp, _dp,
                                          C.POINTER(_Stats)]
        L.tdpo_fft3.argtypes = [_dp, C.c_int, C.c_int, C.c_int, C.c_int]
        L.tdpo_fft3.restype = None
        L.tdpo_bspline.argtypes = [C.c_double, C.c_int, _dp, _dp, _dp]
        L.tdpo_bspline.restype = None
        L.tdpo_cutoff_pairs.argtypes = [C.POINTER(_Sys), _dp, C.c_double, C.c_longlong, _ip]
        L.tdpo_cutoff_pairs.restype = C.c_longlong
        _lib = L
    return _lib


def _p(a, t=_dp):
    return a.ctypes.data_as(t)


class OracleSystem:
    """Packs a TholeDipoleForce (+ box) into the oracle's C struct; keeps the arrays alive."""

    def __init__(self, force, box=None, polarization=None, target_eps=None, max_iter=None):
        a = force.packed_arrays()
        self._arrays = a
        self.n = force.getNumParticles()
        s = _Sys()
        s.n = self.n
        s.charge = _p(a["charge"]); s.dipole = _p(a["dipole"]); s.polarity = _p(a["polarity"])
        s.thole = _p(a["thole"])
        s.axis_type = _p(a["axis_type"], _ip); s.atom_z = _p(a["atom_z"], _ip)
        s.atom_x = _p(a["atom_x"], _ip); s.atom_y = _p(a["atom_y"], _ip)
        s.cov_offsets = _p(a["cov_offsets"], _ip); s.cov_atoms = _p(a["cov_atoms"], _ip)
        s.method = force.getNonbondedMethod()
        b = np.zeros(3) if box is None else np.diag(np.asarray(box, np.float64).reshape(3, 3))
        for k in range(3):
            s.box[k] = b[k]
        s.polarization = force.getPolarizationType() if polarization is None else polarization
        s.max_iter = force.getMutualInducedMaxIterations() if max_iter is None else max_iter
        s.target_eps = force.getMutualInducedTargetEpsilon() if target_eps is None else target_eps
        self._ex = np.asarray(force.getExtrapolationCoefficients(), np.float64)
        s.extrap = _p(self._ex) if self._ex.size else None
        s.n_extrap = int(self._ex.size)
        self.s = s


def _check(rc):
    if rc != 0:
        raise RuntimeError("oracle rejected the input (code %d)" % rc)


def force_energy(force, positions, box=None, threads=1, **kw):
    """Returns dict(energy, forces, induced, labdip, torques, fixed_field, iterations, epsilon, converged)."""
    L = lib()
    L.tdpo_set_threads(int(threads))
    os_ = OracleSystem(force, box, **kw)
    n = os_.n
    pos = np.ascontiguousarray(positions, np.float64).reshape(-1)
    f = np.zeros(3 * n)
    ind = np.zeros(3 * n); lab = np.zeros(3 * n); tq = np.zeros(3 * n); e0 = np.zeros(3 * n)
    st = _Stats()
    _check(L.tdpo_force_energy(C.byref(os_.s), _p(pos), _p(f), C.byref(st), _p(ind), _p(lab), _p(tq), _p(e0)))
    L.tdpo_set_threads(1)
    return dict(energy=st.energy, forces=f.reshape(n, 3), induced=ind.reshape(n, 3), labdip=lab.reshape(n, 3),
                torques=tq.reshape(n, 3), fixed_field=e0.reshape(n, 3), iterations=st.iterations,
                epsilon=st.epsilon, converged=st.converged)


def induced_dipoles(force, positions, box=None, threads=1, **kw):
    L = lib()
    L.tdpo_set_threads(int(threads))
    os_ = OracleSystem(force, box, **kw)
    n = os_.n
    pos = np.ascontiguousarray(positions, np.float64).reshape(-1)
    ind = np.zeros(3 * n); lab = np.zeros(3 * n); e0 = np.zeros(3 * n)
    st = _Stats()
    _check(L.tdpo_induced_dipoles(C.byref(os_.s), _p(pos), _p(ind), _p(lab), _p(e0), C.byref(st)))
    L.tdpo_set_threads(1)
    return dict(induced=ind.reshape(n, 3), labdip=lab.reshape(n, 3), fixed_field=e0.reshape(n, 3),
                iterations=st.iterations, epsilon=st.epsilon, converged=st.converged)


def potential(force, positions, grid, box=None):
    os_ = OracleSystem(force, box)
    pos = np.ascontiguousarray(positions, np.float64).reshape(-1)
    g = np.ascontiguousarray(grid, np.float64).reshape(-1)
    out = np.zeros(g.size // 3)
    _check(lib().tdpo_potential(C.byref(os_.s), _p(pos), _p(g), int(out.size), _p(out)))
    return out


def multipole_moments(force, positions, masses, box=None):
    os_ = OracleSystem(force, box)
    pos = np.ascontiguousarray(positions, np.float64).reshape(-1)
    m = np.ascontiguousarray(masses, np.float64)
    out = np.zeros(13)
    _check(lib().tdpo_multipole_moments(C.byref(os_.s), _p(pos), _p(m), _p(out)))
    return out


def pair_scales(force, i, j):
    os_ = OracleSystem(force)
    m = C.c_double(); s = C.c_double()
    _check(lib().tdpo_pair_scales(C.byref(os_.s), i, j, C.byref(m), C.byref(s)))
    return m.value, s.value


def damped_rr(ai, aj, r):
    a = C.c_double(); b = C.c_double()
    lib().tdpo_damped_rr(ai, aj, r, C.byref(a), C.byref(b))
    return a.value, b.value


def qi_frame(pi, pj, box=None):
    pi = np.asarray(pi, np.float64); pj = np.asarray(pj, np.float64)
    d = pj - pi
    if box is not None:
        L = np.asarray(box, np.float64)
        d = d - L * np.floor(d / L + 0.5)
    r = float(np.sqrt(d @ d))
    out = np.zeros(9)
    lib().tdpo_qi_frame(_p(np.ascontiguousarray(pi)), _p(np.ascontiguousarray(pj)), _p(np.ascontiguousarray(d)), r, _p(out))
    return out.reshape(3, 3)


def pair_electrostatic(pi, pk, qi, qk, di, dk, ui, uk, ai, ak, m_scale=1.0, i_scale=1.0, polarization=1, box=None):
    """One calculateElectrostaticPairIxn call; returns (energy, force_on_k, torque_i, torque_k)."""
    arrs = [np.ascontiguousarray(x, np.float64) for x in (pi, pk, di, dk, ui, uk)]
    fk = np.zeros(3); ti = np.zeros(3); tk = np.zeros(3)
    b = None if box is None else np.ascontiguousarray(box, np.float64)
    e = lib().tdpo_pair_electrostatic(_p(arrs[0]), _p(arrs[1]), qi, qk, _p(arrs[2]), _p(arrs[3]), _p(arrs[4]),
                                      _p(arrs[5]), ai, ak, m_scale, i_scale, polarization,
                                      None if b is None else _p(b), _p(fk), _p(ti), _p(tk))
    return e, fk, ti, tk


def lab_dipoles(force, positions):
    os_ = OracleSystem(force)
    pos = np.ascontiguousarray(positions, np.float64).reshape(-1)
    out = np.zeros(pos.size)
    _check(lib().tdpo_lab_dipoles(C.byref(os_.s), _p(pos), _p(out)))
    return out.reshape(-1, 3)


def map_torques(force, positions, torques):
    os_ = OracleSystem(force)
    pos = np.ascontiguousarray(positions, np.float64).reshape(-1)
    t = np.ascontiguousarray(torques, np.float64).reshape(-1)
    f = np.zeros(pos.size)
    _check(lib().tdpo_map_torques(C.byref(os_.s), _p(pos), _p(t), _p(f)))
    return f.reshape(-1, 3)


def rows(force, positions, mu, row_idx, box=None, threads=8):
    """Full-size rows for particles row_idx given induced dipoles mu: dict(e0, eind, force, torque, energy)."""
    L = lib()
    L.tdpo_set_threads(int(threads))
    os_ = OracleSystem(force, box)
    pos = np.ascontiguousarray(positions, np.float64).reshape(-1)
    m = np.ascontiguousarray(mu, np.float64).reshape(-1)
    r = np.ascontiguousarray(row_idx, np.int32)
    k = r.size
    e0 = np.zeros(3 * k); ei = np.zeros(3 * k); f = np.zeros(3 * k); t = np.zeros(3 * k); en = np.zeros(k)
    _check(L.tdpo_rows(C.byref(os_.s), _p(pos), _p(m), _p(r, _ip), k, _p(e0), _p(ei), _p(f), _p(t), _p(en)))
    L.tdpo_set_threads(1)
    return dict(e0=e0.reshape(k, 3), eind=ei.reshape(k, 3), force=f.reshape(k, 3), torque=t.reshape(k, 3), energy=en)


def evaluate_at(force, positions, mu, box=None, threads=8):
    """The reference's pair passes at given induced dipoles: dict(e0, eind, forces, energy)."""
    L = lib()
    L.tdpo_set_threads(int(threads))
    os_ = OracleSystem(force, box)
    n = os_.n
    pos = np.ascontiguousarray(positions, np.float64).reshape(-1)
    m = np.ascontiguousarray(mu, np.float64).reshape(-1)
    e0 = np.zeros(3 * n); ei = np.zeros(3 * n); f = np.zeros(3 * n); e = C.c_double(0.0)
    _check(L.tdpo_evaluate_at(C.byref(os_.s), _p(pos), _p(m), _p(e0), _p(ei), _p(f), C.byref(e)))
    L.tdpo_set_threads(1)
    return dict(e0=e0.reshape(n, 3), eind=ei.reshape(n, 3), forces=f.reshape(n, 3), energy=e.value)


# ---- Ewald sums (tdp_ewald.c): the model of the HIP library's Ewald option
def ewald_evaluate(force, positions, box, cutoff, alpha, grid=(0, 0, 0), order=5, kmax=None, threads=8, **kw):
    """E0, the induced-dipole solve, energy and forces by Ewald summation: dict(energy, forces,
    induced, e0, iterations, epsilon, converged)."""
    L = lib()
    L.tdpo_set_threads(int(threads))
    os_ = OracleSystem(force, box, **kw)
    n = os_.n
    pos = np.ascontiguousarray(positions, np.float64).reshape(-1)
    f = np.zeros(3 * n); mu = np.zeros(3 * n); e0 = np.zeros(3 * n); en = C.c_double(0.0)
    st = _Stats()
    e = _ewald(cutoff, alpha, grid, order, kmax)
    _check(L.tdpo_ewald_evaluate(C.byref(os_.s), C.byref(e), _p(pos), _p(f), C.byref(en), _p(mu), _p(e0),
                                 C.byref(st)))
    L.tdpo_set_threads(1)
    return dict(energy=en.value, forces=f.reshape(n, 3), induced=mu.reshape(n, 3), e0=e0.reshape(n, 3),
                iterations=st.iterations, epsilon=st.epsilon, converged=st.converged)


def ewald_energy_at(force, positions, mu, box, cutoff, alpha, grid=(0, 0, 0), order=5, kmax=None, threads=8):
    """Energy, forces and E0 by Ewald summation at GIVEN induced dipoles: dict(energy, forces, e0)."""
    L = lib()
    L.tdpo_set_threads(int(threads))
    os_ = OracleSystem(force, box)
    n = os_.n
    pos = np.ascontiguousarray(positions, np.float64).reshape(-1)
    m = np.ascontiguousarray(mu, np.float64).reshape(-1)
    f = np.zeros(3 * n); e0 = np.zeros(3 * n); en = C.c_double(0.0)
    e = _ewald(cutoff, alpha, grid, order, kmax)
    _check(L.tdpo_ewald_energy_at(C.byref(os_.s), C.byref(e), _p(pos), _p(m), _p(f), C.byref(en), _p(e0)))
    L.tdpo_set_threads(1)
    return dict(energy=en.value, forces=f.reshape(n, 3), e0=e0.reshape(n, 3))


def ewald_fields(force, positions, mu, box, cutoff, alpha, grid=(0, 0, 0), order=5, kmax=None, threads=8):
    """E0 and T mu (with the Ewald self-field) by Ewald summation at given dipoles: (e0, tmu)."""
    L = lib()
    L.tdpo_set_threads(int(threads))
    os_ = OracleSystem(force, box)
    n = os_.n
    pos = np.ascontiguousarray(positions, np.float64).reshape(-1)
    m = np.ascontiguousarray(mu, np.float64).reshape(-1)
    e0 = np.zeros(3 * n); tm = np.zeros(3 * n)
    e = _ewald(cutoff, alpha, grid, order, kmax)
    _check(L.tdpo_ewald_fields(C.byref(os_.s), C.byref(e), _p(pos), _p(m), _p(e0), _p(tm)))
    L.tdpo_set_threads(1)
    return e0.reshape(n, 3), tm.reshape(n, 3)


def cutoff_pairs(force, positions, box, cutoff):
    """Pairs (i < j) within the cutoff by the library's FP64 neighbour criterion, i then j ascending."""
    L = lib()
    os_ = OracleSystem(force, box)
    pos = np.ascontiguousarray(positions, np.float64).reshape(-1)
    n = L.tdpo_cutoff_pairs(C.byref(os_.s), _p(pos), float(cutoff), 0, None)
    out = np.zeros(2 * max(n, 1), np.int32)
    L.tdpo_cutoff_pairs(C.byref(os_.s), _p(pos), float(cutoff), n, _p(out, _ip))
    return out[:2 * n].reshape(-1, 2).astype(np.int64)


def fft3(a, sign):
    """Unnormalised 3D DFT of a complex array (sign -1 forward, +1 inverse)."""
    b = np.ascontiguousarray(a, np.complex128).copy()
    nx, ny, nz = b.shape
    lib().tdpo_fft3(b.ctypes.data_as(_dp), nx, ny, nz, int(sign))
    return b


def bspline(w, order):
    th = np.zeros(order); d1 = np.zeros(order); d2 = np.zeros(order)
    lib().tdpo_bspline(float(w), int(order), _p(th), _p(d1), _p(d2))
    return th, d1, d2


def time_pass_slices(force, positions, mu, box, which, slices, threads=1):
    """Times reference pair-pass slices (rows [i0, i1), partners j > i); returns (seconds, pairs)."""
    import time
    L = lib()
    L.tdpo_set_threads(int(threads))
    os_ = OracleSystem(force, box)
    pos = np.ascontiguousarray(positions, np.float64).reshape(-1)
    m = np.ascontiguousarray(mu, np.float64).reshape(-1)
    out = np.zeros(pos.size)
    secs = 0.0
    pairs = 0.0
    for i0, i1 in slices:
        t = time.perf_counter()
        p = L.tdpo_pass_slice(C.byref(os_.s), _p(pos), _p(m), int(which), int(i0), int(i1), _p(out))
        secs += time.perf_counter() - t
        if p < 0:
            raise RuntimeError("oracle rejected the input")
        pairs += p
    L.tdpo_set_threads(1)
    return secs, pairs
