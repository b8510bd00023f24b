"""ctypes binding of libtholedipole_hip.so (C ABI declared in include/tholedipole_hip.h).

There is no fallback: if the library is missing or cannot be loaded, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libtholedipole_hip.so")

TDP_OK = 0
TDP_ERR_INVALID_ARGUMENT = 1
TDP_ERR_HIP = 2
TDP_ERR_NOT_PME = 3
TDP_ERR_BOX_TOO_SMALL = 4
TDP_ERR_PARTICLE_COUNT = 5
TDP_ERR_NO_DEVICE = 6
TDP_ERR_NOT_CONVERGED = 7

FLAG_REQUIRE_CONVERGENCE = 1
FLAG_EWALD = 2

PRECISION = {"double": 0, "mixed": 1}
SOLVER = {"jacobi": 0, "pcg": 1}
DIPOLES_INDUCED, DIPOLES_LAB_PERMANENT, DIPOLES_TOTAL = 0, 1, 2

# Every symbol include/tholedipole_hip.h declares.
EXPORTED_SYMBOLS = [
    "tdp_create", "tdp_destroy", "tdp_last_error", "tdp_last_create_error", "tdp_copy_parameters",
    "tdp_execute", "tdp_execute_device", "tdp_execute_device_noalloc", "tdp_get_neighbor_pairs", "tdp_get_ewald_info",
    "tdp_get_dipoles",
    "tdp_get_electrostatic_potential",
    "tdp_get_system_multipole_moments", "tdp_get_pme_parameters", "tdp_get_solver_stats", "tdp_set_timing",
    "tdp_autotune", "tdp_get_work_schedule", "tdp_set_work_schedule", "tdp_set_warm_start",
    "tdp_batch_create", "tdp_batch_num_systems", "tdp_batch_execute", "tdp_batch_execute_device",
    "tdp_batch_get_solver_stats", "tdp_batch_solve_responses",
]

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class ForceDesc(C.Structure):
    _fields_ = [("num_particles", C.c_int32), ("charges", _dp), ("dipoles", _dp), ("polarizabilities", _dp),
                ("thole", _dp), ("axis_types", _ip), ("atom_z", _ip), ("atom_x", _ip), ("atom_y", _ip),
                ("covalent_offsets", _ip), ("covalent_atoms", _ip), ("nonbonded_method", C.c_int32),
                ("polarization_type", C.c_int32), ("cutoff", C.c_double), ("ewald_error_tolerance", C.c_double),
                ("pme_alpha", C.c_double), ("pme_grid", C.c_int32 * 3), ("mutual_max_iterations", C.c_int32),
                ("mutual_target_epsilon", C.c_double), ("extrapolation_coefficients", _dp),
                ("num_extrapolation_coefficients", C.c_int32), ("default_box", C.c_double * 9)]


class Options(C.Structure):
    _fields_ = [("precision", C.c_int32), ("solver", C.c_int32), ("device", C.c_int32), ("flags", C.c_int32)]


class DenseStats(C.Structure):
    _fields_ = [("steps", C.c_int32), ("max_relative_residual", C.c_double), ("factor_failed", C.c_int32),
                ("assemble_ms", C.c_double), ("factor_ms", C.c_double)]


class SolverStats(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("epsilon", C.c_double), ("converged", C.c_int32),
                ("matvec_ms", C.c_double), ("matvec_launches", C.c_int32), ("matvec_ms_full", C.c_double),
                ("matvec_full_launches", C.c_int32)]


class EwaldInfo(C.Structure):
    _fields_ = [("alpha", C.c_double), ("grid", C.c_int32 * 3), ("cutoff", C.c_double), ("r_fp64", C.c_double),
                ("stored", C.c_int32), ("entries32", C.c_int64), ("entries64", C.c_int64), ("slots32", C.c_int64),
                ("slots64", C.c_int64), ("bytes_per_entry32", C.c_int32), ("bytes_per_entry64", C.c_int32),
                ("tiles", C.c_int32), ("max_neighbor_tiles", C.c_int32), ("mean_neighbor_tiles", C.c_double)]


_lib = None


class TdpError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(message)
        self.code = code


def load():
    """Loads the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libtholedipole_hip.so is not built (run __graft_entry__.build() or make -C "
                           "tholedipoleplugin_amd/csrc); there is no CPU fallback")
    L = C.CDLL(LIB_PATH)
    ctx_p = C.c_void_p
    L.tdp_create.argtypes = [C.POINTER(ForceDesc), C.POINTER(Options), C.POINTER(ctx_p)]
    L.tdp_destroy.argtypes = [ctx_p]
    L.tdp_destroy.restype = None
    L.tdp_last_error.argtypes = [ctx_p]
    L.tdp_last_error.restype = C.c_char_p
    L.tdp_last_create_error.argtypes = []
    L.tdp_last_create_error.restype = C.c_char_p
    L.tdp_copy_parameters.argtypes = [ctx_p, C.POINTER(ForceDesc)]
    L.tdp_execute.argtypes = [ctx_p, _dp, _dp, C.c_int32, C.c_int32, _dp, _dp]
    L.tdp_execute_device.argtypes = [ctx_p, C.c_void_p, _dp, C.c_void_p, C.c_void_p, C.c_void_p]
    L.tdp_execute_device_noalloc.argtypes = [ctx_p, C.c_void_p, _dp, C.c_void_p, C.c_void_p, C.c_void_p]
    L.tdp_get_dipoles.argtypes = [ctx_p, _dp, _dp, C.c_int32, _dp]
    L.tdp_get_neighbor_pairs.argtypes = [ctx_p, _dp, _dp, C.c_int64, _ip, C.POINTER(C.c_int64)]
    L.tdp_get_ewald_info.argtypes = [ctx_p, C.POINTER(EwaldInfo)]
    L.tdp_get_electrostatic_potential.argtypes = [ctx_p, _dp, _dp, _dp, C.c_int32, _dp]
    L.tdp_get_system_multipole_moments.argtypes = [ctx_p, _dp, _dp, _dp, _dp]
    L.tdp_get_pme_parameters.argtypes = [ctx_p, _dp, _ip, _ip, _ip]
    L.tdp_get_solver_stats.argtypes = [ctx_p, C.POINTER(SolverStats)]
    L.tdp_set_timing.argtypes = [ctx_p, C.c_int32]
    L.tdp_autotune.argtypes = [ctx_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_int32)]
    L.tdp_get_work_schedule.argtypes = [ctx_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
    L.tdp_set_work_schedule.argtypes = [ctx_p, C.POINTER(C.c_int32)]
    L.tdp_set_warm_start.argtypes = [ctx_p, C.c_int32]
    L.tdp_batch_create.argtypes = [C.POINTER(ForceDesc), C.c_int32, C.POINTER(Options), C.POINTER(ctx_p)]
    L.tdp_batch_num_systems.argtypes = [ctx_p]
    L.tdp_batch_execute.argtypes = [ctx_p, _dp, _dp, _dp]
    L.tdp_batch_execute_device.argtypes = [ctx_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.tdp_batch_get_solver_stats.argtypes = [ctx_p, C.c_int32, C.POINTER(SolverStats)]
    L.tdp_batch_solve_responses.argtypes = [ctx_p, _dp, C.c_int32, C.c_double, C.c_int32, _dp, C.POINTER(DenseStats)]
    for name in EXPORTED_SYMBOLS:
        if name not in ("tdp_destroy", "tdp_last_error", "tdp_last_create_error"):
            getattr(L, name).restype = C.c_int
    _lib = L
    return L


def _ptr(a, t=_dp):
    return a.ctypes.data_as(t)


class PackedForce:
    """tdp_force_desc over numpy arrays owned by this object."""

    def __init__(self, force, default_box):
        a = force.packed_arrays()
        self.arrays = a
        self.extrap = np.asarray(force.getExtrapolationCoefficients(), np.float64)
        d = ForceDesc()
        d.num_particles = force.getNumParticles()
        d.charges = _ptr(a["charge"]); d.dipoles = _ptr(a["dipole"]); d.polarizabilities = _ptr(a["polarity"])
        d.thole = _ptr(a["thole"])
        d.axis_types = _ptr(a["axis_type"], _ip); d.atom_z = _ptr(a["atom_z"], _ip)
        d.atom_x = _ptr(a["atom_x"], _ip); d.atom_y = _ptr(a["atom_y"], _ip)
        d.covalent_offsets = _ptr(a["cov_offsets"], _ip); d.covalent_atoms = _ptr(a["cov_atoms"], _ip)
        d.nonbonded_method = force.getNonbondedMethod()
        d.polarization_type = force.getPolarizationType()
        d.cutoff = force.getCutoffDistance()
        d.ewald_error_tolerance = force.getEwaldErrorTolerance()
        alpha, nx, ny, nz = force.getPMEParameters()
        d.pme_alpha = alpha
        d.pme_grid[0], d.pme_grid[1], d.pme_grid[2] = nx, ny, nz
        d.mutual_max_iterations = force.getMutualInducedMaxIterations()
        d.mutual_target_epsilon = force.getMutualInducedTargetEpsilon()
        d.extrapolation_coefficients = _ptr(self.extrap) if self.extrap.size else None
        d.num_extrapolation_coefficients = int(self.extrap.size)
        b = np.asarray(default_box, np.float64).reshape(9)
        for k in range(9):
            d.default_box[k] = b[k]
        self.desc = d


class HipContextHandle:
    """Owns one tdp_context; methods raise TdpError on a nonzero status."""

    def __init__(self, force, default_box, precision="double", solver="jacobi", device=0, flags=0):
        self.L = load()
        self.n = force.getNumParticles()
        pf = PackedForce(force, default_box)
        opts = Options(PRECISION[precision], SOLVER[solver], int(device), int(flags))
        h = C.c_void_p()
        rc = self.L.tdp_create(C.byref(pf.desc), C.byref(opts), C.byref(h))
        if rc != TDP_OK:
            raise TdpError(rc, self.L.tdp_last_create_error().decode())
        self.h = h
        self.precision = precision
        self.solver = solver

    def _check(self, rc):
        if rc != TDP_OK:
            raise TdpError(rc, self.L.tdp_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.L.tdp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _box(box):
        if box is None:
            return None
        return np.ascontiguousarray(np.asarray(box, np.float64).reshape(9))

    def copy_parameters(self, force, default_box):
        pf = PackedForce(force, default_box)
        self._check(self.L.tdp_copy_parameters(self.h, C.byref(pf.desc)))

    def execute(self, positions, box, forces=None, include_forces=True, include_energy=True):
        pos = np.ascontiguousarray(positions, np.float64).reshape(-1)
        f = np.zeros(pos.size) if forces is None else forces
        e = C.c_double(0.0)
        b = self._box(box)
        self._check(self.L.tdp_execute(self.h, _ptr(pos), None if b is None else _ptr(b), int(bool(include_forces)),
                                       int(bool(include_energy)), _ptr(f), C.byref(e)))
        return e.value, f.reshape(-1, 3)

    def execute_device(self, d_positions_ptr, box, d_forces_ptr, d_energy_ptr, stream_ptr):
        b = self._box(box)
        self._check(self.L.tdp_execute_device(self.h, C.c_void_p(d_positions_ptr), None if b is None else _ptr(b),
                                              C.c_void_p(d_forces_ptr), C.c_void_p(d_energy_ptr),
                                              C.c_void_p(stream_ptr)))

    def execute_device_noalloc(self, d_positions_ptr, box, d_forces_ptr, d_energy_ptr, stream_ptr, check=True):
        """tdp_execute_device_noalloc: the device entry point with its partial sums in the library's static
        scratch (no allocation).  With check=False the status is returned instead of raised."""
        b = self._box(box)
        rc = self.L.tdp_execute_device_noalloc(self.h, C.c_void_p(d_positions_ptr), None if b is None else _ptr(b),
                                               C.c_void_p(d_forces_ptr), C.c_void_p(d_energy_ptr),
                                               C.c_void_p(stream_ptr))
        if not check:
            return rc
        self._check(rc)
        return rc

    def neighbor_pairs(self, positions, box):
        """Pairs (i < j) within the cutoff as the Ewald option's real-space passes find them, sorted."""
        pos = np.ascontiguousarray(positions, np.float64).reshape(-1)
        b = self._box(box)
        cnt = C.c_int64(0)
        self._check(self.L.tdp_get_neighbor_pairs(self.h, _ptr(pos), None if b is None else _ptr(b), 0, None,
                                                  C.byref(cnt)))
        out = np.zeros(2 * max(cnt.value, 1), np.int32)
        self._check(self.L.tdp_get_neighbor_pairs(self.h, _ptr(pos), None if b is None else _ptr(b), cnt.value,
                                                  _ptr(out, _ip), C.byref(cnt)))
        p = out[:2 * cnt.value].reshape(-1, 2).astype(np.int64)
        return p[np.lexsort((p[:, 1], p[:, 0]))]

    def dipoles(self, positions, box, kind):
        pos = np.ascontiguousarray(positions, np.float64).reshape(-1)
        out = np.zeros(pos.size)
        b = self._box(box)
        self._check(self.L.tdp_get_dipoles(self.h, _ptr(pos), None if b is None else _ptr(b), kind, _ptr(out)))
        return out.reshape(-1, 3)

    def potential(self, positions, box, grid):
        pos = np.ascontiguousarray(positions, np.float64).reshape(-1)
        g = np.ascontiguousarray(grid, np.float64).reshape(-1)
        out = np.zeros(g.size // 3)
        b = self._box(box)
        self._check(self.L.tdp_get_electrostatic_potential(self.h, _ptr(pos), None if b is None else _ptr(b),
                                                           _ptr(g), int(out.size), _ptr(out)))
        return out

    def multipole_moments(self, positions, box, masses):
        pos = np.ascontiguousarray(positions, np.float64).reshape(-1)
        m = np.ascontiguousarray(masses, np.float64)
        out = np.zeros(13)
        b = self._box(box)
        self._check(self.L.tdp_get_system_multipole_moments(self.h, _ptr(pos), None if b is None else _ptr(b),
                                                            _ptr(m), _ptr(out)))
        return out

    def pme_parameters(self):
        a = C.c_double(); nx = C.c_int32(); ny = C.c_int32(); nz = C.c_int32()
        self._check(self.L.tdp_get_pme_parameters(self.h, C.byref(a), C.byref(nx), C.byref(ny), C.byref(nz)))
        return a.value, nx.value, ny.value, nz.value

    def ewald_info(self):
        """Layout of the Ewald option after the last evaluation (stored T entries and padded slots)."""
        info = EwaldInfo()
        self._check(self.L.tdp_get_ewald_info(self.h, C.byref(info)))
        return {k: (list(getattr(info, k)) if k == "grid" else getattr(info, k)) for k, _ in EwaldInfo._fields_}

    def solver_stats(self):
        st = SolverStats()
        self._check(self.L.tdp_get_solver_stats(self.h, C.byref(st)))
        return dict(iterations=st.iterations, epsilon=st.epsilon, converged=st.converged,
                    matvec_ms=st.matvec_ms, matvec_launches=st.matvec_launches, matvec_ms_full=st.matvec_ms_full,
                    matvec_full_launches=st.matvec_full_launches)

    def set_timing(self, enable=True):
        self._check(self.L.tdp_set_timing(self.h, 1 if enable else 0))

    def set_warm_start(self, enable=True):
        """Later Mutual solves start from the previous solve's induced dipoles (tdp_set_warm_start)."""
        self._check(self.L.tdp_set_warm_start(self.h, 1 if enable else 0))

    def autotune(self, positions, box=None):
        """Times the pair kernels' work-item lengths on these positions and keeps the fastest
        (tdp_autotune); returns the tiles per work item chosen."""
        pos = np.ascontiguousarray(positions, np.float64).reshape(-1)
        b = self._box(box)
        out = C.c_int32(0)
        self._check(self.L.tdp_autotune(self.h, _ptr(pos), None if b is None else _ptr(b), C.byref(out)))
        return out.value

    def work_schedule(self):
        """Tiles per work item of each size bucket and each bucket's largest tile count
        (tdp_get_work_schedule)."""
        per = (C.c_int32 * 5)()
        bound = (C.c_int32 * 5)()
        self._check(self.L.tdp_get_work_schedule(self.h, per, bound))
        return list(per), list(bound)

    def set_work_schedule(self, tiles_per_item):
        """Sets the tiles per work item of the five size buckets (tdp_set_work_schedule)."""
        if len(tiles_per_item) != 5:
            raise ValueError("one length per size bucket (5)")
        per = (C.c_int32 * 5)(*[int(v) for v in tiles_per_item])
        self._check(self.L.tdp_set_work_schedule(self.h, per))


class HipBatchHandle(HipContextHandle):
    """Owns one batch context over several NoCutoff systems (tdp_batch_create); particles are
    numbered consecutively across the systems."""

    def __init__(self, forces, precision="double", solver="pcg", device=0, flags=0):
        self.L = load()
        self.packed = [PackedForce(f, np.zeros(9)) for f in forces]
        self.sizes = [f.getNumParticles() for f in forces]
        self.n = int(sum(self.sizes))
        descs = (ForceDesc * len(forces))(*[pf.desc for pf in self.packed])
        opts = Options(PRECISION[precision], SOLVER[solver], int(device), int(flags))
        h = C.c_void_p()
        rc = self.L.tdp_batch_create(descs, len(forces), C.byref(opts), C.byref(h))
        if rc != TDP_OK:
            raise TdpError(rc, self.L.tdp_last_create_error().decode())
        self.h = h
        self.precision = precision
        self.solver = solver

    def execute(self, positions, box=None, forces=None):
        pos = np.ascontiguousarray(positions, np.float64).reshape(-1)
        f = np.zeros(pos.size) if forces is None else forces
        e = np.zeros(len(self.sizes))
        self._check(self.L.tdp_batch_execute(self.h, _ptr(pos), _ptr(f), _ptr(e)))
        return e, f.reshape(-1, 3)

    def execute_device(self, d_positions_ptr, d_forces_ptr, d_energies_ptr, stream_ptr):
        self._check(self.L.tdp_batch_execute_device(self.h, C.c_void_p(d_positions_ptr), C.c_void_p(d_forces_ptr),
                                                    C.c_void_p(d_energies_ptr), C.c_void_p(stream_ptr)))

    def system_stats(self, system):
        st = SolverStats()
        self._check(self.L.tdp_batch_get_solver_stats(self.h, int(system), C.byref(st)))
        return dict(iterations=st.iterations, epsilon=st.epsilon, converged=st.converged)

    RESPONSE_METHODS = {"fp16": 0, "bf16": 1, "cg": 2}

    def responses(self, positions, method="fp16", tolerance=1e-10, max_steps=50, check=True):
        """Induced dipoles of every system for E0 and unit fields along x, y, z: array [4, N, 3], and the
        statistics.  With check=False a nonzero status is returned as a third value instead of raised."""
        pos = np.ascontiguousarray(positions, np.float64).reshape(-1)
        out = np.zeros(4 * pos.size)
        st = DenseStats()
        rc = self.L.tdp_batch_solve_responses(self.h, _ptr(pos), self.RESPONSE_METHODS[method], float(tolerance),
                                              int(max_steps), _ptr(out), C.byref(st))
        stats = dict(steps=st.steps, max_relative_residual=st.max_relative_residual, factor_failed=st.factor_failed,
                     assemble_ms=st.assemble_ms, factor_ms=st.factor_ms)
        if not check:
            return out.reshape(4, -1, 3), stats, rc
        self._check(rc)
        return out.reshape(4, -1, 3), stats
