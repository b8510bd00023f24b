"""Seeded synthetic systems: rigid 3-site water boxes, ion solutions and water clusters.

Geometry and parameters: O-H 0.09572 nm, H-O-H 104.52 deg, q(O) = -0.834 e, q(H) = +0.417 e,
alpha(O) = 1.44e-3 nm^3, alpha(H) = 0.41e-3 nm^3, Thole a = 0.39, zero molecular dipoles,
NoAxisType.  Each water carries the covalent-map pattern the reference tests use for water
(platforms/reference/tests/TestReferenceTholeDipoleForce.cpp:370-389):
O: 1-2 = [H1, H2]; H1: 1-2 = [O], 1-3 = [H2]; H2: 1-2 = [O], 1-3 = [H1].

Random numbers come from numpy's PCG64 generator seeded per system, so a given
(kind, sizes, seed) always produces the same coordinates on any host.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from .force import TholeDipoleForce

OH = 0.09572
HOH_DEG = 104.52
Q_O, Q_H = -0.834, 0.417
ALPHA_O, ALPHA_H = 1.44e-3, 0.41e-3
THOLE = 0.39
Q_ION = 1.0
ALPHA_NA, ALPHA_CL = 0.24e-3, 3.25e-3
DEBYE_ENM = 0.020819434  # 1 Debye in e*nm

_half = math.radians(HOH_DEG) / 2.0
WATER_LOCAL = np.array([[0.0, 0.0, 0.0],
                        [OH * math.sin(_half), 0.0, OH * math.cos(_half)],
                        [-OH * math.sin(_half), 0.0, OH * math.cos(_half)]])


@dataclass
class SyntheticSystem:
    """Positions and force parameters of one synthetic system."""
    name: str
    force: TholeDipoleForce
    positions: np.ndarray            # (n, 3) nm
    box: Optional[np.ndarray]        # (3, 3) box vectors or None for non-periodic
    masses: np.ndarray
    notes: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return self.positions.shape[0]


def random_rotations(rng: np.random.Generator, k: int) -> np.ndarray:
    """k uniformly distributed rotation matrices (Shoemake's quaternion method)."""
    u1, u2, u3 = rng.random(k), rng.random(k), rng.random(k)
    q0 = np.sqrt(1 - u1) * np.sin(2 * np.pi * u2)
    q1 = np.sqrt(1 - u1) * np.cos(2 * np.pi * u2)
    q2 = np.sqrt(u1) * np.sin(2 * np.pi * u3)
    q3 = np.sqrt(u1) * np.cos(2 * np.pi * u3)
    w, x, y, z = q3, q0, q1, q2
    R = np.empty((k, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - z * w)
    R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w)
    R[:, 2, 1] = 2 * (y * z + x * w)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def _add_water(force: TholeDipoleForce, thole=(THOLE, THOLE, THOLE)) -> None:
    o = force.getNumParticles()
    zero = [0.0, 0.0, 0.0]
    force.addParticle(Q_O, zero, ALPHA_O, thole[0])
    force.addParticle(Q_H, zero, ALPHA_H, thole[1])
    force.addParticle(Q_H, zero, ALPHA_H, thole[2])
    force.setCovalentMap(o, TholeDipoleForce.Covalent12, [o + 1, o + 2])
    force.setCovalentMap(o + 1, TholeDipoleForce.Covalent12, [o])
    force.setCovalentMap(o + 1, TholeDipoleForce.Covalent13, [o + 2])
    force.setCovalentMap(o + 2, TholeDipoleForce.Covalent12, [o])
    force.setCovalentMap(o + 2, TholeDipoleForce.Covalent13, [o + 1])


def _water_sites(centers: np.ndarray, rng: np.random.Generator) -> np.ndarray:
    R = random_rotations(rng, centers.shape[0])
    sites = centers[:, None, :] + np.einsum("kab,sb->ksa", R, WATER_LOCAL)
    return sites.reshape(-1, 3)


def water_box(n_side_points: int, n_waters: int, box_length: float, seed: int, jitter: float = 0.05,
              method: int = TholeDipoleForce.PME, cutoff: float = 0.9, name: str = "water_box",
              n_ion_pairs: int = 0, thole_range=None) -> SyntheticSystem:
    """Waters (and optional Na+/Cl- pairs) on a seeded choice of cubic-lattice points.

    The lattice has ``n_side_points``^3 points with spacing box_length/n_side_points; the
    molecule centres (O sites) are the chosen points plus a uniform jitter in [-jitter, jitter].
    """
    rng = np.random.default_rng(seed)
    spacing = box_length / n_side_points
    g = (np.arange(n_side_points) + 0.5) * spacing
    pts = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    n_mol = n_waters + 2 * n_ion_pairs
    if n_mol > pts.shape[0]:
        raise ValueError("not enough lattice points")
    if n_mol < pts.shape[0]:
        choice = np.sort(rng.choice(pts.shape[0], n_mol, replace=False))
        pts = pts[choice]
    pts = pts + rng.uniform(-jitter, jitter, pts.shape)
    ion_slots = np.sort(rng.choice(n_mol, 2 * n_ion_pairs, replace=False)) if n_ion_pairs else np.zeros(0, int)
    is_ion = np.zeros(n_mol, bool)
    is_ion[ion_slots] = True
    force = TholeDipoleForce()
    force.setNonbondedMethod(method)
    force.setCutoffDistance(cutoff)
    positions: List[np.ndarray] = []
    masses: List[float] = []
    water_centres = pts[~is_ion]
    wsites = _water_sites(water_centres, rng).reshape(-1, 3, 3)
    w = 0
    ion_count = 0
    for m in range(n_mol):
        if is_ion[m]:
            na = ion_count % 2 == 0
            ion_count += 1
            th = THOLE if thole_range is None else rng.uniform(*thole_range)
            force.addParticle(Q_ION if na else -Q_ION, [0.0, 0.0, 0.0], ALPHA_NA if na else ALPHA_CL, th)
            positions.append(pts[m][None, :])
            masses.append(22.99 if na else 35.45)
        else:
            th = (THOLE,) * 3 if thole_range is None else tuple(rng.uniform(*thole_range, size=3))
            _add_water(force, th)
            positions.append(wsites[w])
            masses.extend([15.999, 1.008, 1.008])
            w += 1
    pos = np.concatenate(positions, 0)
    box = np.diag([box_length] * 3).astype(np.float64)
    return SyntheticSystem(name, force, pos, box, np.asarray(masses),
                           notes=dict(n_waters=n_waters, n_ion_pairs=n_ion_pairs, box=box_length, seed=seed))


def water_cluster(n_mol: int, seed: int, spacing: float = 0.31, jitter: float = 0.05,
                  name: str = "water_cluster") -> SyntheticSystem:
    """Non-periodic water cluster: the n_mol cubic-lattice points nearest the centre
    (spacing 0.31 nm = liquid density), jittered, randomly oriented."""
    rng = np.random.default_rng(seed)
    k = int(math.ceil((n_mol * 2.0) ** (1.0 / 3.0))) + 2
    g = (np.arange(k) - (k - 1) / 2.0) * spacing
    pts = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    d = np.linalg.norm(pts, axis=1) + 1e-9 * rng.random(pts.shape[0])
    pts = pts[np.argsort(d, kind="stable")[:n_mol]]
    pts = pts + rng.uniform(-jitter, jitter, pts.shape)
    force = TholeDipoleForce()
    force.setNonbondedMethod(TholeDipoleForce.NoCutoff)
    for _ in range(n_mol):
        _add_water(force)
    pos = _water_sites(pts, rng)
    masses = np.tile([15.999, 1.008, 1.008], n_mol)
    return SyntheticSystem(name, force, pos, None, masses, notes=dict(n_mol=n_mol, seed=seed))


# ------------------------------------------------------------------ the five benchmark workloads
def config0(seed: int = 20261203) -> SyntheticSystem:
    """216 waters (648 sites), 6x6x6 lattice, cubic box 1.86 nm, reference PME semantics."""
    s = water_box(6, 216, 1.86, seed, method=TholeDipoleForce.PME, cutoff=0.9, name="water216")
    s.force.setEwaldErrorTolerance(5e-4)
    s.force.setPMEParameters(0.0, 24, 24, 24)
    s.force.setMutualInducedTargetEpsilon(1e-6 * DEBYE_ENM)
    s.force.setMutualInducedMaxIterations(1000)
    return s


def pme_alpha(force) -> float:
    """The Ewald parameter OpenMM derives from the error tolerance: sqrt(-ln(2 tol)) / cutoff."""
    return math.sqrt(-math.log(2.0 * force.getEwaldErrorTolerance())) / force.getCutoffDistance()


def config1(seed: int = 20261204) -> SyntheticSystem:
    """12,288 waters (36,864 sites) in a 4.95 nm box: seeded 12,288 of the 24^3 lattice points."""
    s = water_box(24, 12288, 4.95, seed, method=TholeDipoleForce.PME, cutoff=1.0, name="water12288")
    # the Ewald parameter of the force's error tolerance with the stated 64^3 grid (alpha 0 would
    # make the grid follow the tolerance too)
    s.force.setPMEParameters(pme_alpha(s.force), 64, 64, 64)
    s.force.setMutualInducedTargetEpsilon(1e-6 * DEBYE_ENM)
    s.force.setMutualInducedMaxIterations(1000)
    return s


def config2(seed: int = 20261205) -> SyntheticSystem:
    """96,000 sites: 31,978 waters + 33 Na+/Cl- pairs in a 7.1 nm box, Thole a ~ U[0.3, 0.45]."""
    s = water_box(32, 31978, 7.1, seed, method=TholeDipoleForce.PME, cutoff=1.0, name="saline96000",
                  n_ion_pairs=33, thole_range=(0.3, 0.45))
    s.force.setPMEParameters(pme_alpha(s.force), 96, 96, 96)
    s.force.setMutualInducedTargetEpsilon(1e-6 * DEBYE_ENM)
    s.force.setMutualInducedMaxIterations(1000)
    return s


def cluster_sizes_config3(count: int = 512, seed: int = 20261206) -> np.ndarray:
    """Molecule counts drawn log-uniformly from [20, 400]."""
    rng = np.random.default_rng(seed)
    return np.rint(np.exp(rng.uniform(math.log(20), math.log(400), count))).astype(int)


def cluster_sizes_config4(count: int = 256, seed: int = 20261207) -> np.ndarray:
    """Molecule counts drawn uniformly from {17..333} (3N in [153, 2997] unknowns)."""
    rng = np.random.default_rng(seed)
    return rng.integers(17, 334, count)
