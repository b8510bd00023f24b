"""Host-side mirror of the reference's public force API.

``TholeDipoleForce`` follows openmmapi/include/TholeDipoleForce.h:54-428 and
openmmapi/src/TholeDipoleForce.cpp:43-235 (same method names, argument order, defaults and
error behaviour).  ``validate_force`` restates the checks of
TholeDipoleForceImpl::initialize (openmmapi/src/TholeDipoleForceImpl.cpp:52-108).

The context queries (getInducedDipoles, getTotalDipoles, ...) forward to the HIP kernel that
a :class:`tholedipoleplugin_amd.context.Context` created for this force.
"""
from __future__ import annotations

import copy
from typing import List, Sequence

import numpy as np


class OpenMMException(RuntimeError):
    """Error type raised where the reference throws OpenMM::OpenMMException."""


class Force:
    """Minimal stand-in for OpenMM::Force: only the force-group bookkeeping is needed."""

    def __init__(self):
        self._force_group = 0

    def getForceGroup(self) -> int:
        return self._force_group

    def setForceGroup(self, group: int) -> None:
        if group < 0 or group > 31:
            raise OpenMMException("Force group must be between 0 and 31")
        self._force_group = int(group)


class _ParticleInfo:
    """TholeDipoleForce::ParticleInfo (openmmapi/include/TholeDipoleForce.h:399-425)."""

    __slots__ = ("charge", "molecularDipole", "polarizability", "tholeDamping", "axisType",
                 "multipoleAtomZ", "multipoleAtomX", "multipoleAtomY", "covalentInfo")

    def __init__(self, charge, dipole, polarizability, thole, axisType, z, x, y):
        self.charge = float(charge)
        d = list(dipole)
        self.molecularDipole = [float(d[0]), float(d[1]), float(d[2])]
        self.polarizability = float(polarizability)
        self.tholeDamping = float(thole)
        self.axisType = int(axisType)
        self.multipoleAtomZ = int(z)
        self.multipoleAtomX = int(x)
        self.multipoleAtomY = int(y)
        self.covalentInfo: List[List[int]] = [[], [], [], []]


class TholeDipoleForce(Force):
    """Thole/Applequist polarizable point dipoles (charge + dipole, one induced dipole per site)."""

    # NonbondedMethod (TholeDipoleForce.h:58-70)
    NoCutoff = 0
    PME = 1
    # PolarizationType (TholeDipoleForce.h:72-92)
    Direct = 0
    Mutual = 1
    Extrapolated = 2
    # MultipoleAxisTypes (TholeDipoleForce.h:94-102)
    ZThenX = 0
    Bisector = 1
    ZBisect = 2
    ThreeFold = 3
    ZOnly = 4
    NoAxisType = 5
    LastAxisTypeIndex = 6
    # CovalentType (TholeDipoleForce.h:104-110)
    Covalent12 = 0
    Covalent13 = 1
    Covalent14 = 2
    Covalent15 = 3
    CovalentEnd = 4

    def __init__(self):
        # Defaults: openmmapi/src/TholeDipoleForce.cpp:43-51
        super().__init__()
        self._nonbondedMethod = TholeDipoleForce.NoCutoff
        self._polarizationType = TholeDipoleForce.Mutual
        self._pmeBSplineOrder = 5
        self._cutoffDistance = 1.0
        self._ewaldErrorTol = 1e-4
        self._mutualInducedMaxIterations = 60
        self._mutualInducedTargetEpsilon = 1e-5
        self._alpha = 0.0
        self._nx = self._ny = self._nz = 0
        self._extrapolationCoefficients = [-0.154, 0.017, 0.658, 0.474]
        self._particles: List[_ParticleInfo] = []
        self._contexts = []  # contexts created for this force (weak bookkeeping for queries)

    # ---- scalar parameters --------------------------------------------------------------
    def getNumParticles(self) -> int:
        return len(self._particles)

    def getNonbondedMethod(self) -> int:
        return self._nonbondedMethod

    def setNonbondedMethod(self, method: int) -> None:
        # TholeDipoleForce.cpp:57-61
        if method < 0 or method > 1:
            raise OpenMMException("TholeDipoleForce: Illegal value for nonbonded method")
        self._nonbondedMethod = int(method)

    def getPolarizationType(self) -> int:
        return self._polarizationType

    def setPolarizationType(self, ptype: int) -> None:
        self._polarizationType = int(ptype)

    def setExtrapolationCoefficients(self, coefficients: Sequence[float]) -> None:
        self._extrapolationCoefficients = [float(c) for c in coefficients]

    def getExtrapolationCoefficients(self) -> List[float]:
        return list(self._extrapolationCoefficients)

    def getCutoffDistance(self) -> float:
        return self._cutoffDistance

    def setCutoffDistance(self, distance: float) -> None:
        self._cutoffDistance = float(distance)

    def getPMEParameters(self):
        """Returns (alpha, nx, ny, nz) (TholeDipoleForce.cpp:87-92)."""
        return self._alpha, self._nx, self._ny, self._nz

    def setPMEParameters(self, alpha: float, nx: int, ny: int, nz: int) -> None:
        self._alpha, self._nx, self._ny, self._nz = float(alpha), int(nx), int(ny), int(nz)

    def getPmeBSplineOrder(self) -> int:
        return self._pmeBSplineOrder

    def getPMEParametersInContext(self, context):
        return context._kernel_for(self).getPMEParameters()

    def getMutualInducedMaxIterations(self) -> int:
        return self._mutualInducedMaxIterations

    def setMutualInducedMaxIterations(self, n: int) -> None:
        self._mutualInducedMaxIterations = int(n)

    def getMutualInducedTargetEpsilon(self) -> float:
        return self._mutualInducedTargetEpsilon

    def setMutualInducedTargetEpsilon(self, eps: float) -> None:
        self._mutualInducedTargetEpsilon = float(eps)

    def getEwaldErrorTolerance(self) -> float:
        return self._ewaldErrorTol

    def setEwaldErrorTolerance(self, tol: float) -> None:
        self._ewaldErrorTol = float(tol)

    def usesPeriodicBoundaryConditions(self) -> bool:
        return self._nonbondedMethod == TholeDipoleForce.PME

    # ---- particles ---------------------------------------------------------------------
    def addParticle(self, charge, molecularDipole, polarizability, tholeDamping,
                    axisType=NoAxisType, multipoleAtomZ=-1, multipoleAtomX=-1, multipoleAtomY=-1) -> int:
        self._particles.append(_ParticleInfo(charge, molecularDipole, polarizability, tholeDamping,
                                             axisType, multipoleAtomZ, multipoleAtomX, multipoleAtomY))
        return len(self._particles) - 1

    def getParticleParameters(self, index: int):
        """Returns (charge, dipole[3], polarizability, thole, axisType, atomZ, atomX, atomY)."""
        p = self._particles[index]
        return (p.charge, list(p.molecularDipole), p.polarizability, p.tholeDamping, p.axisType,
                p.multipoleAtomZ, p.multipoleAtomX, p.multipoleAtomY)

    def setParticleParameters(self, index, charge, molecularDipole, polarizability, tholeDamping,
                              axisType, multipoleAtomZ, multipoleAtomX, multipoleAtomY) -> None:
        p = self._particles[index]
        p.charge = float(charge)
        p.polarizability = float(polarizability)
        p.tholeDamping = float(tholeDamping)
        p.molecularDipole = [float(molecularDipole[0]), float(molecularDipole[1]), float(molecularDipole[2])]
        p.axisType = int(axisType)
        p.multipoleAtomZ = int(multipoleAtomZ)
        p.multipoleAtomX = int(multipoleAtomX)
        p.multipoleAtomY = int(multipoleAtomY)

    def setCovalentMap(self, index: int, typeId: int, covalentAtoms: Sequence[int]) -> None:
        self._particles[index].covalentInfo[int(typeId)] = [int(a) for a in covalentAtoms]

    def getCovalentMap(self, index: int, typeId: int) -> List[int]:
        return list(self._particles[index].covalentInfo[int(typeId)])

    def getCovalentMaps(self, index: int) -> List[List[int]]:
        return [list(l) for l in self._particles[index].covalentInfo]

    # ---- context queries (TholeDipoleForce.cpp:206-235) ------------------------------
    def getInducedDipoles(self, context) -> np.ndarray:
        return context._kernel_for(self).getInducedDipoles(context)

    def getLabFramePermanentDipoles(self, context) -> np.ndarray:
        return context._kernel_for(self).getLabFramePermanentDipoles(context)

    def getTotalDipoles(self, context) -> np.ndarray:
        return context._kernel_for(self).getTotalDipoles(context)

    def getElectrostaticPotential(self, inputGrid, context) -> np.ndarray:
        return context._kernel_for(self).getElectrostaticPotential(context, inputGrid)

    def getSystemMultipoleMoments(self, context) -> np.ndarray:
        return context._kernel_for(self).getSystemMultipoleMoments(context)

    def updateParametersInContext(self, context) -> None:
        # TholeDipoleForceImpl.cpp:188-191: copy parameters, then systemChanged()
        context._kernel_for(self).copyParametersToContext(context, self)

    # ---- packing helpers used by the kernels ---------------------------------------------
    def snapshot(self) -> "TholeDipoleForce":
        """Deep copy of the parameters (the kernel copies at initialize, ReferenceTholeDipoleKernels.cpp:83-125)."""
        c = TholeDipoleForce.__new__(TholeDipoleForce)
        c.__dict__.update({k: copy.deepcopy(v) for k, v in self.__dict__.items() if k != "_contexts"})
        c._contexts = []
        return c

    def packed_arrays(self):
        """Structure-of-arrays view of the per-particle parameters and covalent maps (CSR).

        Returns a dict with float64 charge/dipole/polarity/thole, int32 axis/atom indices and
        ``cov_offsets`` ([4n+1]) / ``cov_atoms``: list t of particle i occupies
        cov_atoms[cov_offsets[4i+t]:cov_offsets[4i+t+1]].
        """
        n = len(self._particles)
        q = np.empty(n, np.float64)
        d = np.empty((n, 3), np.float64)
        al = np.empty(n, np.float64)
        th = np.empty(n, np.float64)
        ax = np.empty(n, np.int32)
        az = np.empty(n, np.int32)
        axx = np.empty(n, np.int32)
        ay = np.empty(n, np.int32)
        offs = np.zeros(4 * n + 1, np.int32)
        atoms = []
        k = 0
        for i, p in enumerate(self._particles):
            q[i] = p.charge
            d[i] = p.molecularDipole
            al[i] = p.polarizability
            th[i] = p.tholeDamping
            ax[i] = p.axisType
            az[i] = p.multipoleAtomZ
            axx[i] = p.multipoleAtomX
            ay[i] = p.multipoleAtomY
            for t in range(4):
                offs[4 * i + t] = k
                atoms.extend(p.covalentInfo[t])
                k += len(p.covalentInfo[t])
        offs[4 * n] = k
        return dict(charge=q, dipole=d.reshape(-1), polarity=al, thole=th, axis_type=ax, atom_z=az,
                    atom_x=axx, atom_y=ay, cov_offsets=offs,
                    cov_atoms=np.asarray(atoms if atoms else [0], np.int32))


CUTOFF_TOO_LARGE = "The cutoff distance cannot be greater than half the periodic box size."


def validate_force(force: TholeDipoleForce, num_system_particles: int, default_box) -> None:
    """Checks of TholeDipoleForceImpl::initialize (openmmapi/src/TholeDipoleForceImpl.cpp:52-108)."""
    n = force.getNumParticles()
    if n != num_system_particles:
        raise OpenMMException("TholeDipoleForce must have exactly as many particles as the System it belongs to.")
    if force.getNonbondedMethod() == TholeDipoleForce.PME:
        box = np.asarray(default_box, np.float64).reshape(3, 3)
        c = force.getCutoffDistance()
        if c > 0.5 * box[0, 0] or c > 0.5 * box[1, 1] or c > 0.5 * box[2, 2]:
            raise OpenMMException("TholeDipoleForce: " + CUTOFF_TOO_LARGE)
    T = TholeDipoleForce
    for ii in range(n):
        (_, _, _, _, axisType, z, x, y) = force.getParticleParameters(ii)
        if axisType not in (T.ZThenX, T.Bisector, T.ZBisect, T.ThreeFold, T.ZOnly, T.NoAxisType):
            raise OpenMMException(
                "TholeDipoleForce: axis type=%d not currently handled - only axisTypes[ %d, %d, %d, %d, %d, %d] "
                "(ZThenX, Bisector, Z-Bisect, ThreeFold, ZOnly, NoAxisType) currently handled."
                % (axisType, T.ZThenX, T.Bisector, T.ZBisect, T.ThreeFold, T.ZOnly, T.NoAxisType))
        if axisType != T.NoAxisType and (z < 0 or z >= n):
            raise OpenMMException("TholeDipoleForce: invalid z axis particle: %d" % z)
        if axisType != T.NoAxisType and axisType != T.ZOnly and (x < 0 or x >= n):
            raise OpenMMException("TholeDipoleForce: invalid x axis particle: %d" % x)
        if axisType in (T.ZBisect, T.ThreeFold) and (y < 0 or y >= n):
            raise OpenMMException("TholeDipoleForce: invalid y axis particle: %d" % y)
