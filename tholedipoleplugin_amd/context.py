"""Minimal System / Context / State host objects around the HIP kernel.

OpenMM's Context machinery is not part of this package; these classes carry just what the
TholeDipoleForce path needs (particle masses, default box, positions, the force's kernel) so
that programs and tests read like the reference's tests
(platforms/reference/tests/TestReferenceTholeDipoleForce.cpp).

``HipCalcTholeDipoleForceKernel`` implements the ten virtuals of CalcTholeDipoleForceKernel
(openmmapi/include/TholeDipoleKernels.h:46-122) over the C ABI.
"""
from __future__ import annotations

from typing import List, Optional

import numpy as np

from . import _lib
from .force import OpenMMException, TholeDipoleForce, validate_force


class System:
    """Particles (masses), default periodic box and forces.  OpenMM's default box is 2 nm cubic."""

    def __init__(self):
        self._masses: List[float] = []
        self._box = np.diag([2.0, 2.0, 2.0]).astype(np.float64)
        self._forces = []

    def addParticle(self, mass: float) -> int:
        self._masses.append(float(mass))
        return len(self._masses) - 1

    def getNumParticles(self) -> int:
        return len(self._masses)

    def getParticleMass(self, i: int) -> float:
        return self._masses[i]

    def setDefaultPeriodicBoxVectors(self, a, b, c) -> None:
        self._box = np.array([a, b, c], np.float64)

    def getDefaultPeriodicBoxVectors(self):
        return self._box.copy()

    def addForce(self, force) -> int:
        self._forces.append(force)
        return len(self._forces) - 1

    def getForces(self):
        return list(self._forces)


class State:
    def __init__(self, positions, forces, energy):
        self._positions = positions
        self._forces = forces
        self._energy = energy

    def getPositions(self):
        return self._positions.copy()

    def getForces(self):
        if self._forces is None:
            raise OpenMMException("Invoked getForces() on a State which does not contain forces.")
        return self._forces.copy()

    def getPotentialEnergy(self):
        if self._energy is None:
            raise OpenMMException("Invoked getPotentialEnergy() on a State which does not contain energy.")
        return self._energy


class HipCalcTholeDipoleForceKernel:
    """CalcTholeDipoleForceKernel over libtholedipole_hip.so."""

    Name = "CalcTholeDipoleForce"

    def __init__(self, precision="double", solver="jacobi", device=0, flags=0):
        self.precision = precision
        self.solver = solver
        self.device = device
        self.flags = flags
        self.handle: Optional[_lib.HipContextHandle] = None
        self.force = None

    def initialize(self, system: System, force: TholeDipoleForce) -> None:
        self.force = force.snapshot()
        self.default_box = system.getDefaultPeriodicBoxVectors()
        self.masses = np.asarray([system.getParticleMass(i) for i in range(system.getNumParticles())], np.float64)
        try:
            self.handle = _lib.HipContextHandle(self.force, self.default_box, self.precision, self.solver, self.device,
                                                self.flags)
        except _lib.TdpError as e:
            raise OpenMMException(str(e)) from None

    def _call(self, fn, *args):
        try:
            return fn(*args)
        except _lib.TdpError as e:
            raise OpenMMException(str(e)) from None

    def execute(self, context, includeForces=True, includeEnergy=True):
        return self._call(self.handle.execute, context._positions, context._box)

    def copyParametersToContext(self, context, force: TholeDipoleForce) -> None:
        snap = force.snapshot()
        self._call(self.handle.copy_parameters, snap, self.default_box)
        # Only per-particle parameters are updated (ReferenceTholeDipoleKernels.cpp:317-339).
        for i in range(snap.getNumParticles()):
            self.force.setParticleParameters(i, *snap.getParticleParameters(i))

    def getPMEParameters(self):
        return self._call(self.handle.pme_parameters)

    def getInducedDipoles(self, context):
        return self._call(self.handle.dipoles, context._positions, context._box, _lib.DIPOLES_INDUCED)

    def getLabFramePermanentDipoles(self, context):
        return self._call(self.handle.dipoles, context._positions, context._box, _lib.DIPOLES_LAB_PERMANENT)

    def getTotalDipoles(self, context):
        return self._call(self.handle.dipoles, context._positions, context._box, _lib.DIPOLES_TOTAL)

    def getElectrostaticPotential(self, context, inputGrid):
        return self._call(self.handle.potential, context._positions, context._box, inputGrid)

    def getSystemMultipoleMoments(self, context):
        return self._call(self.handle.multipole_moments, context._positions, context._box, self.masses)

    def solverStats(self):
        return self._call(self.handle.solver_stats)


class Context:
    """Holds positions/box and one HIP kernel per TholeDipoleForce of the System.

    ``precision`` is "double" or "mixed", ``solver`` "jacobi" (the reference iteration) or "pcg".
    ``ewald=True`` sums the PME method's electrostatics by Ewald splitting (real space within the
    cutoff, smooth PME) instead of the reference's minimum image over all pairs (TDP_FLAG_EWALD).
    """

    def __init__(self, system: System, precision: str = "double", solver: str = "jacobi", device: int = 0,
                 ewald: bool = False):
        self._system = system
        self._precision = precision
        self._solver = solver
        self._device = device
        self._flags = _lib.FLAG_EWALD if ewald else 0
        self._positions = np.zeros((system.getNumParticles(), 3))
        self._box = system.getDefaultPeriodicBoxVectors()
        self._kernels = {}
        self._init_kernels()

    def _init_kernels(self):
        self._kernels = {}
        for f in self._system.getForces():
            if isinstance(f, TholeDipoleForce):
                validate_force(f, self._system.getNumParticles(), self._system.getDefaultPeriodicBoxVectors())
                k = HipCalcTholeDipoleForceKernel(self._precision, self._solver, self._device, self._flags)
                k.initialize(self._system, f)
                self._kernels[id(f)] = (f, k)

    def _kernel_for(self, force) -> HipCalcTholeDipoleForceKernel:
        try:
            return self._kernels[id(force)][1]
        except KeyError:
            raise OpenMMException("This Force is not present in the Context") from None

    def getSystem(self) -> System:
        return self._system

    def setPositions(self, positions) -> None:
        p = np.asarray(positions, np.float64).reshape(-1, 3)
        if p.shape[0] != self._system.getNumParticles():
            raise OpenMMException("Called setPositions() on a Context with the wrong number of positions")
        self._positions = p.copy()

    def setPeriodicBoxVectors(self, a, b, c) -> None:
        self._box = np.array([a, b, c], np.float64)

    def reinitialize(self, preserveState: bool = False) -> None:
        pos, box = self._positions, self._box
        self._init_kernels()
        if preserveState:
            self._positions, self._box = pos, box

    def getState(self, getForces: bool = False, getEnergy: bool = False, getPositions: bool = False,
                 groups: int = 0xFFFFFFFF) -> State:
        n = self._system.getNumParticles()
        forces = np.zeros((n, 3)) if getForces else None
        energy = 0.0
        if getForces or getEnergy:
            f_acc = np.zeros(3 * n)
            for force, k in self._kernels.values():
                # TholeDipoleForceImpl::calcForcesAndEnergy force-group gate (TholeDipoleForceImpl.cpp:114-118)
                if (groups & (1 << force.getForceGroup())) == 0:
                    continue
                e, f = k.execute(self)
                energy += e
                f_acc += f.reshape(-1)
            if getForces:
                forces = f_acc.reshape(n, 3)
        return State(self._positions.copy(), forces, energy if getEnergy else None)
