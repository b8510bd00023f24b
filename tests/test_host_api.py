"""CPU tests of the host-side API mirror, the serialization format and the C ABI library."""
import ctypes
import os
import re

import numpy as np
import pytest

from tholedipoleplugin_amd import _lib
from tholedipoleplugin_amd.force import OpenMMException, TholeDipoleForce, validate_force
from tholedipoleplugin_amd.serialization import XmlSerializer

T = TholeDipoleForce
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_defaults():
    """TholeDipoleForce.cpp:43-51."""
    f = T()
    assert f.getNonbondedMethod() == T.NoCutoff
    assert f.getPolarizationType() == T.Mutual
    assert f.getPmeBSplineOrder() == 5
    assert f.getCutoffDistance() == 1.0
    assert f.getEwaldErrorTolerance() == 1e-4
    assert f.getMutualInducedMaxIterations() == 60
    assert f.getMutualInducedTargetEpsilon() == 1e-5
    assert f.getPMEParameters() == (0.0, 0, 0, 0)
    assert f.getExtrapolationCoefficients() == [-0.154, 0.017, 0.658, 0.474]
    assert not f.usesPeriodicBoundaryConditions()
    with pytest.raises(OpenMMException, match="Illegal value for nonbonded method"):
        f.setNonbondedMethod(2)


def test_particles_and_covalent_maps():
    f = T()
    i = f.addParticle(0.5, [1, 2, 3], 0.001, 0.39)
    assert i == 0
    assert f.getParticleParameters(0) == (0.5, [1.0, 2.0, 3.0], 0.001, 0.39, T.NoAxisType, -1, -1, -1)
    f.setParticleParameters(0, -0.5, [3, 2, 1], 0.002, 0.3, T.ZThenX, 1, 2, 3)
    assert f.getParticleParameters(0) == (-0.5, [3.0, 2.0, 1.0], 0.002, 0.3, T.ZThenX, 1, 2, 3)
    f.setCovalentMap(0, T.Covalent13, [4, 5])
    assert f.getCovalentMap(0, T.Covalent13) == [4, 5]
    assert f.getCovalentMaps(0) == [[], [4, 5], [], []]


def test_validation_messages():
    """TholeDipoleForceImpl::initialize checks (TholeDipoleForceImpl.cpp:52-108)."""
    f = T()
    f.addParticle(0.0, [0, 0, 0], 0.001, 0.39, T.ZThenX, 1, -1, -1)
    f.addParticle(0.0, [0, 0, 0], 0.001, 0.39)
    box = np.eye(3) * 2.0
    with pytest.raises(OpenMMException, match="exactly as many particles"):
        validate_force(f, 3, box)
    with pytest.raises(OpenMMException, match="invalid x axis particle: -1"):
        validate_force(f, 2, box)
    f.setParticleParameters(0, 0.0, [0, 0, 0], 0.001, 0.39, 9, 1, 1, -1)
    with pytest.raises(OpenMMException, match="axis type=9 not currently handled"):
        validate_force(f, 2, box)
    f.setParticleParameters(0, 0.0, [0, 0, 0], 0.001, 0.39, T.ZBisect, 1, 1, -1)
    with pytest.raises(OpenMMException, match="invalid y axis particle: -1"):
        validate_force(f, 2, box)
    f.setParticleParameters(0, 0.0, [0, 0, 0], 0.001, 0.39, T.ZOnly, 5, -1, -1)
    with pytest.raises(OpenMMException, match="invalid z axis particle: 5"):
        validate_force(f, 2, box)
    f.setParticleParameters(0, 0.0, [0, 0, 0], 0.001, 0.39, T.NoAxisType, -1, -1, -1)
    f.setNonbondedMethod(T.PME)
    f.setCutoffDistance(1.1)
    with pytest.raises(OpenMMException, match="cannot be greater than half the periodic box size"):
        validate_force(f, 2, box)


def test_serialization_round_trip():
    """serialization/tests/TestSerializeTholeDipoleForce.cpp:45-156 (exact equality of every field)."""
    f = T()
    f.setForceGroup(3)
    f.setNonbondedMethod(T.PME)
    f.setPolarizationType(T.Mutual)
    f.setCutoffDistance(0.9)
    f.setEwaldErrorTolerance(1.0e-05)
    f.setMutualInducedMaxIterations(200)
    f.setMutualInducedTargetEpsilon(1.0e-05)
    f.setPMEParameters(0.5, 32, 32, 32)
    f.setExtrapolationCoefficients([0.0, -0.1, 1.1])
    for i in range(4):
        idx = f.addParticle(i + 1.0, [0.1 * i, 0.2 * i, 0.3 * i], 0.5 + 0.1 * i, 0.8 + 0.05 * i, T.ZThenX,
                            i + 1, i + 2, i + 3)
        for j in range(T.CovalentEnd):
            f.setCovalentMap(idx, j, [i * j, i * j + 1])
    text = XmlSerializer.serialize(f)
    g = XmlSerializer.deserialize(text)
    assert g.getForceGroup() == 3
    for name in ("getNonbondedMethod", "getPolarizationType", "getCutoffDistance", "getEwaldErrorTolerance",
                 "getMutualInducedMaxIterations", "getMutualInducedTargetEpsilon", "getPmeBSplineOrder",
                 "getPMEParameters", "getExtrapolationCoefficients", "getNumParticles"):
        assert getattr(f, name)() == getattr(g, name)()
    for i in range(f.getNumParticles()):
        assert f.getParticleParameters(i) == g.getParticleParameters(i)
        for j in range(T.CovalentEnd):
            assert f.getCovalentMap(i, j) == g.getCovalentMap(i, j)
    assert '<Force type="TholeDipoleForce" version="1"' in text
    bad = text.replace('version="1"', 'version="2"')
    with pytest.raises(OpenMMException, match="Unsupported version number"):
        XmlSerializer.deserialize(bad)


def test_library_exports_every_declared_symbol():
    """libtholedipole_hip.so loads (no device needed) and exports every function the header declares."""
    header = open(os.path.join(ROOT, "include", "tholedipole_hip.h")).read()
    declared = set(re.findall(r"^(?:int|int32_t|void|const char\*)\s+(tdp_\w+)\s*\(", header, re.M))
    assert declared == set(_lib.EXPORTED_SYMBOLS)
    L = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared:
        assert hasattr(L, name), name


def test_library_reports_errors_without_device():
    """tdp_create validates before touching a device and reports through tdp_last_create_error."""
    f = T()
    f.addParticle(0.0, [0, 0, 0], 0.001, 0.39, T.ZThenX, 1, -1, -1)
    f.addParticle(0.0, [0, 0, 0], 0.001, 0.39)
    with pytest.raises(_lib.TdpError, match="invalid x axis particle"):
        _lib.HipContextHandle(f, np.eye(3) * 2.0)
    g = T()
    g.addParticle(0.0, [0, 0, 0], 0.001, 0.39)
    g.setCovalentMap(0, T.Covalent12, [1, 2, 3, 4, 5, 6])
    with pytest.raises(_lib.TdpError, match="scale tables hold 5"):
        _lib.HipContextHandle(g, np.eye(3) * 2.0)
