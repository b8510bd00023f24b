"""XML serialization of TholeDipoleForce, format version 1.

Mirrors TholeDipoleForceProxy (serialization/src/TholeDipoleForceProxy.cpp:44-183): the same
node names, property names, child layout and version check, written in the layout of OpenMM's
XmlSerializer (a root element carrying a ``type`` attribute).  Doubles are written with the
shortest representation that round-trips exactly.
"""
from __future__ import annotations

import xml.etree.ElementTree as ET

from .force import OpenMMException, TholeDipoleForce


def _d(x: float) -> str:
    return repr(float(x))


def serialize_node(force: TholeDipoleForce) -> ET.Element:
    """TholeDipoleForceProxy::serialize (TholeDipoleForceProxy.cpp:44-112)."""
    node = ET.Element("Force")
    node.set("type", "TholeDipoleForce")
    node.set("version", "1")
    node.set("forceGroup", str(force.getForceGroup()))
    node.set("nonbondedMethod", str(force.getNonbondedMethod()))
    node.set("polarizationType", str(force.getPolarizationType()))
    node.set("cutoffDistance", _d(force.getCutoffDistance()))
    node.set("ewaldErrorTolerance", _d(force.getEwaldErrorTolerance()))
    node.set("mutualInducedMaxIterations", str(force.getMutualInducedMaxIterations()))
    node.set("mutualInducedTargetEpsilon", _d(force.getMutualInducedTargetEpsilon()))
    node.set("pmeBSplineOrder", str(force.getPmeBSplineOrder()))
    alpha, nx, ny, nz = force.getPMEParameters()
    node.set("alpha", _d(alpha))
    node.set("nx", str(nx))
    node.set("ny", str(ny))
    node.set("nz", str(nz))
    ex = ET.SubElement(node, "ExtrapolationCoefficients")
    for c in force.getExtrapolationCoefficients():
        ET.SubElement(ex, "Coefficient").set("c", _d(c))
    parts = ET.SubElement(node, "Particles")
    for i in range(force.getNumParticles()):
        q, d, pol, th, ax, z, x, y = force.getParticleParameters(i)
        p = ET.SubElement(parts, "Particle")
        p.set("charge", _d(q))
        p.set("polarizability", _d(pol))
        p.set("tholeDamping", _d(th))
        p.set("axisType", str(ax))
        p.set("multipoleAtomZ", str(z))
        p.set("multipoleAtomX", str(x))
        p.set("multipoleAtomY", str(y))
        dip = ET.SubElement(p, "Dipole")
        dip.set("x", _d(d[0]))
        dip.set("y", _d(d[1]))
        dip.set("z", _d(d[2]))
        cms = ET.SubElement(p, "CovalentMaps")
        for t in range(TholeDipoleForce.CovalentEnd):
            atoms = force.getCovalentMap(i, t)
            if atoms:
                cm = ET.SubElement(cms, "CovalentMap")
                cm.set("type", str(t))
                for a in atoms:
                    ET.SubElement(cm, "Atom").set("index", str(a))
    return node


def _req(node: ET.Element, name: str) -> str:
    v = node.get(name)
    if v is None:
        raise OpenMMException("Unknown property '%s' in node '%s'" % (name, node.tag))
    return v


def _child(node: ET.Element, name: str) -> ET.Element:
    c = node.find(name)
    if c is None:
        raise OpenMMException("Unknown child '%s' in node '%s'" % (name, node.tag))
    return c


def deserialize_node(node: ET.Element) -> TholeDipoleForce:
    """TholeDipoleForceProxy::deserialize (TholeDipoleForceProxy.cpp:114-183)."""
    if int(_req(node, "version")) != 1:
        raise OpenMMException("Unsupported version number")
    f = TholeDipoleForce()
    f.setForceGroup(int(_req(node, "forceGroup")))
    f.setNonbondedMethod(int(_req(node, "nonbondedMethod")))
    f.setPolarizationType(int(_req(node, "polarizationType")))
    f.setCutoffDistance(float(_req(node, "cutoffDistance")))
    f.setEwaldErrorTolerance(float(_req(node, "ewaldErrorTolerance")))
    f.setMutualInducedMaxIterations(int(_req(node, "mutualInducedMaxIterations")))
    f.setMutualInducedTargetEpsilon(float(_req(node, "mutualInducedTargetEpsilon")))
    f.setPMEParameters(float(_req(node, "alpha")), int(_req(node, "nx")), int(_req(node, "ny")),
                       int(_req(node, "nz")))
    f.setExtrapolationCoefficients([float(_req(c, "c")) for c in _child(node, "ExtrapolationCoefficients")])
    for p in _child(node, "Particles"):
        dip = _child(p, "Dipole")
        idx = f.addParticle(float(_req(p, "charge")),
                            [float(_req(dip, "x")), float(_req(dip, "y")), float(_req(dip, "z"))],
                            float(_req(p, "polarizability")), float(_req(p, "tholeDamping")),
                            int(_req(p, "axisType")), int(_req(p, "multipoleAtomZ")),
                            int(_req(p, "multipoleAtomX")), int(_req(p, "multipoleAtomY")))
        for cm in _child(p, "CovalentMaps"):
            f.setCovalentMap(idx, int(_req(cm, "type")), [int(_req(a, "index")) for a in cm])
    return f


class XmlSerializer:
    """serialize/deserialize with the call shape of OpenMM::XmlSerializer for this one type."""

    @staticmethod
    def serialize(force: TholeDipoleForce) -> str:
        node = serialize_node(force)
        ET.indent(node, space="\t")
        return '<?xml version="1.0" ?>\n' + ET.tostring(node, encoding="unicode") + "\n"

    @staticmethod
    def deserialize(text: str) -> TholeDipoleForce:
        node = ET.fromstring(text)
        t = node.get("type")
        if t is not None and t != "TholeDipoleForce":
            raise OpenMMException("Unknown object type '%s'" % t)
        return deserialize_node(node)
