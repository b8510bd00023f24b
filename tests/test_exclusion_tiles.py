"""Exclusion tiles against the oracle's scale lookup, pair by pair (CPU only).

The pair kernels read the scale maps (setupScaleMaps/getScaleFactor,
ReferenceTholeDipoleForce.cpp:609-644) only through the 64x64 tile bit masks that the library's
host code builds (tdp_tables.cpp, compiled on its own into libtdp_tables.so).  Here the masks are
built for a shuffled internal order with padding slots and decoded back for EVERY pair i < j of
several topologies; each decoded (mScale, iScale) must equal the oracle's tdpo_pair_scales exactly
(values in {0, 0.5, 1}), both orientations of a tile must agree, and every particle's self pair must
be masked (mScale == 0 and iScale == 0).
"""
import ctypes as C
import os

import numpy as np
import pytest

import oracle
import ref_systems
from tholedipoleplugin_amd import systems
from tholedipoleplugin_amd.force import TholeDipoleForce

T = TholeDipoleForce
_LIB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tholedipoleplugin_amd",
                    "libtdp_tables.so")


def _tables():
    if not os.path.exists(_LIB):
        pytest.fail("libtdp_tables.so is not built (make -C tholedipoleplugin_amd/csrc)")
    L = C.CDLL(_LIB)
    ip = C.POINTER(C.c_int32)
    L.tdp_tables_exclusion_tiles.argtypes = [C.c_int, ip, ip, C.c_int, ip, C.c_int, ip, ip,
                                             C.POINTER(C.c_ulonglong)]
    L.tdp_tables_exclusion_tiles.restype = C.c_int
    return L


def _build(force, inv, nblk):
    a = force.packed_arrays()
    off = np.ascontiguousarray(a["cov_offsets"], np.int32)
    atoms = np.ascontiguousarray(a["cov_atoms"], np.int32)
    if atoms.size == 0:
        atoms = np.zeros(1, np.int32)
    inv = np.ascontiguousarray(inv, np.int32)
    n = force.getNumParticles()
    cap = nblk + 2 * atoms.size + 8
    start = np.zeros(nblk + 1, np.int32)
    J = np.zeros(cap, np.int32)
    mask = np.zeros(cap * 192, np.uint64)
    ip = C.POINTER(C.c_int32)
    nt = _tables().tdp_tables_exclusion_tiles(n, off.ctypes.data_as(ip), atoms.ctypes.data_as(ip), nblk,
                                              inv.ctypes.data_as(ip), cap, start.ctypes.data_as(ip),
                                              J.ctypes.data_as(ip), mask.ctypes.data_as(C.POINTER(C.c_ulonglong)))
    assert nt >= 0
    return start, J[:nt], mask[:nt * 192].reshape(nt, 3, 64)


def _decode(force, inv, nblk):
    """Dense (m, i) scales in the caller's order from the tiles; self pairs checked separately."""
    n = force.getNumParticles()
    start, J, mask = _build(force, inv, nblk)
    perm = np.full(64 * nblk, -1)
    perm[inv] = np.arange(n)
    M = np.ones((n, n))
    I_ = np.ones((n, n))
    seen = {}
    cols = np.arange(64, dtype=np.uint64)
    for I in range(nblk):
        for k in range(start[I], start[I + 1]):
            Jt = int(J[k])
            bits = ((mask[k][:, :, None] >> cols[None, None, :]) & np.uint64(1)).astype(bool)   # [3][row][col]
            seen[(I, Jt)] = bits
            for row in range(64):
                a = perm[64 * I + row]
                for kind in range(3):
                    for col in np.nonzero(bits[kind, row])[0]:
                        b = perm[64 * Jt + col]
                        assert a >= 0 and b >= 0, "bit set on a padding slot"
                        if a == b:
                            continue
                        if kind == 0:
                            M[a, b] = 0.0
                        elif kind == 1:
                            assert M[a, b] != 0.0
                            M[a, b] = 0.5
                        else:
                            I_[a, b] = 0.0
    # both orientations of every tile pair carry the same pairs
    for (I, Jt), bits in seen.items():
        assert (Jt, I) in seen
        assert np.array_equal(bits, np.transpose(seen[(Jt, I)], (0, 2, 1)))
    # self pairs: the mScale == 0 and iScale == 0 masks (never evaluated)
    for i in range(n):
        s = inv[i]
        bits = seen[(s // 64, s // 64)]
        assert bits[0, s % 64, s % 64] and bits[2, s % 64, s % 64] and not bits[1, s % 64, s % 64]
    return M, I_


def _oracle_scales(force):
    n = force.getNumParticles()
    os_ = oracle.OracleSystem(force)
    L = oracle.lib()
    M = np.ones((n, n))
    I_ = np.ones((n, n))
    m = C.c_double()
    s = C.c_double()
    for i in range(n):
        for j in range(i + 1, n):
            assert L.tdpo_pair_scales(C.byref(os_.s), i, j, C.byref(m), C.byref(s)) == 0
            M[i, j] = M[j, i] = m.value
            I_[i, j] = I_[j, i] = s.value
    return M, I_


def _random_topology(n, seed):
    """Particles with random 1-2 and 1-3 lists of up to five atoms, duplicates included (the last
    occurrence wins) and atoms on both sides of the particle's index."""
    rng = np.random.default_rng(seed)
    f = T()
    for _ in range(n):
        f.addParticle(0.1, [0, 0, 0], 1e-3, 0.39)
    for i in range(n):
        for t in range(2):
            k = int(rng.integers(0, 6))
            lst = [int(x) for x in rng.integers(0, n, size=k)]
            if k >= 2 and rng.random() < 0.5:
                lst[-1] = lst[0]   # duplicate
            f.setCovalentMap(i, t, lst)
    return f


def _topologies():
    out = {}
    for name in ("ammonia", "water4", "triclinic", "zbisect", "zonly"):
        out[name] = getattr(ref_systems, name)().force
    out["water216"] = systems.config0().force
    out["random150"] = _random_topology(150, 3)
    # a batch: several clusters concatenated with offset covalent maps (as tdp_batch_create does)
    grouped = T()
    for k, m in enumerate((7, 20, 3)):
        c = systems.water_cluster(m, seed=900 + k).force
        base = grouped.getNumParticles()
        for p in range(c.getNumParticles()):
            grouped.addParticle(*c.getParticleParameters(p))
        for p in range(c.getNumParticles()):
            for t in range(4):
                grouped.setCovalentMap(base + p, t, [base + a for a in c.getCovalentMap(p, t)])
    out["grouped"] = grouped
    return out


@pytest.mark.parametrize("name", ["ammonia", "water4", "triclinic", "zbisect", "zonly", "water216",
                                  "random150", "grouped"])
def test_tiles_equal_oracle_scales_every_pair(name):
    force = _topologies()[name]
    n = force.getNumParticles()
    nblk = max(1, (n + 63) // 64) + 1                # one extra tile: padding slots everywhere
    rng = np.random.default_rng(n)
    inv = rng.permutation(64 * nblk)[:n]             # shuffled internal order with holes
    M, I_ = _decode(force, inv, nblk)
    Mo, Io = _oracle_scales(force)
    iu = np.triu_indices(n, 1)
    # the kernels read the (lower caller index, higher) lookup: both triangles must hold it
    assert np.array_equal(M[iu], Mo[iu]), name
    assert np.array_equal(M.T[iu], Mo[iu]), name
    assert np.array_equal(I_[iu], Io[iu]), name
    assert np.array_equal(I_.T[iu], Io[iu]), name
    assert set(np.unique(Mo[iu])) <= {0.0, 0.5, 1.0}
