"""Opt-in elastomer edge bonds (fm_add_edge_bonds): the bonds the
reference's Particles::addElastomer creates on every triangle edge
(femmd2D/src/particles.cpp:335-337, Bonds::createBond with r0 = the current
distance, femmd2D/src/bonds.cpp:14-29), host-side only (no GPU needed).

  * "all": three bonds per triangle in the reference's order, shared edges
    twice (348 for the 116-triangle disk mesh);
  * "dedup": each edge once; for a triangulated disk (one boundary loop,
    no holes) Euler's formula fixes the count: E = V + F - 1 (186 for the
    disk: 71 + 116 - 1), and the boundary edges (those of one triangle)
    number 2E - 3F;
  * bad input is rejected whole (nothing appended);
  * (tests/test_gpu_edge_bonds.py: r0 is the edge's length at creation and
    k the given stiffness, through the bond energy against the oracle).
"""
from pathlib import Path

import numpy as np
import pytest

from femmd_old_amd import FemmdError, System
from femmd_old_amd.io import load_mesh

MESH = Path(__file__).resolve().parent / "golden" / "meshes"


def _disk_system(edge_bonds, k=7.5, mesh="disk"):
    v, t = load_mesh(str(MESH / f"{mesh}Vertices.txt"), str(MESH / f"{mesh}Triangles.txt"))
    s = System(40.0, 40.0, dt=0.005)
    s.add_particles([1.0, 2.0], [1.0, 1.0])  # two solvent particles first
    first = s.add_elastomer(v.ravel(), t.ravel(), 20.0, 20.0, edge_bonds=edge_bonds, bond_k=k,
                            node_mass=1.0)
    return s, v, t, first


@pytest.mark.parametrize("mesh", ["disk", "square"])
def test_edge_bond_counts(mesh):
    s, v, t, first = _disk_system("all", mesh=mesh)
    nv, nt = v.shape[0], t.shape[0]
    assert s.n_bonds() == 3 * nt
    pairs = s.bond_pairs()
    ref = []
    for a, b, c in (t - 1 + first):
        ref += [(a, b), (a, c), (b, c)]
    np.testing.assert_array_equal(pairs, np.array(ref, np.int32))
    d, _, _, _ = _disk_system("dedup", mesh=mesh)
    E = d.n_bonds()
    assert E == nv + nt - 1  # Euler, one boundary loop
    und = {tuple(sorted(p)) for p in pairs.tolist()}
    assert len(und) == E
    assert {tuple(sorted(p)) for p in d.bond_pairs().tolist()} == und
    # edges of one triangle only form the boundary: 2E - 3F of them
    cnt = {}
    for p in pairs.tolist():
        q = tuple(sorted(p))
        cnt[q] = cnt.get(q, 0) + 1
    assert sum(1 for c in cnt.values() if c == 1) == 2 * E - 3 * nt
    if mesh == "disk":
        assert 3 * nt == 348 and E == 186


def test_no_edge_bonds_by_default_and_bad_input():
    s, v, t, first = _disk_system(None)
    assert s.n_bonds() == 0
    with pytest.raises(ValueError):
        _disk_system("some")
    lib, h = s._lib, s._h
    import ctypes as C
    bad = np.array([1, 2, 999], np.int32)  # node out of range
    rc = lib.fm_add_edge_bonds(h, first, 1, bad.ctypes.data_as(C.POINTER(C.c_int32)), 1,
                               C.c_float(1.0), 1)
    assert rc != 0 and s.n_bonds() == 0
    deg = np.array([1, 1, 2], np.int32)  # degenerate triangle
    rc = lib.fm_add_edge_bonds(h, first, 1, deg.ctypes.data_as(C.POINTER(C.c_int32)), 1,
                               C.c_float(1.0), 0)
    assert rc != 0 and s.n_bonds() == 0
