"""On-disk formats on the CPU: the reference's mesh files read back as the
reference reads them, and the writers' layouts on a stand-in system.

  * the disk and square meshes of femmd2D/meshes (fixtures, unchanged):
    every token parsed, indices 1-based in range, all triangles CCW (the
    Neo-Hookean energy uses the signed area), areas summing to the polygon
    area of the boundary-free mesh (the disk's triangles tile a convex
    region: positive and no gaps, checked against the convex hull area);
  * XYZ / VTF / cluster VTF / cluster XYZ / VTK text as
    femmd2D/src/md3dsystem.cpp:232-377 prints it (C %f), frames appended
    with s != 0.
"""
from pathlib import Path

import numpy as np
import pytest

from femmd_old_amd import configs, io

MESH = Path(__file__).resolve().parent / "golden" / "meshes"


@pytest.mark.parametrize("name,nv,nt", [("disk", 71, 116), ("square", 67, 100)])
def test_load_reference_mesh(name, nv, nt):
    v, t = io.load_mesh(str(MESH / f"{name}Vertices.txt"), str(MESH / f"{name}Triangles.txt"))
    assert v.shape == (nv, 2) and v.dtype == np.float32
    assert t.shape == (nt, 3) and t.dtype == np.int32
    assert t.min() == 1 and t.max() == nv
    a = configs.signed_areas(v.astype(np.float64), t - 1)
    assert np.all(a > 0)
    # no overlaps or holes: the triangles tile the convex hull of the nodes
    from scipy.spatial import ConvexHull
    assert a.sum() == pytest.approx(ConvexHull(v.astype(np.float64)).volume, rel=1e-9)
    # every node is used
    assert np.unique(t).size == nv


def test_read_file_kinds(tmp_path):
    p = tmp_path / "f.txt"
    p.write_text("1 2\n3\t4\n")
    assert io.read_file(str(p), io.FILECONTENT.INT).tolist() == [1, 2, 3, 4]
    f = io.read_file(str(p), io.FILECONTENT.FLOAT)
    assert f.dtype == np.float32 and f.tolist() == [1.0, 2.0, 3.0, 4.0]
    assert io.read_file(str(p), io.FILECONTENT.TEXT) == "1 2\n3\t4\n"
    bad = tmp_path / "t.txt"
    bad.write_text("1 2 5\n")
    with pytest.raises(ValueError):
        io.load_mesh(str(p), str(bad))


class _Stand:
    def __init__(self):
        self.x = np.array([1.5, 2.25], np.float32)
        self.y = np.array([0.125, 3.0], np.float32)
        self.vx = np.array([0.5, -1.0], np.float32)
        self.vy = np.array([0.0, 2.0], np.float32)

    def get_state(self):
        return self.x, self.y, self.vx, self.vy

    def get_box(self):
        return {"x": 10.0, "y": 12.0}

    def energies(self):
        return {"steps": 42}

    def particle_types(self):
        return np.array([0, 1], np.int32)

    def particle_sigmas(self):
        return np.array([1.0, 2.0], np.float32)

    def bond_pairs(self):
        return np.array([[0, 1]], np.int32)

    def triangles(self):
        return np.zeros((0, 3), np.int32)

    def cluster_count(self):
        return 1

    def cluster_props(self, unfolded=False):
        r = {"mass": np.array([2.0]), "cm": np.array([[1.875, 1.5625]]),
             "vcm": np.array([[-0.25, 1.0]]), "centroid": np.array([[1.875, 1.5625]])}
        if unfolded:
            r["ux"], r["uy"] = self.x, self.y
        return r


def test_writers_layout(tmp_path):
    s = _Stand()
    p = str(tmp_path / "a.xyz")
    io.save_xyz(s, p, 0)
    io.save_xyz(s, p, 1)
    assert Path(p).read_text() == ("2\n\nA\t1.500000\t0.125000\t0.000000\nA\t2.250000\t3.000000\t0.000000\n" * 2)
    p = str(tmp_path / "a.vtf")
    io.save_vtf(s, p, 0)
    io.save_vtf(s, p, 1)
    assert Path(p).read_text() == (
        "atom 0 type A radius 0.500000\natom 1 type B radius 1.000000\nbond 0:1\n"
        "pbc 10.000000 12.000000 1.000000\n"
        "timestep\n1.500000 0.125000 0.000000\n2.250000 3.000000 0.000000\n"
        "timestep\n1.500000 0.125000 0.000000\n2.250000 3.000000 0.000000\n")
    p = str(tmp_path / "c.vtf")
    io.save_cluster_vtf(s, p, 0)
    assert Path(p).read_text() == ("atom 0 type A radius 1.000000\npbc 10.000000 12.000000 1.000000\n"
                                   "timestep\n1.875000 1.562500 0.000000\n")
    p = str(tmp_path / "c.xyz")
    io.save_cluster_xyz(s, p, 0)
    assert Path(p).read_text() == (
        "ITEM: TIMESTEP\n42\nITEM: NUMBER OF ATOMS\n1\n"
        "ITEM: BOX BOUNDS pp pp pp\n0.000000 10.000000\n0.000000 12.000000\n0.000000 0.000000\n"
        "ITEM: ATOMS id type x y z vx vy vz\n"
        "0 0 1.875000 1.562500 0.000000 -0.250000 1.000000 0.000000\n")
    name = io.save_vtk(s, str(tmp_path / "m"))
    assert name.endswith("m000000042.vtk")
    txt = Path(name).read_text()
    assert txt.startswith("# vtk DataFile Version 3.1\nReally cool data\nASCII\n"
                          "DATASET UNSTRUCTURED_GRID\n\nPOINTS 2 FLOAT\n1.500000 0.125000 0.000000\n")
    assert "SCALARS VelocitySquare FLOAT\nLOOKUP_TABLE default\n0.250000\n5.000000\n" in txt
    assert txt.endswith("VECTORS Velocity FLOAT\n0.500000 0.000000 0.000000\n-1.000000 2.000000 0.000000\n")
