"""On-disk formats of femmd2D: the mesh files it reads and the trajectory
files it writes, over a ``System``.

Input (femmd2D/src/interface.cpp:135-195, ``read_file``): whitespace
separated numbers, read as FP32 (vertex files: x y per line) or int32
(triangle files: three 1-based vertex indices per line,
femmd2D/src/particles.cpp:327-329).  ``load_mesh`` pairs the two files of a
mesh in femmd2D/meshes/ (``<name>Vertices.txt``, ``<name>Triangles.txt``).

Output (femmd2D/src/md3dsystem.cpp:232-377), byte-compatible layouts:
  * ``save_xyz``          XYZ frames, "A\\tx\\ty\\t0" per particle (:232-246)
  * ``save_cluster_xyz``  LAMMPS-style dump of the cluster centres (:248-273)
  * ``save_vtk``          legacy VTK unstructured grid of the unfolded
                          positions, triangles and velocities (:275-325)
  * ``save_vtf``          VMD VTF: atoms (radius sigma / 2), bonds, pbc, then
                          one "timestep" frame per call (:330-352)
  * ``save_cluster_vtf``  VTF frames of the cluster centres (:354-377)
``s`` = 0 starts a file, anything else appends a frame, as in the reference.
Numbers are printed with C's %f (six decimals).

Where the reference writes something other than what its names say, the
writers write what the names say: ``save_cluster_xyz`` lists the cluster
centres of mass and their velocities (the reference prints the first NC
particles' positions), and ``save_vtk`` takes the cell indices of each
element from the global node ids.
"""
from __future__ import annotations

import os
from typing import Tuple

import numpy as np

__all__ = ["FILECONTENT", "read_file", "load_mesh", "save_xyz", "save_cluster_xyz",
           "save_vtk", "save_vtf", "save_cluster_vtf"]


class FILECONTENT:
    """read_file kinds (femmd2D/src/interface.cpp:145-195)."""
    INT = 0
    FLOAT = 1
    TEXT = 2


def read_file(path: str, kind: int):
    """femmd2D/src/interface.cpp:135-195: every whitespace-separated token as
    int32 (INT) or FP32 (FLOAT), or the text (TEXT)."""
    with open(path, "r") as f:
        text = f.read()
    if kind == FILECONTENT.TEXT:
        return text
    toks = text.split()
    if kind == FILECONTENT.INT:
        return np.array([int(t) for t in toks], dtype=np.int32)
    if kind == FILECONTENT.FLOAT:
        return np.array([float(t) for t in toks], dtype=np.float64).astype(np.float32)
    raise ValueError(f"unknown file content kind {kind}")


def load_mesh(vertices: str, triangles: str) -> Tuple[np.ndarray, np.ndarray]:
    """(verts (nv, 2) FP32, tris (nt, 3) int32 1-based) of a reference mesh,
    checked: every index in 1..nv.  Pass them to ``System.add_elastomer``
    with its default ``index_base=1``."""
    v = read_file(vertices, FILECONTENT.FLOAT)
    t = read_file(triangles, FILECONTENT.INT)
    if v.size % 2 or t.size % 3:
        raise ValueError("a mesh needs pairs of coordinates and triples of indices")
    v = v.reshape(-1, 2)
    t = t.reshape(-1, 3)
    if t.size and (t.min() < 1 or t.max() > v.shape[0]):
        raise ValueError("triangle index outside 1..nv")
    return v, t


def _open(path: str, s: int):
    return open(path, "w" if s == 0 else "a")


def _f(v) -> str:
    return "%f" % float(v)


def save_xyz(sys, path: str, s: int = 0):
    """MDSystem::saveXYZ (femmd2D/src/md3dsystem.cpp:232-246)."""
    x, y, _, _ = sys.get_state()
    with _open(path, s) as fp:
        fp.write("%d\n\n" % len(x))
        fp.writelines("A\t%s\t%s\t%s\n" % (_f(a), _f(b), _f(0.0)) for a, b in zip(x, y))


def save_cluster_xyz(sys, path: str, s: int = 0):
    """MDSystem::saveClusterXYZ (femmd2D/src/md3dsystem.cpp:248-273): one
    LAMMPS-style frame of the cluster centres of mass and velocities."""
    box = sys.get_box()
    steps = sys.energies()["steps"]
    c = sys.cluster_props() if sys.cluster_count() else None
    nc = 0 if c is None else len(c["mass"])
    typ = sys.particle_types()
    with _open(path, s) as fp:
        fp.write("ITEM: TIMESTEP\n%d\n" % steps)
        fp.write("ITEM: NUMBER OF ATOMS\n%d\n" % nc)
        fp.write("ITEM: BOX BOUNDS pp pp pp\n%s %s\n%s %s\n%s %s\n"
                 % (_f(0), _f(box["x"]), _f(0), _f(box["y"]), _f(0), _f(0)))
        fp.write("ITEM: ATOMS id type x y z vx vy vz\n")
        for i in range(nc):
            t = int(typ[i]) if i < len(typ) else 0
            fp.write("%d %d %s %s %s %s %s %s\n" % (
                i, t, _f(c["cm"][i, 0]), _f(c["cm"][i, 1]), _f(0.0),
                _f(c["vcm"][i, 0]), _f(c["vcm"][i, 1]), _f(0.0)))


def save_vtk(sys, prefix: str) -> str:
    """MDSystem::saveVtk (femmd2D/src/md3dsystem.cpp:275-325): writes
    ``<prefix><steps, 9 digits>.vtk`` (nothing without clusters, as the
    reference) and returns its name."""
    if sys.cluster_count() == 0:
        return ""
    steps = sys.energies()["steps"]
    name = "%s%.9d.vtk" % (prefix, steps)
    c = sys.cluster_props(unfolded=True)
    ux, uy = c["ux"], c["uy"]
    _, _, vx, vy = sys.get_state()
    tri = sys.triangles()
    n = len(ux)
    vv = vx.astype(np.float32) * vx.astype(np.float32) + vy.astype(np.float32) * vy.astype(np.float32)
    with open(name, "w") as fp:
        fp.write("# vtk DataFile Version 3.1\n")
        fp.write("Really cool data\n")
        fp.write("ASCII\n")
        fp.write("DATASET UNSTRUCTURED_GRID\n\n")
        fp.write("POINTS %d FLOAT\n" % n)
        fp.writelines("%s %s %s\n" % (_f(a), _f(b), _f(0.0)) for a, b in zip(ux, uy))
        fp.write("\n")
        if len(tri):
            fp.write("CELLS %d %d\n" % (len(tri), 4 * len(tri)))
            fp.writelines("%d %d %d %d\n" % (3, a, b, d) for a, b, d in tri)
            fp.write("\n")
            fp.write("CELL_TYPES %d\n" % len(tri))
            fp.write("5 " * len(tri))
            fp.write("\n\n")
        fp.write("POINT_DATA %d\n" % n)
        fp.write("SCALARS VelocitySquare FLOAT\n")
        fp.write("LOOKUP_TABLE default\n")
        fp.writelines("%s\n" % _f(v) for v in vv)
        fp.write("\n\n")
        fp.write("VECTORS Velocity FLOAT\n")
        fp.writelines("%s %s %s\n" % (_f(a), _f(b), _f(0.0)) for a, b in zip(vx, vy))
    return name


_TYPE_NAMES = "ABCDEF"


def save_vtf(sys, path: str, s: int = 0):
    """MDSystem::saveVTF (femmd2D/src/md3dsystem.cpp:330-352)."""
    x, y, _, _ = sys.get_state()
    with _open(path, s) as fp:
        if s == 0:
            typ = sys.particle_types()
            sig = sys.particle_sigmas()
            for i in range(len(x)):
                fp.write("atom %d type %s radius %s\n" % (i, _TYPE_NAMES[int(typ[i]) % 6],
                                                         _f(float(sig[i]) / 2.0)))
            for a, b in sys.bond_pairs():
                fp.write("bond %d:%d\n" % (a, b))
            box = sys.get_box()
            fp.write("pbc %s %s %s\n" % (_f(box["x"]), _f(box["y"]), _f(1.0)))
        fp.write("timestep\n")
        fp.writelines("%s %s %s\n" % (_f(a), _f(b), _f(0.0)) for a, b in zip(x, y))


def save_cluster_vtf(sys, path: str, s: int = 0, scale: float = 1.0):
    """MDSystem::saveClusterVTF (femmd2D/src/md3dsystem.cpp:354-377): the
    cluster centres of mass, radius 0.5 * scale + 0.5."""
    c = sys.cluster_props() if sys.cluster_count() else None
    nc = 0 if c is None else len(c["mass"])
    with _open(path, s) as fp:
        if s == 0:
            typ = sys.particle_types()
            for i in range(nc):
                t = int(typ[i]) if i < len(typ) else 0
                fp.write("atom %d type %s radius %s\n" % (i, _TYPE_NAMES[t % 6], _f(0.5 * scale + 0.5)))
            box = sys.get_box()
            fp.write("pbc %s %s %s\n" % (_f(box["x"]), _f(box["y"]), _f(1.0)))
        fp.write("timestep\n")
        for i in range(nc):
            fp.write("%s %s %s\n" % (_f(c["cm"][i, 0]), _f(c["cm"][i, 1]), _f(0.0)))
