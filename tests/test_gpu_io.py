"""Reference mesh files through the engine, and the trajectory writers on a
running system.

  * the femmd2D disk mesh (1-based, as read_file reads it), scaled to radius
    5, added with add_elastomer(index_base=1) into c1's solvent: forces
    within 1e-5 of the FP64 oracle built from the same mesh (0-based), and a
    run stays finite;
  * XYZ and VTF frames read back give the engine's positions to the
    printed precision (6 decimals), the VTK file holds the unfolded
    positions and the mesh's cells.
"""
from pathlib import Path

import numpy as np
import pytest

import oracle as O
from femmd_old_amd import configs, io

pytestmark = pytest.mark.gpu

MESH = Path(__file__).resolve().parent / "golden" / "meshes"


def _mesh_system():
    v, t = io.load_mesh(str(MESH / "diskVertices.txt"), str(MESH / "diskTriangles.txt"))
    v = (v * np.float32(10.0)).astype(np.float32)
    w = configs.c1(with_disk=False)
    L = w.Lx
    keep = np.hypot(w.x - L / 2, w.y - L / 2) > 6.5
    el = configs.Elastomer(v.astype(np.float64), (t - 1).astype(np.int32), L / 2, L / 2,
                           np.full(len(t), 50.0), np.full(len(t), 0.3), 1.0)
    w = configs.Workload("mesh", L, L, w.dt, w.x[keep], w.y[keep], w.vx[keep], w.vy[keep],
                         w.mass[keep], w.sigma[keep], [el], seed=1)
    from femmd_old_amd import System
    s = System(L, L, dt=w.dt)
    s.set_pair_potential([w.eps], [w.rc], [w.ushift])
    s.add_particles(w.x, w.y, w.vx, w.vy, w.mass, w.sigma)
    # the mesh as read from the files: flat, 1-based
    s.add_elastomer(v.ravel(), t.ravel(), L / 2, L / 2, 0, 0, 1.0, 0.0, True, node_mass=1.0,
                    E=np.full(len(t), 50.0), nu=np.full(len(t), 0.3))
    s.set_youngs_modulus(50.0, 0.3)
    s.allocate_cuda()
    return w, s


def test_reference_mesh_forces():
    w, s = _mesh_system()
    x, y, _, _ = s.get_state()
    fx, fy = s.get_forces()
    rx, ry, _, fs = O.forces(O.from_workload(w, x, y), pred_f32=True, scale=True)
    err = O.force_error(fx, fy, rx, ry, fs)
    assert err.max() <= 1e-5, err.max()
    s.run(100)
    e = s.energies()
    assert np.isfinite(e["ke"] + e["u_lj"] + e["u_fea"]) and s.device_errors() == 0


def test_writers_read_back(tmp_path):
    w, s = _mesh_system()
    s.run(30)
    x, y, _, _ = s.get_state()
    p = tmp_path / "t.xyz"
    s.save_xyz(str(p), 0)
    s.run(5)
    s.save_xyz(str(p), 1)
    x2, y2, _, _ = s.get_state()
    lines = p.read_text().splitlines()
    n = len(x)
    assert int(lines[0]) == n and int(lines[n + 2]) == n
    f1 = np.array([[float(t) for t in l.split()[1:3]] for l in lines[2:n + 2]])
    f2 = np.array([[float(t) for t in l.split()[1:3]] for l in lines[n + 4:2 * n + 4]])
    assert np.max(np.abs(f1 - np.stack([x, y], 1))) <= 5.1e-7 * max(1.0, w.Lx)
    assert np.max(np.abs(f2 - np.stack([x2, y2], 1))) <= 5.1e-7 * max(1.0, w.Lx)
    p = tmp_path / "t.vtf"
    s.save_vtf(str(p), 0)
    txt = p.read_text().splitlines()
    assert sum(l.startswith("atom ") for l in txt) == n
    i = txt.index("timestep")
    fr = np.array([[float(t) for t in l.split()[:2]] for l in txt[i + 1:i + 1 + n]])
    assert np.max(np.abs(fr - np.stack([x2, y2], 1))) <= 5.1e-7 * max(1.0, w.Lx)
    name = s.save_vtk(str(tmp_path / "m"))
    vt = Path(name).read_text()
    assert "CELLS 116 464" in vt and "CELL_TYPES 116" in vt
    c = s.cluster_props(unfolded=True)
    pts = vt.split("POINTS %d FLOAT\n" % n)[1].split("\n\n")[0].splitlines()
    up = np.array([[float(t) for t in l.split()[:2]] for l in pts])
    assert np.max(np.abs(up[:, 0] - c["ux"])) <= 5.1e-7 * w.Lx
    s.save_cluster_vtf(str(tmp_path / "c.vtf"), 0)
    s.save_cluster_xyz(str(tmp_path / "c.xyz"), 0)
    cx = (tmp_path / "c.xyz").read_text().splitlines()
    assert cx[3] == "1" and cx[-1].split()[0] == "0"
