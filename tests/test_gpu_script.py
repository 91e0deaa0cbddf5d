"""A femmd2D/src/test.js-style study through the script front end, scaled
down: elastomers from the reference square mesh on a lattice (the script's
add_elastomers with fccFlag = 1), WCA, set_velocities, a run with
properties sampled, the per-run file outputs, get_props / print_props."""
import math
from pathlib import Path

import numpy as np
import pytest

from femmd_old_amd import script

pytestmark = pytest.mark.gpu

MESH = Path(__file__).resolve().parent / "golden" / "meshes"


def test_test_js_study(tmp_path, monkeypatch, capsys):
    monkeypatch.chdir(tmp_path)
    S = script.ScriptSystem()
    S.sample_every = 50
    ucells, dens, T = 2, 0.4, 1.0
    N = 2 * ucells * ucells
    L = math.sqrt(N / dens)
    v0 = math.sqrt(2.0 * T * (1.0 - 1.0 / N))
    S.init(L, L)
    S.set_dt(0.001)
    S.set_equilibrium_steps(0)
    S.set_average_average_steps(100)
    S.set_youngs_modulus(50.0)
    S.set_pair_potential([1.0], [2 ** (1 / 6)], [1.0])
    nodes = script.read_file(str(MESH / "squareVertices.txt"), script.FILECONTENT.FLOAT)
    cells = script.read_file(str(MESH / "squareTriangles.txt"), script.FILECONTENT.INT)
    nd = nodes.size // 2
    r0 = 0.5 * math.sqrt(0.91 / (nd * math.pi))
    scale = 1.0 / r0
    S.set_scale(scale)
    S.set_box(scale * L, scale * L)
    nodes = nodes * np.float32(scale)
    ncells = int(math.floor(math.sqrt(N / 2)))
    b = scale * L / ncells
    for ny in range(ncells):
        for nx in range(ncells):
            for a in (0.0, 0.5):
                S.add_elastomer(nodes, cells, (nx + a) * b, (ny + a) * b, 0, 0, 1.0, N, True)
    S.set_velocities(v0)
    assert S.info()["nelastomers"] == 2 * ncells * ncells
    S.allocate_cuda()
    S.init_neighbor_list()
    script.make_dir("data")
    script.change_dir("data")
    S.save_trajectory(True)
    S.set_trajectory_steps(100)
    S.save_vtk("initial_state")
    S.save_vtf("movie.vtf", 0)
    S.calc_cluster_props()
    S.save_cluster_vtf("movieCluster.vtf", 0)
    S.save_cluster_xyz("movie.xyz", 0)
    for _ in range(2):
        S.run(200)
        S.print_props()
        p = S.get_props()
        script.append_to_file("data.dat", "%d %s\n" % (p["steps"], p["uKinetic"]["avgval"]))
        S.save_vtk("movie")
        S.save_vtf("movie.vtf", 1)
        S.save_cluster_vtf("movieCluster.vtf", 1)
        S.save_cluster_xyz("movie.xyz", 1)
    e = S.engine
    assert e.device_errors() == 0
    p = S.get_props()
    assert p["steps"] == 400
    assert np.isfinite(p["uTotal"]["avgval"]) and p["uKinetic"]["avgval"] > 0
    out = capsys.readouterr().out.strip().splitlines()
    assert len(out) >= 4 and int(out[-2].split()[0]) == 400
    assert len(Path("data.dat").read_text().splitlines()) == 2
    assert Path("traj.xyz").exists() and len(list(Path(".").glob("movie*.vtk"))) == 2
