"""A femmd2D/src/test.js-style study through the script front end, scaled
down: elastomers from the reference square mesh on a lattice (the script's
add_elastomers with fccFlag = 1), WCA, set_velocities, a run with
properties sampled, the per-run file outputs, get_props / print_props.

The averaged properties get_props reports after each run (uKinetic,
uPotential, uTotal, virial, feaVirial, pressure: the mean of the samples
of the last averaging block, femmd2D/src/properties.cpp:15-98) are checked
against the same quantities formed from an FP64 oracle rollout of the same
study (same initial state, masses, elements and exclusions; 400 steps),
relative 1e-4 (the virial and the pressure relative to the sum of the
magnitudes of the element virial's terms, which cancel).  Run once without
bonds and once with System.set_kBond(k > 0), which (as in the reference,
femmd2D/src/interface.cpp:455-465 with femmd2D/src/particles.cpp:335-337)
bonds every triangle edge of every elastomer, shared edges twice, rest
length the edge's length at creation; the oracle then carries those bonds
independently built from the mesh."""
import math
import sys
from pathlib import Path

import numpy as np
import pytest

import oracle as O
from femmd_old_amd import script

pytestmark = pytest.mark.gpu

MESH = Path(__file__).resolve().parent / "golden" / "meshes"


@pytest.mark.parametrize("kbond", [0.0, 20.0])
def test_test_js_study(tmp_path, monkeypatch, capsys, kbond):
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
    if kbond:
        S.set_kBond(kbond)
    assert S.info()["nelastomers"] == 2 * ncells * ncells
    S.allocate_cuda()
    S.init_neighbor_list()
    ntri = cells.size // 3 * 2 * ncells * ncells
    assert S.engine.n_bonds() == (3 * ntri if kbond else 0)
    oracle_props = _oracle_study(S, nodes, cells, ncells, b, N, scale * L, kbond)
    script.make_dir("data")
    script.change_dir("data")
    S.save_trajectory(True)
    S.set_trajectory_steps(100)
    S.save_vtk("initial_state")
    S.save_vtf("movie.vtf", 0)
    S.calc_cluster_props()
    S.save_cluster_vtf("movieCluster.vtf", 0)
    S.save_cluster_xyz("movie.xyz", 0)
    for r in range(2):
        S.run(200)
        S.print_props()
        p = S.get_props()
        # the averaging block ending at this run's last step: samples at
        # steps 200 (r + 1) - 50 and 200 (r + 1)
        ref, scale = oracle_props([200 * (r + 1) - 50, 200 * (r + 1)])
        for k, v in ref.items():
            print(f"  step {p['steps']} {k}: engine {p[k]['avgval']:.9e} oracle {v:.9e} "
                  f"(scale {scale[k]:.3e})", file=sys.stderr)
            assert abs(p[k]["avgval"] - v) <= 1e-4 * (abs(v) + scale[k]) + 1e-7, k
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


def _oracle_study(S, nodes, cells, ncells, b, N, L, kbond):
    """An FP64 oracle of the script's engine system (the same elastomers in
    the same order: masses spread as rho A / 3 per corner, one exclusion
    group each, reference shapes at their placement; with kbond > 0 a bond
    of stiffness kbond on each of every triangle's three edges, rest length
    the edge's length in the reference shape) and a function giving the
    averaged properties over the samples at the given steps."""
    e = S.engine
    x0, y0, vx0, vy0 = e.get_state()
    n = x0.shape[0]
    verts = np.asarray(nodes, np.float32).reshape(-1, 2)
    tri1 = np.asarray(cells, np.int64).reshape(-1, 3) - 1
    nv = verts.shape[0]
    # the element areas of the mesh and the per-node mass of one elastomer
    p0, p1, p2 = (verts[tri1[:, k]].astype(np.float64) for k in range(3))
    area = 0.5 * ((p2[:, 1] - p0[:, 1]) * (p1[:, 0] - p0[:, 0]) -
                  (p1[:, 1] - p0[:, 1]) * (p2[:, 0] - p0[:, 0]))
    rho = N / area.sum()
    m1 = np.zeros(nv)
    for k in range(3):
        np.add.at(m1, tri1[:, k], rho * area / 3.0)
    mass = np.tile(m1.astype(np.float32), n // nv).astype(np.float64)
    excl = np.repeat(np.arange(n // nv), nv).astype(np.int32)
    centres = [((nx + a) * b, (ny + a) * b) for ny in range(ncells) for nx in range(ncells)
               for a in (0.0, 0.5)]
    refx = np.concatenate([verts[:, 0] + np.float32(cx) for cx, _ in centres])
    refy = np.concatenate([verts[:, 1] + np.float32(cy) for _, cy in centres])
    bonds = None
    if kbond:
        edges = np.concatenate([tri1[:, [0, 1]], tri1[:, [0, 2]], tri1[:, [1, 2]]])
        bi = np.concatenate([edges[:, 0] + c * nv for c in range(len(centres))])
        bj = np.concatenate([edges[:, 1] + c * nv for c in range(len(centres))])
        r0 = np.hypot(refx[bi].astype(np.float64) - refx[bj], refy[bi].astype(np.float64) - refy[bj])
        bonds = (bi, bj, np.full(bi.shape[0], kbond), r0)
    Lf = O.box_f32(L)
    osys = O.OracleSystem(Lf, Lf, x0, y0, excl=excl, eps=[1.0], rc=[2 ** (1 / 6)], ushift=[1.0],
                          bonds=bonds, tri=e.triangles(), refx=refx, refy=refy, E=50.0, nu=0.3,
                          model=0)
    _, _, _, _, hist = O.run(osys, x0, y0, vx0, vy0, mass, 400, float(np.float32(0.001)),
                             history=True)
    V = Lf * Lf

    def props(steps):
        h = hist[np.asarray(steps) - 1]
        ke = h[:, O.OR_NEN]
        up = (h[:, O.OR_U_LJ] + h[:, O.OR_U_BOND] + h[:, O.OR_U_FEA]) / n
        vals = {"uKinetic": ke / n, "uPotential": up, "uTotal": ke / n + up,
                "virial": h[:, O.OR_W_LJ], "feaVirial": h[:, O.OR_W_FEA],
                "pressure": (ke + 0.5 * (h[:, O.OR_W_LJ] + h[:, O.OR_W_FEA])) / V}
        # the virials cancel: judged on the sum of the magnitudes of their
        # terms (OR_WS_FEA), as the single-step parity tests judge them
        ws = float(np.mean(h[:, O.OR_WS_FEA]))
        scale = {k: 0.0 for k in vals}
        scale["feaVirial"] = ws
        scale["pressure"] = 0.5 * ws / V
        return {k: float(np.mean(v)) for k, v in vals.items()}, scale

    return props
