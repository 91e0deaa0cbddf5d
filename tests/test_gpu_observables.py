"""Observables on the GPU against the oracle's restatements.

  * cluster properties (fm_cluster_props): mass, unfolded centre of mass and
    its velocity, centroid, unfolded positions, for the c1 disk after a run
    and for a disk placed across the periodic edge (either side);
  * cluster virial (fm_cluster_virial) against the O(n^2) oracle on a small
    multi-disk system in solvent, in both pair paths, within 1e-5 of the sum
    of the magnitudes of its terms (FP32 pair forces);
  * temperature rescale (fm_set_temperature): all particles (T = KE / N) and
    a group (the reference's 2 KE_group / (3 N));
  * the Properties layer over a running system.
"""
import numpy as np
import pytest

import oracle as O
from femmd_old_amd import configs
from femmd_old_amd.observables import Properties

pytestmark = pytest.mark.gpu


def _ranges(w):
    r, off = [], w.n_solvent
    for el in w.elastomers:
        n = el.verts.shape[0]
        r.append((off, n))
        off += n
    return r


def _check_props(w, s):
    x, y, vx, vy = s.get_state()
    m = w.global_arrays()["mass"]
    m = np.where(np.isfinite(m), m, 0.0)
    ref, rux, ruy = O.cluster_props(x, y, vx, vy, m, _ranges(w), w.Lx, w.Ly)
    got = s.cluster_props(unfolded=True)
    assert got["mass"] == pytest.approx(ref[:, 0], rel=1e-12)
    L = O.box_f32(w.Lx)
    assert np.max(np.abs(got["cm"] - ref[:, 1:3])) < 1e-9 * L
    assert np.max(np.abs(got["vcm"] - ref[:, 3:5])) < 1e-9 * (1 + np.abs(ref[:, 3:5]).max())
    assert np.max(np.abs(got["centroid"] - ref[:, 5:7])) < 1e-7 * L
    nodes = np.concatenate([np.arange(a, a + n) for a, n in _ranges(w)])
    assert np.array_equal(got["ux"][nodes], rux[nodes])
    assert np.array_equal(got["uy"][nodes], ruy[nodes])


def test_cluster_props_c1():
    w = configs.c1()
    s = configs.build_system(w)
    s.run(50)
    assert s.cluster_count() == 1
    _check_props(w, s)


@pytest.mark.parametrize("cx", [0.5, 73.5])
def test_cluster_props_across_the_edge(cx):
    w = configs.c1(with_disk=False)
    keep = np.hypot(np.mod(w.x - cx + 37, 74) - 37, w.y - 37) > 11.5
    w = configs.Workload("edge", w.Lx, w.Ly, w.dt, w.x[keep], w.y[keep], w.vx[keep], w.vy[keep],
                         w.mass[keep], w.sigma[keep],
                         [configs._disk_elastomer(10.0, cx, 37.0, 50.0, 0.3)], seed=1)
    s = configs.build_system(w)
    got = s.cluster_props()
    assert abs(got["cm"][0, 0] - cx) < 1e-4 and abs(got["cm"][0, 1] - 37.0) < 1e-4
    s.run(20)
    _check_props(w, s)


def _touching_disks(seed=5):
    """Four disks of radius 5 whose rims are 1 apart (inside the cutoff),
    in solvent at density ~0.8."""
    rng = np.random.default_rng(seed)
    L = 48.0
    centres = [(18.0, 18.0), (29.0, 18.0), (18.0, 29.0), (29.5, 29.5)]
    m = 43
    g = (np.arange(m) + 0.5) * (L / m)
    X, Y = np.meshgrid(g, g, indexing="xy")
    X = X.ravel() + rng.uniform(-0.1, 0.1, m * m)
    Y = Y.ravel() + rng.uniform(-0.1, 0.1, m * m)
    keep = np.ones(X.shape, bool)
    for cx, cy in centres:
        keep &= np.hypot(X - cx, Y - cy) > 6.0
    x, y = np.mod(X[keep], L), np.mod(Y[keep], L)
    n = x.shape[0]
    vx, vy = configs.maxwell(rng, n, 1.0, np.ones(n))
    els = [configs._disk_elastomer(5.0, cx, cy, 50.0, 0.3) for cx, cy in centres]
    return configs.Workload("touch", L, L, 0.005, x, y, vx, vy, np.ones(n), np.ones(n), els,
                            seed=seed)


@pytest.mark.parametrize("skin", [0.6, 0.0])
def test_cluster_virial_matches_oracle(skin):
    w = _touching_disks()
    s = configs.build_system(w, skin=skin)
    s.run(30)
    x, y, _, _ = s.get_state()
    ga = w.global_arrays()
    cid = -np.ones(w.n_total, np.int64)
    for c, (a, n) in enumerate(_ranges(w)):
        cid[a:a + n] = c
    cm = s.cluster_props()["cm"]
    ref, ab = O.cluster_virial(x, y, ga["sigma"], ga["excl"], cid, cm, w.Lx, w.Ly, w.eps, w.rc)
    got = s.cluster_virial()
    assert ab > 0
    assert abs(got - ref) <= 1e-5 * ab, (got, ref, ab)
    # the query leaves the dynamics as they were
    s.run(5)
    assert s.device_errors() == 0


def test_set_temperature():
    w = configs.c1()
    s = configs.build_system(w)
    s.run(10)
    s.set_temperature(1.7)
    e = s.energies()
    assert e["ke"] / e["n"] == pytest.approx(1.7, rel=1e-6)
    _, _, vx, vy = s.get_state()
    m = w.global_arrays()["mass"]
    ke = 0.5 * np.sum(m * (vx.astype(np.float64) ** 2 + vy.astype(np.float64) ** 2))
    assert ke / w.n_total == pytest.approx(1.7, rel=1e-6)
    ids = np.arange(w.n_solvent)
    s.set_temperature(0.9, ids)
    _, _, vx, vy = s.get_state()
    keg = 0.5 * np.sum(m[ids] * (vx[ids].astype(np.float64) ** 2 + vy[ids].astype(np.float64) ** 2))
    assert 2 * keg / (3 * w.n_total) == pytest.approx(0.9, rel=1e-6)
    s.run(10)
    assert np.isfinite(s.energies()["ke"])


def test_properties_layer():
    w = configs.c1()
    s = configs.build_system(w)
    p = Properties(s, step_avg=20, step_equi=0, skip_steps_rdf=10, avg_count_rdf=2)
    p.run(40, sample_every=10)
    assert p.count in (0, 2)
    assert np.isfinite(p.uTotal.avgval) and p.uKinetic.avgval > 0
    assert np.isfinite(p.clusterVirial.avgval)
