"""Re-ordering of the internal order (no reference counterpart).

A re-sort gives the particles that no element, bond or cluster refers to
new internal ids in the current cell order.  It is a pure relabelling: the
state, the forces, the cluster properties, the interacting sets, the cell
keys and J w read in caller order are bitwise the same after it (the
counting sort ranks a cell's members by caller index, so the sorted order
and every sum over it do not depend on the internal order).  After further
steps the run stays within the parity bars, re-sorts inside a run
(neighbour lists and the per-step cell walk) leave it bitwise the run
without them, and a batch and the same systems stepped alone still agree
bitwise when both re-sort on the same schedule.
"""
import numpy as np
import pytest

import oracle as O
from femmd_old_amd import Batch, configs

pytestmark = pytest.mark.gpu


def _shuffled(w, skin=None, every=0):
    """The workload with its solvent's positions and velocities shuffled
    among the solvent slots (the internal order no longer follows the
    cells, as after a long run)."""
    s = configs.build_system(w, commit=False, skin=skin)
    s.set_reorder_interval(every)
    s.allocate_cuda()
    x, y, vx, vy = s.get_state()
    ns = w.n_solvent
    p = np.random.default_rng(3).permutation(ns)
    for a in (x, y, vx, vy):
        a[:ns] = a[:ns][p]
    s.set_state(x, y, vx, vy)
    return s


@pytest.mark.parametrize("skin", [0.6, 0.0])
def test_reorder_is_a_pure_relabelling(skin):
    w = configs.c2(scale=0.02)
    s = _shuffled(w, skin)
    s.run(3)
    # state, forces and cluster properties: bitwise the same in caller order
    before = (*s.get_state(), *s.get_forces(), s.cluster_props()["cm"])
    s.reorder()
    after = (*s.get_state(), *s.get_forces(), s.cluster_props()["cm"])
    for u, v in zip(before, after):
        np.testing.assert_array_equal(u, v)
    # interacting sets and cell keys (order independent)
    before = (*s.pair_sets(), s.cells()[2])
    s.reorder()
    after = (*s.pair_sets(), s.cells()[2])
    for u, v in zip(before, after):
        np.testing.assert_array_equal(u, v)
    # J w: the same bits (same sorted order, same sums)
    rng = np.random.default_rng(4)
    wx = rng.standard_normal(s.N).astype(np.float32)
    wy = rng.standard_normal(s.N).astype(np.float32)
    j0 = s.force_jvp(wx, wy)[:2]
    s.reorder()
    j1 = s.force_jvp(wx, wy)[:2]
    for u, v in zip(j0, j1):
        np.testing.assert_array_equal(u, v)
    assert s.reorder_info()["count"] == 3
    # further steps: the forces still meet the parity bar
    for k in range(2):
        s.run(30)
        x, y, _, _ = s.get_state()
        fx, fy = s.get_forces()
        rx, ry, en, fs = O.forces(O.from_workload(w, x, y), pred_f32=True, scale=True)
        err = O.force_error(fx, fy, rx, ry, fs)
        print(f"[reorder skin {skin}] max rel force err {err.max():.3e}")
        assert err.max() <= 1e-5
    assert s.device_errors() == 0


def test_reorder_schedule_batched_equals_looped():
    ws = configs.c4(n_replicas=5, n_max=4000)
    singles = []
    for w in ws:
        s = configs.build_system(w, commit=False)
        s.set_reorder_interval(7)
        s.allocate_cuda()
        s.run(60)
        assert s.reorder_info()["count"] >= 3
        singles.append(s.get_state())
    systems = [configs.build_system(w, commit=False) for w in ws]
    b = Batch(systems)
    b.set_reorder_interval(7)  # the batch's own entry point
    b.run(60)
    for r, st in enumerate(singles):
        for a, c in zip(st, b.get_state(r)):
            np.testing.assert_array_equal(a, c)
    assert b.device_errors() == 0


@pytest.mark.parametrize("skin", [0.0, 0.6])
def test_reorder_inside_a_run(skin):
    """Re-sorts between the pieces of a run (every 3 steps here), after a
    drift: with the per-step cell walk the drift has already binned the new
    positions under the old ids, which the re-sort must not leave behind.
    The sorted order ranks a cell's members by their caller index, so it
    does not depend on the internal order: the run is bitwise the run
    without re-sorts, and its final forces meet the parity bar."""
    w = configs.c2(scale=0.02)
    runs = []
    for every in (0, 3):
        s = _shuffled(w, skin, every)
        s.run(40)
        runs.append(s)
    a, b = runs
    assert a.reorder_info()["count"] == 0 and b.reorder_info()["count"] >= 4
    xa, ya, _, _ = a.get_state()
    xb, yb, _, _ = b.get_state()
    L = np.float32(w.Lx)
    dx = xb - xa
    dx -= L * np.round(dx / L)
    dy = yb - ya
    dy -= L * np.round(dy / L)
    d = np.hypot(dx, dy)
    print(f"[reorder in run, skin {skin}] max |dx| vs no re-sort {d.max():.3e}")
    for u, v in zip(a.get_state(), b.get_state()):
        np.testing.assert_array_equal(u, v)
    fx, fy = b.get_forces()
    rx, ry, en, fs = O.forces(O.from_workload(w, xb, yb), pred_f32=True, scale=True)
    assert O.force_error(fx, fy, rx, ry, fs).max() <= 1e-5
    ea, eb = a.energies(), b.energies()
    for k in ("u_lj", "ke"):
        assert abs(eb[k] - ea[k]) <= 1e-12 * abs(ea[k])
    assert b.device_errors() == 0
