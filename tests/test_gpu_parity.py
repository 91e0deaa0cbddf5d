"""Parity of the HIP engine with the CPU oracle, through the C ABI.

Bars:
  * cell keys, cell_start and the interacting sets: bit-exact against the
    oracle's FP32 restatement on the same positions;
  * per-step forces: |f - f_ref| <= 1e-5 * max(|f_ref|, sum_j |f_ij,ref|,
    0.1 sum_j r_ij |df_ij/dr_ij|) per particle against the FP64 oracle (same
    positions, FP32 decisions): relative to the scale of the cancelling
    terms of that force, and never below what FP32 inputs can resolve
    (oracle.force_error states the rule);
  * energies and virials: relative 1e-5 of the FP64 oracle sums;
  * 100-step trajectories against an FP64 oracle rollout;
  * run-twice, graph/no-graph, tile-shape and batched-vs-looped runs are
    bitwise identical.

Both pair paths run every case: the neighbour lists (the engine default
skin 0.6, and the reference's 0.4) and the per-step cell walk (skin 0).  Forces checked after some
steps in list mode come from lists built steps earlier, so a list that
missed an interacting pair fails the force bar.
"""
import numpy as np
import pytest

import oracle as O
from femmd_old_amd import Batch, FemmdError, System, _lib, configs
from parity_helpers import check_energies

pytestmark = pytest.mark.gpu

FORCE_TOL = 1e-5


SKINS = [0.6, 0.4, 0.0]


def engine_and_oracle(w, skin=None):
    s = configs.build_system(w, skin=skin)
    x, y, vx, vy = s.get_state()
    return s, O.from_workload(w, x, y), (x, y, vx, vy)


def force_err(s, osys):
    fx, fy = s.get_forces()
    rx, ry, en, fs = O.forces(osys, pred_f32=True, scale=True)
    return O.force_error(fx, fy, rx, ry, fs), en


# ---------------------------------------------------------------------------
def test_cells_and_counting_sort_bit_exact():
    w = configs.c1()
    s, osys, (x, y, _, _) = engine_and_oracle(w)
    for rep in range(2):
        ncx, ncy, key, start, perm = s.cells()
        okey = O.cell_keys(x, y, w.Lx, w.Ly, ncx, ncy)
        np.testing.assert_array_equal(key, okey)
        ostart, operm = O.counting_sort(okey, ncx * ncy)
        np.testing.assert_array_equal(start, ostart)
        # the whole sorted order: each cell's members by particle index, as
        # the reference's cell fill appends them (femmd2D/src/nn.cpp:8-20)
        np.testing.assert_array_equal(perm, operm)
        s.run(50)
        x, y, _, _ = s.get_state()


@pytest.mark.parametrize("skin", SKINS)
@pytest.mark.parametrize("cfg", ["c1", "c2s", "c3s", "c5s"])
def test_pair_sets_bit_exact(cfg, skin):
    """The interacting sets the force pass uses, at the current positions:
    with neighbour lists (skin > 0) the sets the lists give, lists built
    steps earlier (femmd2D/src/nn.cpp:22-62 walk, interactionsCuda.cu:88-90
    predicate on the current positions); with skin 0 the per-step cell walk.
    c5s is WCA (f -> 0 at the cutoff), where a pair missed near the cutoff
    would not show in the force bar."""
    w = {"c1": lambda: configs.c1(), "c2s": lambda: configs.c2(scale=0.01),
         "c3s": lambda: configs.c3(scale=0.005), "c5s": lambda: configs.c5(scale=0.01)}[cfg]()
    s, osys, _ = engine_and_oracle(w, skin)
    for rep in range(4):
        cnt, s1, s2 = s.pair_sets()
        x, y, _, _ = s.get_state()
        ocnt, os1, os2 = O.pair_sets(osys.with_positions(x, y), 1)
        np.testing.assert_array_equal(cnt, ocnt)
        np.testing.assert_array_equal(s1, os1)
        np.testing.assert_array_equal(s2, os2)
        s.run(7 + 5 * rep)
    if skin > 0:
        assert s.list_stats()["rebuilds"] >= 2


@pytest.mark.parametrize("skin", [0.6, 0.4])
def test_pair_sets_from_stale_lists_hot_system(skin):
    # three times the thermal speed: lists are rebuilt often and used up to
    # the skin / 2 displacement bound; every step's sets are checked
    w = configs.c2(scale=0.01)
    s = configs.build_system(w, commit=False, skin=skin)
    x, y, vx, vy = s.get_state()
    s.set_state(None, None, 3.0 * vx, 3.0 * vy)
    s.allocate_cuda()
    osys = O.from_workload(w)
    for step in range(24):
        cnt, s1, s2 = s.pair_sets()
        x, y, _, _ = s.get_state()
        ocnt, os1, os2 = O.pair_sets(osys.with_positions(x, y), 1)
        np.testing.assert_array_equal(cnt, ocnt)
        np.testing.assert_array_equal(s1, os1)
        np.testing.assert_array_equal(s2, os2)
        s.run(1)
    assert s.list_stats()["rebuilds"] >= 4


@pytest.mark.parametrize("skin", SKINS)
@pytest.mark.parametrize("cfg", ["c1", "c2s", "c3s", "c5s"])
def test_forces_within_1e5_of_fp64(cfg, skin):
    w = {"c1": lambda: configs.c1(), "c2s": lambda: configs.c2(scale=0.01),
         "c3s": lambda: configs.c3(scale=0.005), "c5s": lambda: configs.c5(scale=0.01)}[cfg]()
    s, osys, _ = engine_and_oracle(w, skin)
    for step in range(3):
        x, y, _, _ = s.get_state()
        err, en = force_err(s, osys.with_positions(x, y))
        print(f"[{cfg} skin {skin} step {step * 30}] max rel force err {err.max():.3e}")
        assert err.max() <= FORCE_TOL
        check_energies(s, en, w)
        s.run(30)


@pytest.mark.parametrize("model", [_lib.FM_NEO_HOOKEAN, _lib.FM_SVK, _lib.FM_LEGACY_GPU])
def test_material_models(model):
    w = configs.c1()
    w.model = model
    # strain the disk a little so every model produces forces
    el = w.elastomers[0]
    el_v = el.verts.copy()
    s = configs.build_system(w, commit=False)
    x, y, vx, vy = s.get_state()
    n0 = w.n_solvent
    x[n0:] = (x[n0:] - el.cx) * 1.03 + el.cx
    y[n0:] = (y[n0:] - el.cy) * 0.98 + el.cy
    s.set_state(x, y, None, None)
    s.allocate_cuda()
    x, y, _, _ = s.get_state()
    osys = O.from_workload(w, x, y)
    err, en = force_err(s, osys)
    print(f"[model {model}] max rel force err {err.max():.3e}, U_fea {en[O.OR_U_FEA]:.6f}")
    assert en[O.OR_U_FEA] > 0
    assert err.max() <= FORCE_TOL
    check_energies(s, en, w)


@pytest.mark.parametrize("skin", SKINS)
def test_trajectory_100_steps_and_energy_drift(skin):
    w = configs.c1()
    s, osys, (x0, y0, vx0, vy0) = engine_and_oracle(w, skin)
    ga = w.global_arrays()
    e0 = s.energies()
    s.run(100)
    x, y, vx, vy = s.get_state()
    ox, oy, ovx, ovy, hist = O.run(osys, x0, y0, vx0, vy0, ga["mass"], 100, w.dt, history=True)
    dx = x - ox
    dx -= w.Lx * np.round(dx / w.Lx)
    dy = y - oy
    dy -= w.Ly * np.round(dy / w.Ly)
    dpos = np.hypot(dx, dy)
    dvel = np.hypot(vx - ovx, vy - ovy)
    e1 = s.energies()
    E0 = e0["ke"] + e0["u_lj"] + e0["u_fea"]
    E1 = e1["ke"] + e1["u_lj"] + e1["u_fea"]
    oE1 = hist[-1, O.OR_H_KE] + hist[-1, O.OR_U_LJ] + hist[-1, O.OR_U_FEA]
    print(f"[traj] max |dx| {dpos.max():.3e} rms {np.sqrt((dpos ** 2).mean()):.3e} "
          f"max |dv| {dvel.max():.3e}; E drift gpu {(E1 - E0) / abs(E0):.3e} "
          f"oracle {(oE1 - E0) / abs(E0):.3e}")
    assert dpos.max() < 2e-3
    assert np.sqrt((dpos ** 2).mean()) < 1e-4
    assert abs(E1 - oE1) / abs(E0) < 1e-4


@pytest.mark.parametrize("skin", SKINS)
def test_run_twice_bitwise_and_graph_invariance(skin):
    w = configs.c2(scale=0.005)
    res = []
    for graph, tile, block in [(1, 0, 256), (1, 0, 256), (0, 0, 256), (1, 8, 128), (1, 3, 64)]:
        s = configs.build_system(w, commit=False, skin=skin)
        s.set_tuning(tile, block, graph)
        s.allocate_cuda()
        s.run(37)
        res.append(s.get_state())
        s.close()
    for r in res[1:]:
        for a, b in zip(res[0], r):
            np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("nrep", [6, 20])
@pytest.mark.parametrize("skin", SKINS)
def test_batched_equals_looped_bitwise(skin, nrep):
    # 20 replicas: the rebuild decision is taken by its own kernel and the
    # sort skips the blocks of replicas that do not rebuild
    ws = configs.c4(n_replicas=nrep, n_max=3000)
    singles = []
    for w in ws:
        s = configs.build_system(w, skin=skin)
        s.run(60)
        singles.append((s.get_state(), s.energies()))
    systems = [configs.build_system(w, commit=False, skin=skin) for w in ws]
    b = Batch(systems)
    b.run(60)
    for r, (st, en) in enumerate(singles):
        bs = b.get_state(r)
        for a, c in zip(st, bs):
            np.testing.assert_array_equal(a, c)
        be = b.energies(r)
        for k in ("ke", "u_lj", "w_lj", "u_fea"):
            assert be[k] == pytest.approx(en[k], rel=1e-12, abs=1e-12)


def test_cell_overflow_is_reported_not_faulting():
    # 200 particles in one 3x3-cell box corner: more than a cell can hold
    s = System(8.0, 8.0)
    s.set_pair_potential([1.0], [2.5], [0.0])
    rng = np.random.default_rng(0)
    s.add_particles(rng.uniform(0.1, 2.0, 200), rng.uniform(0.1, 2.0, 200))
    s.set_lj(False)
    s.allocate_cuda()
    with pytest.raises(FemmdError) as e:
        s.calc_forces()
    assert e.value.code == _lib.FM_ERR_DEVICE
    assert s.device_errors() & _lib.FM_DEV_CELL_OVERFLOW
    s.clear_errors()
    assert s.device_errors() == 0


def test_inverted_element_is_reported():
    w = configs.c1()
    s = configs.build_system(w, commit=False)
    x, y, _, _ = s.get_state()
    el = w.elastomers[0]
    n0 = w.n_solvent
    t = el.tris[0]
    # mirror one node of the first triangle through the opposite edge
    p = np.array([x[n0 + t[0]], y[n0 + t[0]]])
    a = np.array([x[n0 + t[1]], y[n0 + t[1]]])
    b = np.array([x[n0 + t[2]], y[n0 + t[2]]])
    d = b - a
    proj = a + d * np.dot(p - a, d) / np.dot(d, d)
    q = 2 * proj - p
    x[n0 + t[0]], y[n0 + t[0]] = q
    s.set_state(x, y, None, None)
    s.allocate_cuda()
    with pytest.raises(FemmdError):
        s.calc_forces()
    assert s.device_errors() & _lib.FM_DEV_INVERTED


def test_hot_system_lists_stay_complete():
    """Three times the thermal speed: many rebuilds; forces from lists built
    steps earlier must still meet the bar at every check."""
    w = configs.c2(scale=0.005)
    s, osys, (x, y, vx, vy) = engine_and_oracle(w, 0.4)
    s.set_state(None, None, 3.0 * vx, 3.0 * vy)
    worst = 0.0
    for k in range(12):
        s.run(17)
        x, y, _, _ = s.get_state()
        err, _ = force_err(s, osys.with_positions(x, y))
        worst = max(worst, err.max())
    st = s.list_stats()
    print(f"[hot] worst force err {worst:.3e}, lists {st}")
    assert worst <= FORCE_TOL
    assert 3 <= st["rebuilds"] < 12 * 17
    assert st["mean"] > 5 and st["max"] <= 512
    assert s.device_errors() == 0


def test_list_and_cell_paths_agree():
    """Same interacting sets, different summation order: the two paths stay
    within FP32 rounding of each other over a short run."""
    w = configs.c1()
    a = configs.build_system(w, skin=0.4)
    b = configs.build_system(w, skin=0.0)
    fa = a.get_forces()
    fb = b.get_forces()
    scale = np.abs(np.concatenate(fb)).max()
    assert np.abs(np.concatenate(fa) - np.concatenate(fb)).max() <= 1e-5 * scale
    a.run(50)
    b.run(50)
    xa, ya, _, _ = a.get_state()
    xb, yb, _, _ = b.get_state()
    d = np.hypot((xa - xb + w.Lx / 2) % w.Lx - w.Lx / 2, (ya - yb + w.Ly / 2) % w.Ly - w.Ly / 2)
    assert d.max() < 1e-3
    assert a.list_stats()["rebuilds"] >= 1 and b.list_stats()["rebuilds"] == 0


@pytest.mark.parametrize("skin", [0.6, 0.0])
def test_full_sheet_tiles_fit(skin):
    """The 128 x 256-node sheet (hex lattice at close packing) in solvent:
    staged tiles must fit the LDS capacity the engine sizes."""
    w = configs.c5()
    s = configs.build_system(w, skin=skin)
    s.run(40)
    assert s.device_errors() == 0


def test_large_system_lists():
    """~114k particles (tiles with and without the periodic wrap, many list
    rebuilds): forces from lists built steps earlier meet the bar."""
    w = configs.c2(scale=0.1)
    s = configs.build_system(w, skin=0.6)
    osys = O.from_workload(w, *s.get_state()[:2])
    for k in range(3):
        x, y, _, _ = s.get_state()
        err, _ = force_err(s, osys.with_positions(x, y))
        print(f"[c2m step {k * 25}] max rel force err {err.max():.3e}")
        assert err.max() <= FORCE_TOL
        s.run(25)
    assert s.device_errors() == 0 and s.list_stats()["rebuilds"] >= 2
