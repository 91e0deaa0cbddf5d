"""Edge cases of the device path, through the C ABI.

  * an empty system (argument error) and a box narrower than 3 cells
    (geometry error) are refused at commit
    (the cell stencil needs 3 distinct cells per axis);
  * one particle in free flight across the periodic edge is bit-exact with
    the FP32 restatement x <- wrap(x + dt v) (oracle.wrap_f32, the
    reference's pbcCalc, femmd2D/src/def.h:112-129), including a wrap that
    rounds to exactly L, whose cell is the reference's clamp to ncx - 1
    (femmd2D/src/nnCuda.cu:9-17) and agrees with oracle.cell_keys;
  * the smallest box (3 x 3 cells) with particles in every cell: forces
    against the FP64 oracle, both pair paths;
  * two distinct particles on one spot interact (self is excluded by index,
    as in the reference) with an infinite LJ force, which is reported as a
    non-finite state, not a fault;
  * more list partners than the list capacity are reported as a list
    overflow, not written past the lists.
"""
import numpy as np
import pytest

import oracle as O
from femmd_old_amd import FemmdError, System, _lib

pytestmark = pytest.mark.gpu


def _system(L, skin, dt=0.005):
    s = System(L, L, dt=dt)
    s.set_skin(skin)
    s.set_pair_potential([1.0], [2.5], [0.0])
    return s


def test_empty_system_is_refused():
    s = System(20.0, 20.0)
    with pytest.raises(FemmdError) as e:
        s.allocate_cuda()
    assert e.value.code == _lib.FM_ERR_ARG
    s.close()


@pytest.mark.parametrize("skin", [0.6, 0.0])
def test_box_below_three_cells_is_refused(skin):
    s = _system(2.9 * (2.5 + skin), skin)
    s.add_particles([1.0, 3.0], [1.0, 1.0])
    with pytest.raises(FemmdError) as e:
        s.allocate_cuda()
    assert e.value.code == _lib.FM_ERR_GEOMETRY
    assert "box too small" in str(e.value)
    s.close()


@pytest.mark.parametrize("skin", [0.6, 0.0])
def test_single_particle_free_flight_bit_exact(skin):
    L = np.float32(10.0)
    dt = np.float32(0.005)
    # the first step lands at -1e-7, which wraps to exactly L in FP32
    x0, vx = np.float32(0.005), np.float32(-1.00002)
    y0, vy = np.float32(9.999), np.float32(0.731)
    s = _system(float(L), skin)
    s.add_particles([x0], [y0], vx=[vx], vy=[vy])
    s.allocate_cuda()
    x = np.array([x0], np.float32)
    y = np.array([y0], np.float32)
    for k in range(1, 41):
        s.run(1)
        x = (x + dt * vx).astype(np.float32)
        y = (y + dt * vy).astype(np.float32)
        O.wrap_f32(x, y, float(L), float(L))
        ex, ey, evx, evy = s.get_state()
        assert ex[0] == x[0] and ey[0] == y[0], (k, ex[0], x[0], ey[0], y[0])
        assert evx[0] == vx and evy[0] == vy
        if k == 1:
            assert x[0] == L  # the rounding case of the reference's wrap
            ncx, ncy, key, _, _ = s.cells()
            np.testing.assert_array_equal(key, O.cell_keys(ex, ey, float(L), float(L), ncx, ncy))
            assert key[0] % ncx == ncx - 1
    assert s.device_errors() == 0
    s.close()


@pytest.mark.parametrize("skin", [0.6, 0.0])
def test_smallest_box_forces(skin):
    L = 3.0 * (2.5 + skin) + 0.01  # 3 x 3 cells
    rng = np.random.default_rng(5)
    g = np.arange(0.4, L - 0.4, 1.1)
    gx, gy = np.meshgrid(g, g)
    x = (gx.ravel() + rng.uniform(-0.1, 0.1, gx.size)).astype(np.float32)
    y = (gy.ravel() + rng.uniform(-0.1, 0.1, gy.size)).astype(np.float32)
    s = _system(L, skin)
    s.add_particles(x, y)
    s.allocate_cuda()
    ncx, ncy, key, _, _ = s.cells()
    assert (ncx, ncy) == (3, 3)
    xs, ys, _, _ = s.get_state()
    Lf = O.box_f32(L)
    osys = O.OracleSystem(Lf, Lf, xs, ys, eps=[1.0], rc=[2.5], ushift=[0.0])
    fx, fy = s.get_forces()
    rx, ry, en, fs = O.forces(osys, pred_f32=True, scale=True)
    err = O.force_error(fx, fy, rx, ry, fs)
    print(f"[3x3 cells skin {skin}] n {x.size} max rel force err {err.max():.3e}")
    assert err.max() <= 1e-5
    assert s.device_errors() == 0
    s.close()


@pytest.mark.parametrize("skin", [0.6, 0.0])
def test_coincident_particles_are_reported(skin):
    """Both pair paths leave a particle's own slot out by index, as the
    reference does (i != j), so a distinct particle on the same spot is an
    interacting pair (the oracle's pair sets count it) with an infinite
    force: reported as a non-finite state, not a fault."""
    s = _system(20.0, skin)
    s.add_particles([5.0, 5.0, 12.0], [5.0, 5.0, 12.0])
    s.allocate_cuda()
    x, y, _, _ = s.get_state()
    cnt, s1, _ = s.pair_sets()
    ocnt, os1, _ = O.pair_sets(O.OracleSystem(O.box_f32(20.0), O.box_f32(20.0), x, y), 1)
    np.testing.assert_array_equal(cnt, ocnt)
    np.testing.assert_array_equal(s1, os1)
    assert list(cnt) == [1, 1, 0]
    with pytest.raises(FemmdError) as e:
        s.run(2)
    assert e.value.code == _lib.FM_ERR_DEVICE
    assert s.device_errors() & _lib.FM_DEV_NONFINITE
    s.clear_errors()
    assert s.device_errors() == 0
    s.close()


def test_list_overflow_is_reported_not_faulting():
    """60 particles within 1.5 of a cell corner (15 per cell, under the cell
    capacity of 32) have 59 list partners each, over the list capacity of
    48 (1.3 x the close-packed count of the list disc): reported as
    FM_DEV_LIST_OVERFLOW, no write past the lists."""
    rng = np.random.default_rng(9)
    r = 1.5 * np.sqrt(rng.uniform(0, 1, 60))
    t = rng.uniform(0, 2 * np.pi, 60)
    x = np.concatenate([20.0 + r * np.cos(t), rng.uniform(2, 8, 20)]).astype(np.float32)
    y = np.concatenate([20.0 + r * np.sin(t), rng.uniform(2, 8, 20)]).astype(np.float32)
    s = _system(40.0, 0.6)
    s.add_particles(x, y)
    with pytest.raises(FemmdError) as e:
        s.allocate_cuda()
        s.calc_forces()
    assert e.value.code == _lib.FM_ERR_DEVICE
    assert s.device_errors() & _lib.FM_DEV_LIST_OVERFLOW
    assert not s.device_errors() & _lib.FM_DEV_CELL_OVERFLOW
    s.clear_errors()
    assert s.device_errors() == 0
    s.close()


@pytest.mark.parametrize("tile_cells", [0, 4])
@pytest.mark.parametrize("skin", [0.6, 0.4])
def test_particles_wrapping_between_list_builds(skin, tile_cells):
    """Particles driven across the periodic edges (a column 0.05 from the
    right edge moving right, a row 0.05 above the bottom edge moving down)
    while the lists built before the crossing are still in use: at every
    step the interacting sets equal the oracle's FP32 restatement.  A
    particle of a boundary cell can wrap between two builds; the lists of
    the boundary columns and rows take the minimum image in every entry,
    and a particle that wrapped is farther than rc from every other centre,
    where the plain difference rejects it as the minimum image would.
    Tiles span the whole row (default) or 4 cells (interior tiles too)."""
    L = 40.0
    a = L / 32
    rng = np.random.default_rng(5)
    gx = 1.2 + a * np.arange(32)  # last column at L - 0.05
    gy = 0.05 + a * np.arange(32)  # first row at 0.05
    X, Y = np.meshgrid(gx, gy)
    x = (X.ravel() + rng.uniform(-0.02, 0.02, X.size)).astype(np.float32)
    y = (Y.ravel() + rng.uniform(-0.02, 0.02, Y.size)).astype(np.float32)
    O.wrap_f32(x, y, np.float32(L), np.float32(L))
    vx = rng.normal(0.0, 0.5, x.size).astype(np.float32)
    vy = rng.normal(0.0, 0.5, x.size).astype(np.float32)
    vx[X.ravel() > L - 1.0] = 2.0
    vy[Y.ravel() < 1.0] = -2.0
    s = _system(L, skin)
    s.add_particles(x, y, vx, vy)
    s.allocate_cuda()
    if tile_cells:
        s.set_tuning(tile_cells=tile_cells)
    s.calc_forces()  # lists built (after a retile too) before counting
    osys = O.OracleSystem(O.box_f32(L), O.box_f32(L), x, y)
    r0 = s.list_stats()["rebuilds"]
    wraps_stale = 0
    for step in range(24):
        cnt, s1, s2 = s.pair_sets()
        xc, yc, _, _ = s.get_state()
        ocnt, os1, os2 = O.pair_sets(osys.with_positions(xc, yc), 1)
        np.testing.assert_array_equal(cnt, ocnt)
        np.testing.assert_array_equal(s1, os1)
        np.testing.assert_array_equal(s2, os2)
        s.run(1)
        xn, yn, _, _ = s.get_state()
        wrapped = int(np.sum((np.abs(xn - xc) > L / 2) | (np.abs(yn - yc) > L / 2)))
        if s.list_stats()["rebuilds"] == r0:
            wraps_stale += wrapped
    print(f"[wrap skin {skin} tiles {tile_cells}] {wraps_stale} crossings under the first "
          f"lists, rebuilds {s.list_stats()['rebuilds'] - r0}")
    assert wraps_stale >= 20
    assert s.device_errors() == 0
    s.close()


def test_long_candidate_runs_in_the_list_build():
    """Cells 6.55 wide (skin 4) at hex density 1.04 hold about 45 particles,
    so a row of three staged cells offers up to about 146 candidates: the list
    build buffers each entry as its offset in that run, here past 127 (the
    offsets stay below 3 cell capacities <= 180 < 256).  128-thread
    workgroups keep the buffer in LDS.  The interacting sets at the start
    and after steps equal the oracle's FP32 restatement."""
    L, d = 65.5, 1.04  # 10 cells of 6.55 per axis
    rows = 2 * int(L / (2 * d * np.sqrt(3) / 2))  # even: the stagger is periodic
    cols = int(L / d)
    gx, gy = np.meshgrid(np.arange(cols) * (L / cols), np.arange(rows) * (L / rows))
    gx = gx + 0.5 * (L / cols) * (np.arange(rows)[:, None] % 2)
    rng = np.random.default_rng(17)
    x = (gx.ravel() + rng.uniform(-0.01, 0.01, gx.size)).astype(np.float32)
    y = (gy.ravel() + rng.uniform(-0.01, 0.01, gy.size)).astype(np.float32)
    O.wrap_f32(x, y, np.float32(L), np.float32(L))
    s = _system(L, 4.0)
    s.add_particles(x, y, rng.normal(0, 0.5, x.size), rng.normal(0, 0.5, x.size))
    s.allocate_cuda()
    s.set_tuning(block=128)
    ncx, ncy, _, _, _ = s.cells()
    counts = np.bincount(O.cell_keys(x, y, O.box_f32(L), O.box_f32(L), ncx, ncy),
                         minlength=ncx * ncy).reshape(ncy, ncx)
    runs = counts + np.roll(counts, 1, axis=1) + np.roll(counts, -1, axis=1)
    print(f"[long runs] N={x.size}, {ncx}x{ncy} cells, particles per cell "
          f"{counts.min()}..{counts.max()}, candidate runs up to {runs.max()}")
    assert ncx == 10 and runs.max() > 128
    osys = O.OracleSystem(O.box_f32(L), O.box_f32(L), x, y)
    for step in range(3):
        cnt, s1, s2 = s.pair_sets()
        xc, yc, _, _ = s.get_state()
        ocnt, os1, os2 = O.pair_sets(osys.with_positions(xc, yc), 1)
        np.testing.assert_array_equal(cnt, ocnt)
        np.testing.assert_array_equal(s1, os1)
        np.testing.assert_array_equal(s2, os2)
        s.run(4)
    print(f"[long runs] lists {s.list_stats()}")
    assert s.device_errors() == 0
    s.close()
