"""Pair tables with several types and per-particle sigma, against the FP64
oracle (femmd2D/src/interactionsCuda.cu:43-100: sigma_ij = (sigma_i +
sigma_j) / 2, eps and rc (in sigma units) and the additive shift from the
row-major ntypes x ntypes tables, femmd2D/src/interactions.cpp:65-79).

  * two types with different eps, rc and shift, mixed sigmas, and six
    types (more than the reference's usual two) with uniform and mixed
    sigmas: the neighbour-list path (lists cover the table's largest cutoff,
    each pair is then filtered by its own) and the cell walk;
  * one type with mixed sigmas: the neighbour-list path with per-particle
    sigma (the list radius from the largest sigma pair) and the cell walk;
  * a six-type system stepped 100 steps on the lists against the FP64
    oracle's rollout;
  * replicas with different two- and six-type tables stepped as one batch,
    bitwise equal to each stepped alone.
Interacting sets are bit-exact with the oracle's FP32 predicate; forces
meet the 1e-5 bar of oracle.force_error; energies and virials 1e-5.

  * particle masses: two masses (the kick loads each particle's 1/m) and
    one mass other than 1 for all (the kick takes 1/m from the replica's
    descriptor): 100 steps against the FP64 oracle's KDK rollout
    (femmd2D/src/integrator.cpp:100-162), with the trajectory bars of the
    C1 parity test.
"""
import numpy as np
import pytest

import oracle as O
from femmd_old_amd import System

pytestmark = pytest.mark.gpu


def _liquid(n, L, seed):
    """n particles on a jittered square lattice of the box (spacing >= 1.2,
    jitter 0.08: no close contacts)."""
    rng = np.random.default_rng(seed)
    m = int(np.ceil(np.sqrt(n)))
    g = (np.arange(m) + 0.5) * (L / m)
    gx, gy = np.meshgrid(g, g)
    x = (gx.ravel()[:n] + rng.uniform(-0.08, 0.08, n)).astype(np.float32)
    y = (gy.ravel()[:n] + rng.uniform(-0.08, 0.08, n)).astype(np.float32)
    vx = rng.normal(0, 0.5, n).astype(np.float32)
    vy = rng.normal(0, 0.5, n).astype(np.float32)
    return x, y, vx, vy, rng


def _check(s, osys_of, steps=(0, 25, 25)):
    for k in steps:
        if k:
            s.run(k)
        x, y, _, _ = s.get_state()
        osys = osys_of(x, y)
        cnt, s1, s2 = s.pair_sets()
        ocnt, os1, os2 = O.pair_sets(osys, 1)
        np.testing.assert_array_equal(cnt, ocnt)
        np.testing.assert_array_equal(s1, os1)
        fx, fy = s.get_forces()
        rx, ry, en, fs = O.forces(osys, pred_f32=True, scale=True)
        err = O.force_error(fx, fy, rx, ry, fs)
        e = s.energies()
        print(f"  max rel force err {err.max():.3e}, U {e['u_lj']:.6f} / {en[O.OR_U_LJ]:.6f}")
        assert err.max() <= 1e-5
        assert abs(e["u_lj"] - en[O.OR_U_LJ]) <= 1e-5 * (abs(en[O.OR_U_LJ]) + 1.0)
        assert abs(e["w_lj"] - en[O.OR_W_LJ]) <= 1e-5 * (abs(en[O.OR_W_LJ]) + 1.0)
    assert s.device_errors() == 0


@pytest.mark.parametrize("skin", [0.6, 0.0])
def test_two_type_table_with_mixed_sigma(skin):
    L, n = 60.0, 2400
    x, y, vx, vy, rng = _liquid(n, L, 11)
    sig = rng.choice([0.9, 1.1], n).astype(np.float32)
    typ = rng.integers(0, 2, n).astype(np.int32)
    eps = np.array([1.0, 0.6, 0.6, 1.4], np.float32)
    rc = np.array([2.5, 1.8, 1.8, 2.2], np.float32)
    us = np.array([0.0, 0.05, 0.05, -0.02], np.float32)
    s = System(L, L, dt=0.002)
    s.set_skin(skin)
    s.set_pair_potential(eps, rc, us)
    s.add_particles(x, y, vx, vy, 1.0, sig, typ)
    s.allocate_cuda()
    Lf = O.box_f32(L)
    _check(s, lambda px, py: O.OracleSystem(Lf, Lf, px, py, sigma=sig, type=typ, eps=eps, rc=rc,
                                            ushift=us))
    if skin > 0:
        assert s.list_stats()["rebuilds"] >= 1


def _six_type_tables(rng):
    nt = 6
    e = rng.uniform(0.5, 1.5, (nt, nt))
    r = rng.uniform(1.6, 2.6, (nt, nt))
    u = rng.uniform(-0.05, 0.05, (nt, nt))
    sym = lambda a: ((a + a.T) / 2).astype(np.float32).ravel()  # symmetric like the reference's tables
    return nt, sym(e), sym(r), sym(u)


@pytest.mark.parametrize("sigma", ["uniform", "mixed"])
@pytest.mark.parametrize("skin", [0.6, 0.0])
def test_six_type_table(skin, sigma):
    L, n = 70.0, 3200
    x, y, vx, vy, rng = _liquid(n, L, 21)
    nt, eps, rc, us = _six_type_tables(rng)
    typ = rng.integers(0, nt, n).astype(np.int32)
    sig = (rng.uniform(0.9, 1.1, n) if sigma == "mixed" else np.ones(n)).astype(np.float32)
    s = System(L, L, dt=0.002)
    s.set_skin(skin)
    s.set_pair_potential(eps, rc, us)
    s.add_particles(x, y, vx, vy, 1.0, sig, typ)
    s.allocate_cuda()
    Lf = O.box_f32(L)
    _check(s, lambda px, py: O.OracleSystem(Lf, Lf, px, py, sigma=sig, type=typ, eps=eps, rc=rc,
                                            ushift=us))
    if skin > 0:
        assert s.list_stats()["rebuilds"] >= 1


def test_six_type_lists_trajectory():
    L, n, dt = 70.0, 3200, 0.004
    x, y, vx, vy, rng = _liquid(n, L, 22)
    nt, eps, rc, us = _six_type_tables(rng)
    typ = rng.integers(0, nt, n).astype(np.int32)
    s = System(L, L, dt=dt)
    s.set_skin(0.6)
    s.set_pair_potential(eps, rc, us)
    s.add_particles(x, y, vx, vy, 1.0, 1.0, typ)
    s.allocate_cuda()
    x0, y0, vx0, vy0 = s.get_state()
    Lf = O.box_f32(L)
    osys = O.OracleSystem(Lf, Lf, x0, y0, type=typ, eps=eps, rc=rc, ushift=us)
    s.run(100)
    x1, y1, vx1, vy1 = s.get_state()
    ox, oy, ovx, ovy, _ = O.run(osys, x0, y0, vx0, vy0, np.ones(n), 100, dt)
    dx = x1 - ox
    dx -= Lf * np.round(dx / Lf)
    dy = y1 - oy
    dy -= Lf * np.round(dy / Lf)
    dpos = np.hypot(dx, dy)
    print(f"  six types, lists: max |dx| {dpos.max():.3e} rms {np.sqrt((dpos ** 2).mean()):.3e}")
    assert dpos.max() < 2e-3
    assert np.sqrt((dpos ** 2).mean()) < 1e-4
    assert s.list_stats()["rebuilds"] >= 1
    assert s.device_errors() == 0


@pytest.mark.parametrize("skin", [0.6, 0.4, 0.0])
def test_single_type_mixed_sigma(skin):
    L, n = 75.0, 3600
    x, y, vx, vy, rng = _liquid(n, L, 12)
    sig = rng.uniform(0.85, 1.15, n).astype(np.float32)
    s = System(L, L, dt=0.002)
    s.set_skin(skin)
    s.set_pair_potential([1.0], [2.5], [0.0])
    s.add_particles(x, y, vx, vy, 1.0, sig)
    s.allocate_cuda()
    Lf = O.box_f32(L)
    _check(s, lambda px, py: O.OracleSystem(Lf, Lf, px, py, sigma=sig, eps=[1.0], rc=[2.5],
                                            ushift=[0.0]))
    if skin > 0:
        assert s.list_stats()["rebuilds"] >= 1


@pytest.mark.parametrize("masses", ["mixed", "uniform"])
@pytest.mark.parametrize("skin", [0.6, 0.0])
def test_masses_trajectory(masses, skin):
    L, n, dt = 60.0, 2400, 0.005
    x, y, vx, vy, rng = _liquid(n, L, 13)
    m = (rng.choice([1.0, 2.5], n) if masses == "mixed" else np.full(n, 2.5)).astype(np.float32)
    s = System(L, L, dt=dt)
    s.set_skin(skin)
    s.set_pair_potential([1.0], [2.5], [0.0])
    s.add_particles(x, y, vx, vy, m)
    s.allocate_cuda()
    x0, y0, vx0, vy0 = s.get_state()
    Lf = O.box_f32(L)
    osys = O.OracleSystem(Lf, Lf, x0, y0, eps=[1.0], rc=[2.5], ushift=[0.0])
    s.run(100)
    x1, y1, vx1, vy1 = s.get_state()
    ox, oy, ovx, ovy, _ = O.run(osys, x0, y0, vx0, vy0, m, 100, dt)
    dx = x1 - ox
    dx -= Lf * np.round(dx / Lf)
    dy = y1 - oy
    dy -= Lf * np.round(dy / Lf)
    dpos = np.hypot(dx, dy)
    dvel = np.hypot(vx1 - ovx, vy1 - ovy)
    print(f"  {masses} skin {skin}: max |dx| {dpos.max():.3e} rms "
          f"{np.sqrt((dpos ** 2).mean()):.3e} max |dv| {dvel.max():.3e}")
    assert dpos.max() < 2e-3
    assert np.sqrt((dpos ** 2).mean()) < 1e-4
    assert s.device_errors() == 0


def _typed_system(seed, nt, n, L, skin, commit):
    """A liquid with an nt-type table (1, 2 or 6 types) and mixed sigma."""
    from femmd_old_amd import System as Sys
    x, y, vx, vy, rng = _liquid(n, L, seed)
    if nt == 1:
        eps, rc, us = [1.0], [2.5], [0.0]
    elif nt == 2:
        eps, rc, us = [1.0, 0.6, 0.6, 1.4], [2.5, 1.8, 1.8, 2.2], [0.0, 0.05, 0.05, -0.02]
    else:
        nt, eps, rc, us = _six_type_tables(rng)
    typ = rng.integers(0, nt, n).astype(np.int32)
    sig = rng.uniform(0.9, 1.1, n).astype(np.float32)
    s = Sys(L, L, dt=0.004)
    s.set_skin(skin)
    s.set_pair_potential(eps, rc, us)
    s.add_particles(x, y, vx, vy, 1.0, sig, typ)
    if commit:
        s.allocate_cuda()
    return s


@pytest.mark.parametrize("skin", [0.6, 0.0])
def test_batch_of_typed_tables_equals_looped(skin):
    """Replicas with different pair tables (two and six types, mixed sigma)
    stepped as one batch give the bits of each stepped alone: every replica
    carries its own table and list radius (the largest cutoff of its own
    table)."""
    from femmd_old_amd import Batch
    specs = [(31, 6, 2000, 50.0), (32, 2, 2400, 60.0), (33, 6, 3200, 70.0), (34, 2, 900, 36.0)]
    singles = []
    for seed, nt, n, L in specs:
        s = _typed_system(seed, nt, n, L, skin, True)
        s.run(60)
        assert s.device_errors() == 0
        singles.append((s.get_state(), s.energies()))
    b = Batch([_typed_system(seed, nt, n, L, skin, False) for seed, nt, n, L in specs])
    b.run(60)
    assert b.device_errors() == 0
    for r, (st, en) in enumerate(singles):
        for u, v in zip(st, b.get_state(r)):
            np.testing.assert_array_equal(u, v)
        be = b.energies(r)
        for k in ("ke", "u_lj", "w_lj"):
            assert be[k] == pytest.approx(en[k], rel=1e-12, abs=1e-12)
