"""Pins the CPU oracle (oracle/) against what the mathematics fixes.

The reference ships no tests, golden vectors or fixtures for this path
(its tree has none), and its C++ cannot be compiled here, so each oracle
function is checked by closed forms, invariants, brute force or finite
differences, chosen so that a dropped term, a wrong sign or index, or a
transposed operand fails at least one of them.
"""
import math

import numpy as np
import pytest

import oracle as O
from femmd_old_amd import configs

RC = 2.5
USHIFT = configs.U_SHIFT_25


def two_body(r, eps=1.0, sigma=1.0, rc=RC, ushift=USHIFT, L=50.0):
    x = np.array([10.0, 10.0 + r])
    y = np.array([20.0, 20.0])
    s = O.OracleSystem(L, L, x, y, sigma=[sigma, sigma], eps=[eps], rc=[rc], ushift=[ushift])
    fx, fy, en = O.forces(s, pred_f32=False, all_pairs=True)
    return fx, fy, en


# ---------------------------------------------------------------------------
# LJ (femmd2D/src/interactionsCuda.cu:43-100)
# ---------------------------------------------------------------------------
def test_lj_minimum_is_force_free():
    r = 2 ** (1 / 6)
    fx, fy, en = two_body(r)
    assert abs(fx[0]) < 1e-12 and abs(fx[1]) < 1e-12
    # U counts each ordered pair with 1/2: total = u(r) = -eps + shift
    assert en[O.OR_U_LJ] == pytest.approx(-1.0 + USHIFT, abs=1e-12)


def test_lj_at_sigma():
    fx, fy, en = two_body(1.0)
    assert en[O.OR_U_LJ] == pytest.approx(USHIFT, abs=1e-12)
    # |F| = 24 eps / sigma, repulsive: particle 0 sits left, pushed to -x
    assert fx[0] == pytest.approx(-24.0, rel=1e-12)
    assert fx[1] == pytest.approx(24.0, rel=1e-12)
    # virial f_ij . r_ij = 24 (each ordered pair adds 1/2)
    assert en[O.OR_W_LJ] == pytest.approx(24.0, rel=1e-12)


@pytest.mark.parametrize("sigma,eps", [(1.0, 1.0), (1.3, 0.7), (0.8, 2.0)])
def test_lj_scaling_and_cutoff_in_sigma_units(sigma, eps):
    # cutoff is rc * sigma (interactionsCuda.cu:88): just inside / outside
    inside = two_body(RC * sigma * (1 - 1e-9), eps, sigma)
    outside = two_body(RC * sigma * (1 + 1e-9), eps, sigma)
    assert inside[2][O.OR_NPAIRS] == 2 and outside[2][O.OR_NPAIRS] == 0
    assert outside[0][0] == 0.0
    # reduced units: F(r) sigma / eps depends on r / sigma only
    f = two_body(1.1 * sigma, eps, sigma)[0][0]
    f1 = two_body(1.1, 1.0, 1.0)[0][0]
    assert f * sigma / eps == pytest.approx(f1, rel=1e-12)


def test_lj_energy_force_consistency_fd():
    rng = np.random.default_rng(7)
    L, m = 12.0, 10
    g = (np.arange(m) + 0.5) * (L / m)
    X, Y = np.meshgrid(g, g)
    x = X.ravel() + rng.uniform(-0.15, 0.15, m * m)
    y = Y.ravel() + rng.uniform(-0.15, 0.15, m * m)
    s = O.OracleSystem(L, L, x, y, eps=[1.0], rc=[RC], ushift=[0.0])
    fx, fy, _ = O.forces(s, pred_f32=False)
    h = 1e-6
    for i in [0, 7, 23]:
        for comp, f in (("x", fx), ("y", fy)):
            xp, yp = x.copy(), y.copy()
            xm, ym = x.copy(), y.copy()
            if comp == "x":
                xp[i] += h
                xm[i] -= h
            else:
                yp[i] += h
                ym[i] -= h
            up = O.forces(s.with_positions(xp, yp), pred_f32=False)[2][O.OR_U_LJ]
            um = O.forces(s.with_positions(xm, ym), pred_f32=False)[2][O.OR_U_LJ]
            s.with_positions(x, y)
            g = -(up - um) / (2 * h)
            assert g == pytest.approx(f[i], rel=1e-6, abs=1e-6)


def test_force_scales_match_fd_of_term_magnitudes():
    # row 1 of the oracle's force scale is sum_terms r |df/dr|: check it by
    # central differences of the single-term force magnitude
    for r in (0.95, 1.12, 1.6, 2.3):
        x = np.array([10.0, 10.0 + r])
        y = np.array([20.0, 20.0])
        s = O.OracleSystem(50.0, 50.0, x, y, eps=[1.3], rc=[RC], ushift=[0.0])
        _, _, _, fs = O.forces(s, pred_f32=False, scale=True)
        h = 1e-6
        def mag(rr):
            s2 = O.OracleSystem(50.0, 50.0, [10.0, 10.0 + rr], y, eps=[1.3], rc=[RC], ushift=[0.0])
            return O.forces(s2, pred_f32=False)[0][0]
        dfdr = (mag(r + h) - mag(r - h)) / (2 * h)
        assert fs[1][0] == pytest.approx(r * abs(dfdr), rel=1e-5)
        assert fs[0][0] == pytest.approx(abs(mag(r)), rel=1e-12)
    # bond: r dF/dr = k r
    s = O.OracleSystem(30.0, 30.0, [5.0, 6.2], [5.0, 5.0], lj_on=False,
                       bonds=([0], [1], [40.0], [1.0]))
    _, _, _, fs = O.forces(s, pred_f32=False, scale=True)
    assert fs[1][0] == pytest.approx(40.0 * 1.2, rel=1e-12)
    assert fs[0][0] == pytest.approx(40.0 * 0.2, rel=1e-9)


def test_lj_newton_third_law_and_virial_theorem():
    w = configs.c1(with_disk=False)
    s = O.from_workload(w)
    fx, fy, en = O.forces(s)
    scale = np.abs(fx).sum()
    assert abs(fx.sum()) < 1e-10 * scale and abs(fy.sum()) < 1e-10 * scale


def triangular_lattice(a, m):
    """m x m triangular lattice (rows offset by a/2), periodic box."""
    h = a * math.sqrt(3) / 2
    j, i = np.meshgrid(np.arange(m), np.arange(m), indexing="ij")
    x = a * (i + 0.5 * (j % 2))
    y = h * j
    return x.ravel().astype(float), y.ravel().astype(float), a * m, h * m


def test_lj_perfect_lattice_closed_form():
    # per-particle energy of a perfect triangular lattice = 1/2 sum over the
    # lattice shells inside rc: 6 at a, 6 at sqrt3 a, 6 at 2a, 12 at sqrt7 a
    a, m = 1.1, 24  # m even keeps the offset rows periodic
    x, y, Lx, Ly = triangular_lattice(a, m)
    s = O.OracleSystem(Lx, Ly, x, y, eps=[1.0], rc=[RC], ushift=[USHIFT])
    fx, fy, en = O.forces(s, pred_f32=False)
    u = lambda r: 4 * (r ** -12 - r ** -6) + USHIFT
    shells = [(6, a), (6, math.sqrt(3) * a), (6, 2 * a)]
    if math.sqrt(7) * a < RC:
        shells.append((12, math.sqrt(7) * a))
    assert all(r < RC for _, r in shells) and 3 * a > RC
    e_site = 0.5 * sum(k * u(r) for k, r in shells)
    assert en[O.OR_U_LJ] / (m * m) == pytest.approx(e_site, rel=1e-12)
    assert np.abs(fx).max() < 1e-10 and np.abs(fy).max() < 1e-10
    assert en[O.OR_NPAIRS] == m * m * sum(k for k, _ in shells)


# ---------------------------------------------------------------------------
# interacting sets and binning
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("L,n,seed", [(7.6, 30, 1), (12.0, 150, 2), (31.0, 900, 3)])
def test_pair_sets_cell_walk_equals_all_pairs(L, n, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, L, n).astype(np.float32)
    y = rng.uniform(0, L, n).astype(np.float32)
    excl = np.where(rng.uniform(size=n) < 0.3, rng.integers(0, 3, n), -1)
    sig = rng.choice([0.8, 1.0, 1.2], n)
    s = O.OracleSystem(L, L, x, y, sigma=sig, excl=excl, eps=[1.0], rc=[RC])
    a = O.pair_sets(s, 0)
    b = O.pair_sets(s, 1)
    for u, v in zip(a, b):
        np.testing.assert_array_equal(u, v)


def test_pair_sets_symmetric_and_exclusions():
    rng = np.random.default_rng(11)
    L, n = 15.0, 200
    x = rng.uniform(0, L, n).astype(np.float32)
    y = rng.uniform(0, L, n).astype(np.float32)
    excl = np.where(np.arange(n) < 40, 0, -1)
    s = O.OracleSystem(L, L, x, y, excl=excl, eps=[1.0], rc=[RC])
    cnt, s1, s2 = O.pair_sets(s, 0)
    # symmetry: sum_i sum_{j in N(i)} (j+1) counted both ways equals
    # sum_j cnt_j (j+1) when every pair is mutual
    assert s1.sum() == (cnt.astype(np.int64) * (np.arange(n) + 1)).sum()
    # excluded group members never see each other: their sets only hold
    # ids >= 40 (non-members)
    for i in range(40):
        if cnt[i]:
            assert s1[i] >= cnt[i] * 41


def test_pair_set_predicate_matches_fp64_away_from_cutoff():
    rng = np.random.default_rng(5)
    L, n = 20.0, 400
    x = rng.uniform(0, L, n).astype(np.float32)
    y = rng.uniform(0, L, n).astype(np.float32)
    s = O.OracleSystem(L, L, x, y, eps=[1.0], rc=[RC])
    cnt, _, _ = O.pair_sets(s, 0)
    dx = x[:, None].astype(float) - x[None, :]
    dy = y[:, None].astype(float) - y[None, :]
    dx -= L * np.round(dx / L)
    dy -= L * np.round(dy / L)
    r = np.sqrt(dx * dx + dy * dy)
    np.fill_diagonal(r, np.inf)
    far = np.abs(r - RC) > 1e-4
    near_count = (r < RC).sum(1)
    ok = far.all(1)
    np.testing.assert_array_equal(cnt[ok], near_count[ok])


def test_cell_keys_properties():
    L, ncx = 74.0, 29
    xs = np.array([0.0, -0.0, 1e-7, L / ncx, np.nextafter(np.float32(L), 0), L, 73.99999],
                  np.float32)
    keys = O.cell_keys(xs, np.zeros_like(xs), L, L, ncx, ncx)
    assert keys.min() >= 0 and keys.max() < ncx
    assert keys[5] == ncx - 1  # x == L is clamped (nnCuda.cu:12)
    rng = np.random.default_rng(3)
    x = rng.uniform(0, L, 10000).astype(np.float32)
    y = rng.uniform(0, L, 10000).astype(np.float32)
    k = O.cell_keys(x, y, L, L, ncx, ncx)
    ix, iy = k % ncx, k // ncx
    cs = L / ncx
    # a particle lies in its cell up to the FP32 rounding of x * ncx / L
    assert np.all(x >= ix * cs - 1e-4) and np.all(x < (ix + 1) * cs + 1e-4)
    assert np.all(y >= iy * cs - 1e-4) and np.all(y < (iy + 1) * cs + 1e-4)
    # monotone in x within a row
    order = np.argsort(x)
    assert np.all(np.diff(ix[order]) >= 0)


def test_wrap_into_box():
    L = 74.0
    x = np.array([-1e-9, -0.5, 74.0, 74.5, 148.2, 30.0, -74.0], np.float32)
    y = x.copy()
    O.wrap_f32(x, y, L, L)
    assert np.all(x >= 0) and np.all(x <= L)
    np.testing.assert_allclose(x[1], 73.5, rtol=1e-6)
    np.testing.assert_allclose(x[3], 0.5, rtol=1e-5)
    assert x[5] == np.float32(30.0)


def test_counting_sort_is_stable_sort():
    rng = np.random.default_rng(9)
    key = rng.integers(0, 97, 5000).astype(np.int32)
    start, perm = O.counting_sort(key, 97)
    np.testing.assert_array_equal(perm, np.argsort(key, kind="stable"))
    np.testing.assert_array_equal(np.diff(start), np.bincount(key, minlength=97))
    assert start[0] == 0 and start[-1] == 5000


# ---------------------------------------------------------------------------
# bonds (femmd2D/src/interactions.cpp:143-173)
# ---------------------------------------------------------------------------
def test_bond_rest_length_and_fd():
    L = 30.0
    x = np.array([5.0, 6.0, 29.5])
    y = np.array([5.0, 5.5, 5.2])
    bonds = ([0, 1, 2], [1, 2, 0], [100.0, 30.0, 10.0], [math.hypot(1, 0.5), 1.5, 2.0])
    s = O.OracleSystem(L, L, x, y, lj_on=False, bonds=bonds)
    fx, fy, en = O.forces(s, pred_f32=False)
    assert abs(fx.sum()) < 1e-12 and abs(fy.sum()) < 1e-12
    # bond 0 at rest: only bonds 1, 2 contribute energy
    d1 = math.hypot(6.0 - 29.5 + 30, 0.5 - 0.2) - 1.5
    d2 = math.hypot(29.5 - 30 - 5.0, 0.2) - 2.0
    assert en[O.OR_U_BOND] == pytest.approx(0.5 * 30 * d1 ** 2 + 0.5 * 10 * d2 ** 2, rel=1e-12)
    h = 1e-6
    for i in range(3):
        xp = x.copy()
        xp[i] += h
        xm = x.copy()
        xm[i] -= h
        up = O.forces(s.with_positions(xp, y), pred_f32=False)[2][O.OR_U_BOND]
        um = O.forces(s.with_positions(xm, y), pred_f32=False)[2][O.OR_U_BOND]
        s.with_positions(x, y)
        assert -(up - um) / (2 * h) == pytest.approx(fx[i], rel=1e-6, abs=1e-7)


@pytest.mark.parametrize("r,r0,k", [(1.3, 1.0, 100.0), (0.8, 1.0, 30.0), (2.0, 2.0, 5.0)])
def test_bond_virial_closed_form(r, r0, k):
    # W_bond = f_ij . r_ij = -k (r - r0) / r * r^2 = -k (r - r0) r, whatever
    # the bond's direction and image
    L = 20.0
    for th, x0 in ((0.3, 5.0), (2.1, 19.6)):
        x = np.array([x0, (x0 + r * math.cos(th)) % L])
        y = np.array([7.0, 7.0 + r * math.sin(th)])
        s = O.OracleSystem(L, L, x, y, lj_on=False, bonds=([0], [1], [k], [r0]))
        _, _, en = O.forces(s, pred_f32=False)
        assert en[O.OR_W_BOND] == pytest.approx(-k * (r - r0) * r, rel=1e-9, abs=1e-12)
        assert en[O.OR_U_BOND] == pytest.approx(0.5 * k * (r - r0) ** 2, rel=1e-9, abs=1e-12)


# ---------------------------------------------------------------------------
# CST (femmd2D/src/interactions.cpp:532-689)
# ---------------------------------------------------------------------------
def one_element(F, model, E=50.0, nu=0.3, L=100.0, ref=None, ccw=True):
    ref = np.array([[0.0, 0.0], [1.0, 0.0], [0.3, 0.9]]) if ref is None else ref
    if not ccw:
        ref = ref[[0, 2, 1]]
    cur = ref @ np.asarray(F).T + 40.0
    s = O.OracleSystem(L, L, cur[:, 0], cur[:, 1], lj_on=False, tri=[0, 1, 2],
                       refx=ref[:, 0] + 40, refy=ref[:, 1] + 40, E=[E], nu=[nu], model=model)
    fx, fy, en = O.forces(s, pred_f32=False)
    A = 0.5 * abs((ref[1, 0] - ref[0, 0]) * (ref[2, 1] - ref[0, 1]) -
                  (ref[1, 1] - ref[0, 1]) * (ref[2, 0] - ref[0, 0]))
    return s, cur, fx, fy, en, A


def lame(E, nu):
    return E * nu / ((1 - 2 * nu) * (1 + nu)), E / (2 * (1 + nu))


@pytest.mark.parametrize("model", [0, 1, 2])
def test_cst_identity_is_stress_free(model):
    _, _, fx, fy, en, _ = one_element(np.eye(2), model)
    assert np.abs(fx).max() < 1e-12 and np.abs(fy).max() < 1e-12
    assert abs(en[O.OR_U_FEA]) < 1e-12


@pytest.mark.parametrize("s_", [1.05, 0.9, 1.3])
def test_cst_uniaxial_closed_forms(s_):
    E, nu = 50.0, 0.3
    lam, mu = lame(E, nu)
    F = np.diag([s_, 1.0])
    # Neo-Hookean: psi = mu/2 (s^2 - 1) - mu ln s + lam/2 (ln s)^2
    _, _, _, _, en, A = one_element(F, 0, E, nu)
    psi = 0.5 * mu * (s_ ** 2 - 1) - mu * math.log(s_) + 0.5 * lam * math.log(s_) ** 2
    assert en[O.OR_U_FEA] == pytest.approx(psi * A, rel=1e-12)
    # SVK: E11 = (s^2 - 1)/2, psi = (mu + lam/2) E11^2; legacy = 2 x SVK
    e11 = 0.5 * (s_ ** 2 - 1)
    _, _, _, _, en1, _ = one_element(F, 1, E, nu)
    assert en1[O.OR_U_FEA] == pytest.approx((mu + 0.5 * lam) * e11 ** 2 * A, rel=1e-12)
    _, _, _, _, en2, _ = one_element(F, 2, E, nu)
    assert en2[O.OR_U_FEA] == pytest.approx(2 * en1[O.OR_U_FEA], rel=1e-12)


@pytest.mark.parametrize("model", [0, 1])
def test_cst_forces_are_energy_gradient(model):
    F = np.array([[1.08, 0.05], [-0.03, 0.95]])
    s, cur, fx, fy, en, _ = one_element(F, model)
    x, y = cur[:, 0].copy(), cur[:, 1].copy()
    h = 1e-6
    for i in range(3):
        for comp in (0, 1):
            xp, yp, xm, ym = x.copy(), y.copy(), x.copy(), y.copy()
            (xp if comp == 0 else yp)[i] += h
            (xm if comp == 0 else ym)[i] -= h
            up = O.forces(s.with_positions(xp, yp), pred_f32=False)[2][O.OR_U_FEA]
            um = O.forces(s.with_positions(xm, ym), pred_f32=False)[2][O.OR_U_FEA]
            g = -(up - um) / (2 * h)
            assert g == pytest.approx((fx if comp == 0 else fy)[i], rel=1e-6, abs=1e-8)
    assert abs(fx.sum()) < 1e-12 and abs(fy.sum()) < 1e-12


def test_cst_rotation_invariance_and_small_strain_limit():
    E, nu = 50.0, 0.3
    lam, mu = lame(E, nu)
    F = np.array([[1.002, 0.001], [0.0005, 0.999]])
    th = 0.7
     # rigid rotation: same energy, forces rotated
    R = np.array([[math.cos(th), -math.sin(th)], [math.sin(th), math.cos(th)]])
    _, _, fx, fy, en, A = one_element(F, 0, E, nu)
    _, _, gx, gy, en2, _ = one_element(R @ F, 0, E, nu)
    assert en2[O.OR_U_FEA] == pytest.approx(en[O.OR_U_FEA], rel=1e-9)
    rot = R @ np.stack([fx, fy])
    np.testing.assert_allclose(rot, np.stack([gx, gy]), atol=1e-10)
    # small strain: psi -> mu eps:eps + lam/2 tr(eps)^2 (plane strain)
    eps_ = 0.5 * (F + F.T) - np.eye(2)
    psi_lin = mu * (eps_ * eps_).sum() + 0.5 * lam * np.trace(eps_) ** 2
    assert en[O.OR_U_FEA] / A == pytest.approx(psi_lin, rel=5e-3)


def p_dot_f(F, model, lam, mu):
    """P:F of the material law, in closed form from the invariants: Neo-Hookean
    P = mu (F - F^-T) + lam lnJ F^-T gives mu (I1 - 2) + 2 lam lnJ; SVK
    P = F S with S = 2 mu E + lam tr(E) I, E = (C - I)/2 gives S:C; the
    legacy GPU energy is twice SVK."""
    F = np.asarray(F, float)
    if model == 0:
        J = np.linalg.det(F)
        return mu * ((F * F).sum() - 2.0) + 2.0 * lam * math.log(J)
    C_ = F.T @ F
    E_ = 0.5 * (C_ - np.eye(2))
    S_ = 2 * mu * E_ + lam * np.trace(E_) * np.eye(2)
    return (2.0 if model == 2 else 1.0) * (S_ * C_).sum()


@pytest.mark.parametrize("model", [0, 1, 2])
@pytest.mark.parametrize("F", [np.diag([1.2, 1.0]), np.array([[1.08, 0.05], [-0.03, 0.95]]),
                               np.array([[0.9, 0.2], [0.1, 1.1]])])
def test_cst_virial_closed_form(model, F):
    # sum_k f_k . x_k = tr(H^T Ds) = -A0 tr(Bm P^T F Dm)... = -A0 P:F for a
    # homogeneous deformation (the reference sums f . unfoldedPos,
    # femmd2D/src/interactions.cpp:681; sum_k f_k = 0 makes it independent of
    # the origin)
    E, nu = 50.0, 0.3
    lam, mu = lame(E, nu)
    _, _, _, _, en, A = one_element(F, model, E, nu)
    want = -A * p_dot_f(F, model, lam, mu)
    assert en[O.OR_W_FEA] == pytest.approx(want, rel=1e-9, abs=1e-12)


def test_cst_virial_independent_of_position_and_image():
    F = np.array([[1.05, 0.02], [0.01, 0.97]])
    L = 100.0
    ref = np.array([[0.0, 0.0], [1.0, 0.0], [0.3, 0.9]])
    cur = ref @ F.T
    w = []
    for shift in (40.0, -0.5, 99.6):
        c = cur + np.array([shift, 40.0])
        c[:, 0] %= L
        s = O.OracleSystem(L, L, c[:, 0], c[:, 1], lj_on=False, tri=[0, 1, 2], refx=ref[:, 0],
                           refy=ref[:, 1], E=[50.0], nu=[0.3], model=0)
        w.append(O.forces(s, pred_f32=False)[2][O.OR_W_FEA])
    lam, mu = lame(50.0, 0.3)
    A = 0.45
    for v in w:
        assert v == pytest.approx(-A * p_dot_f(F, 0, lam, mu), rel=1e-9)


def test_cst_orientation_uses_signed_area():
    # femmd2D/src/interactions.cpp:596,686 multiply by the signed vol0: a
    # clockwise element reports the negative energy of its CCW twin
    F = np.diag([1.1, 1.0])
    _, _, _, _, en_ccw, _ = one_element(F, 0)
    _, _, _, _, en_cw, _ = one_element(F, 0, ccw=False)
    assert en_cw[O.OR_U_FEA] == pytest.approx(-en_ccw[O.OR_U_FEA], rel=1e-12)


def test_cst_unwrap_across_boundary():
    # same element straddling x = 0 in a periodic box: identical forces
    F = np.array([[1.05, 0.02], [0.01, 0.97]])
    L = 100.0
    ref = np.array([[0.0, 0.0], [1.0, 0.0], [0.3, 0.9]])
    cur = ref @ F.T
    res = []
    for shift in (40.0, -0.5):
        c = cur + np.array([shift, 40.0])
        c[:, 0] %= L
        s = O.OracleSystem(L, L, c[:, 0], c[:, 1], lj_on=False, tri=[0, 1, 2], refx=ref[:, 0],
                           refy=ref[:, 1], E=[50.0], nu=[0.3], model=0)
        res.append(O.forces(s, pred_f32=False))
    np.testing.assert_allclose(res[0][0], res[1][0], atol=1e-12)
    assert res[0][2][O.OR_U_FEA] == pytest.approx(res[1][2][O.OR_U_FEA], rel=1e-12)


# ---------------------------------------------------------------------------
# integrator (femmd2D/src/integrator.cpp:8-29,100-135)
# ---------------------------------------------------------------------------
def test_kdk_free_flight():
    L = 10.0
    x = np.array([1.0, 9.9])
    y = np.array([2.0, 0.05])
    s = O.OracleSystem(L, L, x, y, lj_on=False)
    vx, vy = np.array([0.3, 0.2]), np.array([-0.1, -0.4])
    xo, yo, vxo, vyo, _ = O.run(s, x, y, vx, vy, [1.0, 1.0], 100, 0.01)
    np.testing.assert_allclose(xo, (x + 100 * 0.01 * vx) % L, atol=1e-12)
    np.testing.assert_allclose(yo, (y + 100 * 0.01 * vy) % L, atol=1e-12)
    np.testing.assert_array_equal(vxo, vx)


def test_kdk_harmonic_oscillator_matches_closed_map():
    # particle 1 (mass 2) on a bond of stiffness k to a pinned particle 0:
    # the KDK map per step is M = [[1 - w2 dt2/2, dt], [-w2 dt (1 - w2 dt2/4), 1 - w2 dt2/2]]
    # acting on (q, v) with q the displacement from rest, w2 = k / m
    k, m, dt, r0 = 40.0, 2.0, 0.01, 1.0
    L = 20.0
    x = np.array([5.0, 5.0 + r0 + 0.1])
    y = np.array([5.0, 5.0])
    s = O.OracleSystem(L, L, x, y, lj_on=False, bonds=([0], [1], [k], [r0]))
    mass = np.array([np.inf, m])
    xo, _, vxo, _, _ = O.run(s, x, y, [0.0, 0.3], [0.0, 0.0], mass, 250, dt)
    w2 = k / m
    M = np.array([[1 - w2 * dt * dt / 2, dt], [-w2 * dt * (1 - w2 * dt * dt / 4), 1 - w2 * dt * dt / 2]])
    q = np.linalg.matrix_power(M, 250) @ np.array([0.1, 0.3])
    assert xo[1] - x[0] - r0 == pytest.approx(q[0], rel=1e-9)
    assert vxo[1] == pytest.approx(q[1], rel=1e-9)
    assert xo[0] == x[0]


def test_kdk_time_reversible():
    w = configs.c1(with_disk=False)
    s = O.from_workload(w)
    x, y = s.keep["x"].copy(), s.keep["y"].copy()
    vx, vy = w.vx.copy(), w.vy.copy()
    m = np.ones(w.n_total)
    x1, y1, vx1, vy1, _ = O.run(s, x, y, vx, vy, m, 40, 0.005)
    x2, y2, vx2, vy2, _ = O.run(s, x1, y1, -vx1, -vy1, m, 40, 0.005)
    d = x2 - x
    d -= w.Lx * np.round(d / w.Lx)
    assert np.abs(d).max() < 1e-9
    np.testing.assert_allclose(-vx2, vx, atol=1e-9)


def test_kdk_energy_conservation_c1():
    w = configs.c1()
    s = O.from_workload(w)
    ga = w.global_arrays()
    vx, vy = configs.node_velocities(w)
    _, _, _, _, hist = O.run(s, s.keep["x"], s.keep["y"], vx, vy, ga["mass"], 200, 0.005,
                             history=True)
    etot = hist[:, O.OR_U_LJ] + hist[:, O.OR_U_FEA] + hist[:, O.OR_H_KE]
    drift = (etot.max() - etot.min()) / abs(etot[0])
    assert drift < 5e-3


def test_baseline_matches_fp64_oracle_short():
    # the FP32 Verlet-list restatement of the reference CPU path computes the
    # same dynamics as the FP64 oracle: a few steps agree closely
    w = configs.c1()
    s = O.from_workload(w)
    ga = w.global_arrays()
    vx, vy = configs.node_velocities(w)
    b = O.Baseline(s, vx, vy, ga["mass"], 0.005, threads=1)
    b.run(10)
    (bx, by, bvx, bvy), props = b.state()
    x, y, _, _, _ = O.run(s, s.keep["x"], s.keep["y"], vx, vy, ga["mass"], 10, 0.005)
    d = bx - x
    d -= w.Lx * np.round(d / w.Lx)
    assert np.abs(d).max() < 1e-4


def test_oracle_under_host_sanitizers():
    """Every oracle entry point on a mixed system (two types, mixed sigma,
    exclusions, bonds, CST elements of each model, the threaded baseline)
    built with AddressSanitizer and UndefinedBehaviorSanitizer: no invalid
    access, overflow or undefined operation (make -C oracle sanitize)."""
    import shutil
    import subprocess
    from pathlib import Path
    if shutil.which("make") is None or shutil.which("g++") is None:
        pytest.skip("make / g++ not available")
    odir = Path(__file__).resolve().parent.parent / "oracle"
    r = subprocess.run(["make", "-s", "-C", str(odir), "sanitize"], capture_output=True, text=True,
                       timeout=600)
    try:
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        assert "oracle sanitizer run ok" in r.stdout
        assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
    finally:
        (odir / "sanitize_check").unlink(missing_ok=True)


def test_force_error_terms_report():
    # particle 0: |f| binds; 1: the cancellation term; 2: the resolution term
    rx, ry = np.array([3.0, 0.1, 0.0]), np.zeros(3)
    fs = (np.array([1.0, 2.0, 0.1]), np.array([1.0, 1.0, 50.0]))
    fx = rx + np.array([3e-6, 1e-5, 1e-5])
    t = O.force_error_terms(fx, ry, rx, ry, fs)
    assert (t["net"], t["cancel"], t["resolution"]) == pytest.approx((1 / 3, 1 / 3, 1 / 3))
    # two-term rule: 1e-6, 5e-6, 1e-4 -> one above 1e-5
    assert t["max_two_term"] == pytest.approx(1e-4)
    assert t["above_tol_two_term"] == 1 and t["n"] == 3
    assert O.force_error(fx, ry, rx, ry, fs).max() == pytest.approx(5e-6)


@pytest.mark.parametrize("L", [1.0, 16.0, 74.0, 117.86813, 1179.468, 1024.0, 3.0e5])
def test_wrap_is_identity_inside_the_box(L):
    # pbcCalc (femmd2D/src/def.h:112) in FP32: for 0 < r < L, r / L rounds
    # to at most 1 - 2^-24, so floor is 0 and the result is r itself; the
    # engine's wrap skips the division there.  Checked on the floats just
    # below L, just above 0 and random ones in between.
    Lf = np.float32(L)
    below = [Lf]
    for _ in range(64):
        below.append(np.nextafter(below[-1], np.float32(0)))
    rng = np.random.default_rng(7)
    r = np.concatenate([np.array(below[1:], np.float32),
                        np.float32(1e-30) * np.arange(1, 65, dtype=np.float32),
                        rng.uniform(0, L, 100000).astype(np.float32)])
    r = r[(r > 0) & (r < Lf)]
    q = (r / Lf).astype(np.float32)
    assert q.dtype == np.float32 and np.all(np.floor(q) == 0)
    out = r - np.floor(q) * Lf
    np.testing.assert_array_equal(out.view(np.uint32), r.view(np.uint32))
