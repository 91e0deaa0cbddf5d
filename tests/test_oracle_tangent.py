"""The oracle's forward-mode rollout derivative (or_tangent) against central
differences of whole FP64 rollouts (or_run): the check that lets the
engine's adjoint be compared with FP64 finite differences along the
trajectory the engine itself took.

or_tangent differentiates the smooth branch each step is on (both force
evaluations of a step use the pairs of the unperturbed configuration).  A
whole-rollout difference also sees pairs that cross the cutoff inside its
interval; the WCA force is continuous there but its slope is not, which
biases the quotient by O(h) (the sigma directions, which move every pair,
show it first): at h = 1e-7 the bias is below the bar, and it shrinks in
proportion to h."""
import numpy as np
import pytest

import oracle as O
from femmd_old_amd import configs


def _setup(nsteps):
    w = configs.c5(scale=0.01)
    ga = w.global_arrays()
    xf, yf = ga["x"].astype(np.float32), ga["y"].astype(np.float32)
    L = O.box_f32(w.Lx)
    O.wrap_f32(xf, yf, L, L)
    x0, y0 = xf.astype(np.float64), yf.astype(np.float64)
    vx0, vy0 = (a.astype(np.float64) for a in configs.node_velocities(w))
    return w, ga, x0, y0, vx0, vy0, float(np.float32(w.dt))


@pytest.mark.parametrize("which", ["E", "eps", "sigma", "all"])
def test_tangent_equals_rollout_fd(which):
    nsteps = 40
    w, ga, x0, y0, vx0, vy0, dt = _setup(nsteps)
    E, sig, eps = ga["E"].astype(np.float64), ga["sigma"].astype(np.float64), float(w.eps)
    m = ga["mass"]
    o = O.from_workload(w, x0, y0)
    tx, ty = O.trajectory(o, x0, y0, vx0, vy0, m, nsteps, dt)
    rng = np.random.default_rng(3)
    dE = rng.standard_normal(E.size) * E if which in ("E", "all") else np.zeros(E.size)
    de = 0.7 * eps if which in ("eps", "all") else 0.0
    ds = -0.4 if which in ("sigma", "all") else 0.0
    dx, dy = O.tangent(o, tx, ty, dt, m, dE, de, ds)
    h = 1e-7

    def final(sg):
        oo = O.from_workload(w, x0, y0, E=E + sg * h * dE, eps=eps + sg * h * de,
                             sigma=sig + sg * h * ds)
        x, y, _, _, _ = O.run(oo, x0, y0, vx0, vy0, m, nsteps, dt)
        return x, y

    (xp, yp), (xm, ym) = final(1), final(-1)
    L = O.box_f32(w.Lx)
    rx = xp - xm
    ry = yp - ym
    rx -= L * np.round(rx / L)
    ry -= L * np.round(ry / L)
    rx /= 2 * h
    ry /= 2 * h
    err = np.hypot(dx - rx, dy - ry)
    ref = np.hypot(rx, ry)
    assert ref.max() > 0
    print(f"[{which}] max |tangent - fd| {err.max():.3e} of max |fd| {ref.max():.3e}")
    assert np.linalg.norm(err) <= 1e-6 * np.linalg.norm(ref)
    # the tangent's own differences are converged: another step agrees
    dx2, dy2 = O.tangent(o, tx, ty, dt, m, dE, de, ds, 1e-5)
    assert np.linalg.norm(np.hypot(dx2 - dx, dy2 - dy)) <= 1e-6 * np.linalg.norm(ref)
    # the trajectory is the rollout's own
    x, y, _, _, _ = O.run(o, x0, y0, vx0, vy0, m, nsteps, dt)
    np.testing.assert_array_equal(tx[-1], x)
    np.testing.assert_array_equal(ty[-1], y)


def test_rollout_fd_bias_shrinks_with_h():
    """Along a sigma direction the whole-rollout quotient's cutoff bias is
    linear in h: 10x smaller h, about 10x closer (until FP64 rounding)."""
    nsteps = 40
    w, ga, x0, y0, vx0, vy0, dt = _setup(nsteps)
    sig, m = ga["sigma"].astype(np.float64), ga["mass"]
    o = O.from_workload(w, x0, y0)
    tx, ty = O.trajectory(o, x0, y0, vx0, vy0, m, nsteps, dt)
    dx, dy = O.tangent(o, tx, ty, dt, m, None, 0.0, -0.4)
    L = O.box_f32(w.Lx)
    errs = []
    for h in (1e-5, 1e-6, 1e-7):
        f = []
        for sg in (1, -1):
            oo = O.from_workload(w, x0, y0, sigma=sig + sg * h * -0.4)
            x, y, _, _, _ = O.run(oo, x0, y0, vx0, vy0, m, nsteps, dt)
            f.append((x, y))
        rx, ry = f[0][0] - f[1][0], f[0][1] - f[1][1]
        rx = (rx - L * np.round(rx / L)) / (2 * h)
        ry = (ry - L * np.round(ry / L)) / (2 * h)
        errs.append(np.linalg.norm(np.hypot(dx - rx, dy - ry)) / np.linalg.norm(np.hypot(rx, ry)))
    print("relative bias at h = 1e-5, 1e-6, 1e-7:", errs)
    assert errs[0] > 3 * errs[1] > 9 * errs[2]
    assert errs[2] <= 1e-6


def test_tangent_one_step_short_misses_the_bar():
    """A derivative one step short (a step dropped at a segment boundary,
    the kind of slip an adjoint sweep can make) differs from the full one far
    beyond the bar the engine's gradient is held to."""
    nsteps = 20
    w, ga, x0, y0, vx0, vy0, dt = _setup(nsteps)
    m = ga["mass"]
    o = O.from_workload(w, x0, y0)
    tx, ty = O.trajectory(o, x0, y0, vx0, vy0, m, nsteps, dt)
    E = ga["E"].astype(np.float64)
    dE = np.random.default_rng(1).standard_normal(E.size) * E
    full = O.tangent(o, tx, ty, dt, m, dE)
    # one step fewer along the same trajectory: the last step's derivative dropped
    short = O.tangent(o, tx[:-1], ty[:-1], dt, m, dE)
    d = np.hypot(full[0] - short[0], full[1] - short[1])
    assert np.linalg.norm(d) > 1e-3 * np.linalg.norm(np.hypot(*full))
