"""The adjoint kernels against central finite differences of the FP64 oracle.

  * force JVP: j = (df/dx) w from the HIP pass against
    (f(x + h w) - f(x - h w)) / 2h of the oracle forces, h = 1e-6;
  * parameter terms w . df/dE_e, w . df/deps, w . df/dsigma against central
    differences of w . f(theta) over each parameter;
  * full discrete adjoint of a rollout of the elastic sheet in solvent
    (WCA, per-element E in [20, 120], bottom row pinned): the gradient of
    the loss over (E_1..E_m, eps, sigma) along 64 random directions against
    central differences of FP64 oracle rollouts, relative 1e-4;
  * the forward part of the adjoint leaves the same state as plain steps,
    and the gradient does not depend on the checkpoint interval (bitwise).

The LJ systems with a plain cutoff (rc = 2.5, no force shift) have a force
jump at the cutoff; a pair that crosses it between x - h w and x + h w has
no derivative there, so those cases allow two particles outside the bar.
"""
import concurrent.futures as cf

import numpy as np
import pytest

import oracle as O
from femmd_old_amd import configs

pytestmark = pytest.mark.gpu

H_X = 1e-6


def _system(cfg):
    w = {"c1": lambda: configs.c1(), "c3s": lambda: configs.c3(scale=0.005),
         "c5s": lambda: configs.c5(scale=0.01)}[cfg]()
    s = configs.build_system(w)
    return w, s


def _oracle_f(w, x, y, **over):
    o = O.from_workload(w, x, y, **over)
    fx, fy, _ = O.forces(o, pred_f32=False)
    return fx, fy


def _wrap(a, L):
    return np.mod(a, L)


@pytest.mark.parametrize("cfg", ["c1", "c3s", "c5s"])
def test_force_jvp_matches_fd(cfg):
    w, s = _system(cfg)
    s.run(20)
    x, y, _, _ = s.get_state()
    rng = np.random.default_rng(11)
    n = x.shape[0]
    wx, wy = rng.standard_normal(n).astype(np.float32), rng.standard_normal(n).astype(np.float32)
    jx, jy, _, _, _ = s.force_jvp(wx, wy)
    L = O.box_f32(w.Lx)
    xd, yd = x.astype(np.float64), y.astype(np.float64)
    fpx, fpy = _oracle_f(w, _wrap(xd + H_X * wx, L), _wrap(yd + H_X * wy, L))
    fmx, fmy = _oracle_f(w, _wrap(xd - H_X * wx, L), _wrap(yd - H_X * wy, L))
    rx, ry = (fpx - fmx) / (2 * H_X), (fpy - fmy) / (2 * H_X)
    d = np.hypot(jx - rx, jy - ry)
    ref = np.hypot(rx, ry)
    rms = np.sqrt(np.mean(ref ** 2))
    bad = np.flatnonzero(d > 1e-4 * np.maximum(ref, rms))
    allowed = 0 if cfg == "c5s" else 2
    assert bad.size <= allowed, (bad[:10], d[bad[:10]], ref[bad[:10]])
    good = np.setdiff1d(np.arange(n), bad)
    assert np.linalg.norm(d[good]) <= 1e-5 * np.linalg.norm(ref[good]) + 1e-30


def test_force_jvp_symmetric():
    """J = df/dx is minus the Hessian of the potential: w1 . J w2 = w2 . J w1."""
    w, s = _system("c5s")
    s.run(10)
    n = s.N
    rng = np.random.default_rng(3)
    a = rng.standard_normal((4, n)).astype(np.float32)
    j1x, j1y, *_ = s.force_jvp(a[0], a[1])
    j2x, j2y, *_ = s.force_jvp(a[2], a[3])
    l = float(np.dot(a[2], j1x.astype(np.float64)) + np.dot(a[3], j1y.astype(np.float64)))
    r = float(np.dot(a[0], j2x.astype(np.float64)) + np.dot(a[1], j2y.astype(np.float64)))
    scale = np.sqrt(np.dot(j1x, j1x) + np.dot(j1y, j1y)) * np.sqrt(2 * n)
    assert abs(l - r) <= 1e-5 * scale, (l, r)


@pytest.mark.parametrize("cfg", ["c1", "c5s"])
def test_parameter_terms_match_fd(cfg):
    w, s = _system(cfg)
    s.run(20)
    x, y, _, _ = s.get_state()
    rng = np.random.default_rng(7)
    n = x.shape[0]
    wx, wy = rng.standard_normal(n).astype(np.float32), rng.standard_normal(n).astype(np.float32)
    _, _, dE, deps, dsig = s.force_jvp(wx, wy)
    xd, yd = x.astype(np.float64), y.astype(np.float64)
    ga = w.global_arrays()

    def wf(**over):
        fx, fy = _oracle_f(w, xd, yd, **over)
        return float(np.dot(wx.astype(np.float64), fx) + np.dot(wy.astype(np.float64), fy))

    # eps and a common shift of every sigma
    h = 1e-6
    fd_eps = (wf(eps=w.eps + h) - wf(eps=w.eps - h)) / (2 * h)
    sig = ga["sigma"].astype(np.float64)
    fd_sig = (wf(sigma=sig + h) - wf(sigma=sig - h)) / (2 * h)
    assert abs(deps - fd_eps) <= 1e-4 * abs(fd_eps) + 1e-6, (deps, fd_eps)
    assert abs(dsig - fd_sig) <= 1e-4 * abs(fd_sig) + 1e-6, (dsig, fd_sig)
    # per-element E on a sample of elements
    E = ga["E"].astype(np.float64)
    for e in rng.choice(E.shape[0], 8, replace=False):
        hE = 1e-6 * E[e]
        Ep, Em = E.copy(), E.copy()
        Ep[e] += hE
        Em[e] -= hE
        fd = (wf(E=Ep) - wf(E=Em)) / (2 * hE)
        assert abs(dE[e] - fd) <= 1e-4 * abs(fd) + 1e-7 * np.abs(dE).max(), (e, dE[e], fd)


# ---------------------------------------------------------------------------
# full adjoint of a rollout
# ---------------------------------------------------------------------------
NSTEPS = 100


def _loss_setup(w):
    ga = w.global_arrays()
    m = ga["mass"]
    nodes = np.arange(w.n_solvent, w.n_total)
    ids = nodes[np.isfinite(m[nodes])].astype(np.int32)
    return ids


def _oracle_loss(w, x0, y0, vx0, vy0, ids, tgt, E, eps, sig):
    o = O.from_workload(w, x0, y0, E=E, eps=eps, sigma=sig)
    m = w.global_arrays()["mass"]
    x, y, _, _, _ = O.run(o, x0, y0, vx0, vy0, m, NSTEPS, float(np.float32(w.dt)))
    L = O.box_f32(w.Lx)
    dx = x[ids] - tgt[:, 0]
    dy = y[ids] - tgt[:, 1]
    dx -= L * np.round(dx / L)
    dy -= L * np.round(dy / L)
    return float(np.mean(dx * dx + dy * dy))


def test_adjoint_rollout_matches_fd():
    w, s = _system("c5s")
    x0, y0, vx0, vy0 = s.get_state()
    ids = _loss_setup(w)
    rng = np.random.default_rng(21)
    tgt = np.stack([x0[ids], y0[ids]], 1).astype(np.float64) + rng.normal(0, 0.3, (ids.size, 2))
    g = s.adjoint(NSTEPS, ids, tgt, ckpt=8)
    ga = w.global_arrays()
    E = ga["E"].astype(np.float64)
    sig = ga["sigma"].astype(np.float64)
    eps = float(w.eps)
    args = (w, x0.astype(np.float64), y0.astype(np.float64), vx0.astype(np.float64),
            vy0.astype(np.float64), ids, tgt)
    l0 = _oracle_loss(*args, E, eps, sig)
    assert abs(g["loss"] - l0) <= 1e-4 * l0, (g["loss"], l0)
    ndir = 64
    dirs = rng.standard_normal((ndir, E.shape[0] + 2))
    h = 1e-6

    def fd(k):
        d = dirs[k]
        dE, de, ds = d[:-2] * E, d[-2] * eps, d[-1] * 1.0
        lp = _oracle_loss(*args, E + h * dE, eps + h * de, sig + h * ds)
        lm = _oracle_loss(*args, E - h * dE, eps - h * de, sig - h * ds)
        pred = float(np.dot(g["dE"], dE) + g["deps"] * de + g["dsigma"] * ds)
        return pred, (lp - lm) / (2 * h)

    with cf.ThreadPoolExecutor(8) as ex:
        res = np.array(list(ex.map(fd, range(ndir))))
    pred, ref = res[:, 0], res[:, 1]
    err = np.abs(pred - ref)
    assert np.linalg.norm(err) <= 1e-4 * np.linalg.norm(ref), (err.max(), np.abs(ref).max())
    assert err.max() <= 1e-4 * np.abs(ref).max(), (err.max(), np.abs(ref).max())
    assert s.device_errors() == 0


@pytest.mark.parametrize("skin", [0.4, 0.0])
def test_adjoint_forward_state_and_checkpoint_invariance(skin):
    """Stored-trajectory adjoint: its forward part is the plain step (bitwise)
    and the state afterwards steps on unchanged.  Checkpointed adjoint: the
    cell walk recomputes each segment bit for bit, so the gradient is the
    same; with neighbour lists every checkpoint starts a fresh list (so the
    recomputation matches its forward segment), which changes the summation
    order against the stored-trajectory run: the gradient agrees to FP32
    rounding."""
    w = configs.c5(scale=0.01)
    sa = configs.build_system(w, skin=skin)
    sb = configs.build_system(w, skin=skin)
    sc = configs.build_system(w, skin=skin)
    x0, y0, _, _ = sa.get_state()
    ids = _loss_setup(w)
    tgt = np.stack([x0[ids] + 0.2, y0[ids]], 1)
    ga = sa.adjoint(60, ids, tgt, ckpt=0)  # whole trajectory stored
    gb = sb.adjoint(60, ids, tgt, ckpt=7)
    sc.run(60)
    for a, c in zip(sa.get_state(), sc.get_state()):
        np.testing.assert_array_equal(a, c)
    if skin == 0.0:
        np.testing.assert_array_equal(ga["dE"], gb["dE"])
        assert ga["deps"] == gb["deps"] and ga["dsigma"] == gb["dsigma"]
        assert ga["loss"] == gb["loss"]
    else:
        scale = np.abs(ga["dE"]).max()
        assert np.abs(ga["dE"] - gb["dE"]).max() <= 1e-4 * scale
        assert abs(ga["deps"] - gb["deps"]) <= 1e-4 * abs(ga["deps"]) + 1e-12
        assert abs(ga["dsigma"] - gb["dsigma"]) <= 1e-4 * abs(ga["dsigma"]) + 1e-12
        assert abs(ga["loss"] - gb["loss"]) <= 1e-6 * ga["loss"]
    # the engine keeps stepping normally after an adjoint sweep (with lists
    # it restarts from a fresh list: same interacting sets, another
    # summation order)
    sa.run(40)
    sc.run(40)
    for a, c in zip(sa.get_state(), sc.get_state()):
        if skin == 0.0:
            np.testing.assert_array_equal(a, c)
        else:
            np.testing.assert_allclose(a, c, rtol=0, atol=1e-3)
