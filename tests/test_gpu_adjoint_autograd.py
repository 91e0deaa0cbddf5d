"""The discrete adjoint for an arbitrary loss of the final positions:
fm_adjoint_forward / fm_adjoint_backward (terminal cotangent) through a
torch.autograd.Function (tests/adjoint_autograd.py).

  * the gradient of an MSE loss formed in torch equals the engine's own MSE
    adjoint (fm_adjoint) on the same rollout;
  * the C5 workload as specified, at 10 % of its size (a 40 x 81-node
    sheet, 6,240 elements, in 20,000 WCA solvent particles; 200 steps, a
    checkpoint every 20): the loss is the MSE between the final node
    positions and a target rollout from the same initial state with E
    perturbed by 10 % (a second smooth field).  Its gradient with respect to
    (E_1..E_m, eps, sigma) along 64 random directions against FP64 central
    differences: (1) of every step along the trajectory the engine took
    (or_tangent: the smooth branch of each step, itself pinned to
    whole-rollout differences on the CPU), relative 1e-4; (2) of whole FP64
    oracle rollouts.  The FP32 trajectory
    departs from the FP64 one by chaotic growth of rounding; the loss's
    residual is small (the target differs by a 10 % stiffness change), so
    (2) carries that departure relative to the residual, and its bar is set
    from the measured departure;
  * the recording is consumed by the backward call and dropped when the
    state moves.
"""
import concurrent.futures as cf

import numpy as np
import pytest
import torch

import oracle as O
from adjoint_autograd import FinalPositions, mse_min_image
from femmd_old_amd import FemmdError, _lib, configs

pytestmark = pytest.mark.gpu


def _free_nodes(w):
    m = w.global_arrays()["mass"]
    nodes = np.arange(w.n_solvent, w.n_total)
    return nodes[np.isfinite(m[nodes])].astype(np.int64)


def _params(w):
    ga = w.global_arrays()
    return (torch.tensor(w.elastomers[0].E, dtype=torch.float64, requires_grad=True),
            torch.tensor(float(w.eps), dtype=torch.float64, requires_grad=True),
            torch.tensor(float(ga["sigma"][0]), dtype=torch.float64, requires_grad=True))


@pytest.mark.parametrize("skin", [0.6, 0.0])
def test_autograd_mse_equals_engine_mse_adjoint(skin):
    w = configs.c5(scale=0.01)
    ids = _free_nodes(w)
    s0 = configs.build_system(w, skin=skin)
    x0, y0, _, _ = s0.get_state()
    rng = np.random.default_rng(4)
    tgt = np.stack([x0[ids], y0[ids]], 1) + rng.normal(0, 0.3, (ids.size, 2))
    tgt = tgt.astype(np.float32).astype(np.float64)  # the entry point takes FP32 targets
    g = s0.adjoint(60, ids.astype(np.int32), tgt, ckpt=7)
    E, eps, sig = _params(w)
    pos = FinalPositions.apply(E, eps, sig, w, 60, 7, skin)
    L = O.box_f32(w.Lx)
    loss = mse_min_image(pos, torch.from_numpy(ids), torch.from_numpy(tgt), L)
    loss.backward()
    assert float(loss) == pytest.approx(g["loss"], rel=1e-5)
    scale = np.abs(g["dE"]).max()
    np.testing.assert_allclose(E.grad.numpy(), g["dE"], rtol=0, atol=1e-5 * scale)
    assert float(eps.grad) == pytest.approx(g["deps"], rel=1e-5, abs=1e-9)
    assert float(sig.grad) == pytest.approx(g["dsigma"], rel=1e-5, abs=1e-9)


def test_autograd_c5_target_rollout_matches_fd():
    w = configs.c5(scale=0.1)
    nsteps, ckpt = 200, 20
    ids = _free_nodes(w)
    L = O.box_f32(w.Lx)
    dt = float(np.float32(w.dt))
    # the target: the same initial state stepped with E (1 + 0.1 field2)
    st = configs.build_system(configs.with_params(w, E=configs.c5_target_E(w)))
    st.run(nsteps)
    xt, yt, _, _ = st.get_state()
    tgt = np.stack([xt[ids], yt[ids]], 1).astype(np.float64)
    st.close()
    # the engine's trajectory, step by step (cell walk: the checkpointed
    # recomputation reproduces it bit for bit)
    s = configs.build_system(w, skin=0.0)
    x0, y0, vx0, vy0 = (a.astype(np.float64) for a in s.get_state())
    tx, ty = np.empty((nsteps + 1, s.N)), np.empty((nsteps + 1, s.N))
    tx[0], ty[0] = x0, y0
    for k in range(nsteps):
        s.run(1)
        x, y, _, _ = s.get_state()
        tx[k + 1], ty[k + 1] = x, y
    s.close()
    E, eps, sig = _params(w)
    pos = FinalPositions.apply(E, eps, sig, w, nsteps, ckpt, 0.0)
    np.testing.assert_array_equal(pos[:, 0].detach().numpy(), tx[-1])
    np.testing.assert_array_equal(pos[:, 1].detach().numpy(), ty[-1])
    loss = mse_min_image(pos, torch.from_numpy(ids), torch.from_numpy(tgt), L)
    loss.backward()
    gE, geps, gsig = E.grad.numpy(), float(eps.grad), float(sig.grad)
    ga = w.global_arrays()
    E0 = ga["E"].astype(np.float64)
    sig0 = ga["sigma"].astype(np.float64)
    eps0 = float(w.eps)
    mass = ga["mass"]
    o = O.from_workload(w, x0, y0)
    res = np.stack([tx[-1][ids], ty[-1][ids]], 1) - tgt
    res -= L * np.round(res / L)
    K = ids.size

    def oracle_final(Ev, ev, sv):
        oo = O.from_workload(w, x0, y0, E=Ev, eps=ev, sigma=sv)
        x, y, _, _, _ = O.run(oo, x0, y0, vx0, vy0, mass, nsteps, dt)
        return x, y

    def loss_of(x, y):
        d = np.stack([x[ids], y[ids]], 1) - tgt
        d -= L * np.round(d / L)
        return float(np.mean((d * d).sum(1)))

    xo, yo = oracle_final(E0, eps0, sig0)
    l0 = loss_of(xo, yo)
    div = np.sqrt(np.mean((xo[ids] - tx[-1][ids]) ** 2 + (yo[ids] - ty[-1][ids]) ** 2))
    rms_res = np.sqrt(np.mean((res * res).sum(1)))
    print(f"[c5 10%] N={w.n_total} elements={E0.size}: loss engine {float(loss):.9e} oracle "
          f"{l0:.9e}; rms |x_fp32 - x_fp64| {div:.3e} against rms residual {rms_res:.3e}",
          flush=True)
    rng = np.random.default_rng(23)
    ndir = 64
    dirs = rng.standard_normal((ndir, E0.size + 2))
    h = 1e-6

    def check(k):
        d = dirs[k]
        dE, de, ds = d[:-2] * E0, d[-2] * eps0, d[-1] * 1.0
        pred = float(np.dot(gE, dE) + geps * de + gsig * ds)
        # FP64 central differences of every step along the engine's trajectory
        dx, dy = O.tangent(o, tx, ty, dt, mass, dE, de, ds, h)
        step_fd = float(2.0 / K * (np.dot(res[:, 0], dx[ids]) + np.dot(res[:, 1], dy[ids])))
        # FP64 central differences of whole FP64 rollouts (a smaller step:
        # the quotient's cutoff-crossing bias is linear in it)
        hr = 1e-7
        lp = loss_of(*oracle_final(E0 + hr * dE, eps0 + hr * de, sig0 + hr * ds))
        lm = loss_of(*oracle_final(E0 - hr * dE, eps0 - hr * de, sig0 - hr * ds))
        return pred, step_fd, (lp - lm) / (2 * hr)

    with cf.ThreadPoolExecutor(16) as ex:
        r = np.array(list(ex.map(check, range(ndir))))
    pred, step_fd, roll_fd = r[:, 0], r[:, 1], r[:, 2]
    e1 = np.abs(pred - step_fd)
    e2 = np.abs(pred - roll_fd)
    r1 = np.linalg.norm(e1) / np.linalg.norm(step_fd)
    r2 = np.linalg.norm(e2) / np.linalg.norm(roll_fd)
    print(f"[c5 10%] 64 directions, along the engine trajectory: norm ratio {r1:.3e}, max "
          f"|err| {e1.max():.3e} of max |fd| {np.abs(step_fd).max():.3e}", flush=True)
    print(f"[c5 10%] 64 directions, whole FP64 rollouts: norm ratio {r2:.3e}, max |err| "
          f"{e2.max():.3e} of max |fd| {np.abs(roll_fd).max():.3e}", flush=True)
    assert r1 <= 1e-4
    assert e1.max() <= 1e-4 * np.abs(step_fd).max()
    # whole rollouts: the FP32 trajectory departs from the FP64 one by `div`
    # against a residual of `rms_res`; the gradient inherits that relative error
    bound = max(1e-4, 20 * div / rms_res)
    assert r2 <= bound
    assert float(loss) == pytest.approx(l0, rel=bound)


def test_adjoint_recording_is_consumed_and_invalidated():
    w = configs.c5(scale=0.01)
    s = configs.build_system(w)
    with pytest.raises(FemmdError) as e:
        s.adjoint_backward(np.zeros(s.N), np.zeros(s.N))
    assert e.value.code == _lib.FM_ERR_STATE
    s.adjoint_forward(10, 0)
    g = s.adjoint_backward(np.ones(s.N), None)
    assert np.isfinite(g["dE"]).all()
    with pytest.raises(FemmdError):  # consumed
        s.adjoint_backward(np.ones(s.N), None)
    s.adjoint_forward(10, 0)
    x, y, vx, vy = s.get_state()
    s.set_state(None, None, vx, vy)  # the state moved (velocities set)
    with pytest.raises(FemmdError):
        s.adjoint_backward(np.ones(s.N), None)
    s.adjoint_forward(10, 0)
    s.run(1)
    with pytest.raises(FemmdError):
        s.adjoint_backward(np.ones(s.N), None)


@pytest.mark.parametrize("skin", [0.6, 0.0])
def test_checkpointed_adjoint_after_history_growth(skin):
    """A checkpointed adjoint, then one storing the whole trajectory (the
    history buffer grows), then the checkpointed one again from the same
    state: the third equals the first bit for bit (the recorded-segment
    graphs follow the history buffer)."""
    w = configs.c5(scale=0.01)
    ids = _free_nodes(w).astype(np.int32)
    s = configs.build_system(w, skin=skin)
    s0 = s.get_state()
    tgt = np.stack([s0[0][ids] + 0.2, s0[1][ids]], 1)
    a = s.adjoint(60, ids, tgt, ckpt=7)
    s.set_state(*s0)
    b = s.adjoint(60, ids, tgt, ckpt=0)
    s.set_state(*s0)
    c = s.adjoint(60, ids, tgt, ckpt=7)
    np.testing.assert_array_equal(a["dE"], c["dE"])
    assert (a["loss"], a["deps"], a["dsigma"]) == (c["loss"], c["deps"], c["dsigma"])
    scale = np.abs(b["dE"]).max()
    assert np.abs(a["dE"] - b["dE"]).max() <= 1e-4 * scale
    assert a["deps"] == pytest.approx(b["deps"], rel=1e-4)
    assert a["dsigma"] == pytest.approx(b["dsigma"], rel=1e-4)
