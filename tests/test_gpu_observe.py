"""Per-step observables formed on the device (Integrator::doStep ->
Properties::compute, femmd2D/src/integrator.cpp:18, properties.cpp:15-98).

Every step of a run forms its energies, virials, kinetic energy and
pressure and adds the averaged quantities (uKinetic, uPotential, uTotal,
pressure, virial, feaVirial, uBonded, uFea) to running sums, as the
reference's Properties::accumulate(1) does.  The device sums are checked
against the same quantities queried after each of n single steps
(fm_get_energies, whose values the parity tests pin to the FP64 oracle),
summed on the host in step order: the same numbers, so they agree to the
last bits.  A run that forms the observables only at its end follows the
same trajectory bit for bit.
"""
import numpy as np
import pytest

from femmd_old_amd import Batch, configs
from femmd_old_amd._lib import AVG_NAMES

pytestmark = pytest.mark.gpu

SKINS = [0.6, 0.0]


def derived(e, n, V):
    """The averaged quantities of one step (femmd2D/src/properties.cpp:18-34)."""
    uk = e["ke"] / n
    up = (e["u_lj"] + e["u_fea"] + e["u_bond"]) / n
    return np.array([uk, up, uk + up, e["pressure"], e["w_lj"], e["w_fea"], e["u_bond"] / n,
                     e["u_fea"] / n])


def host_averages(vals):
    v = np.asarray(vals)
    m = v.mean(axis=0)
    sd = np.sqrt(np.maximum((v * v).mean(axis=0) - m * m, 0.0))
    return m, sd


@pytest.mark.parametrize("skin", SKINS)
@pytest.mark.parametrize("cfg", ["c1", "c3s"])
def test_per_step_averages_equal_stepwise_queries(cfg, skin):
    w = {"c1": lambda: configs.c1(), "c3s": lambda: configs.c3(scale=0.005)}[cfg]()
    n_steps = 37
    a = configs.build_system(w, skin=skin)
    b = configs.build_system(w, skin=skin)
    a.run(n_steps)
    avg = a.averages()
    V = float(np.float32(w.Lx)) * float(np.float32(w.Ly))
    vals = []
    for _ in range(n_steps):
        b.run(1)
        vals.append(derived(b.energies(), w.n_total, V))
    for u, v in zip(a.get_state(), b.get_state()):
        np.testing.assert_array_equal(u, v)
    m, sd = host_averages(vals)
    assert avg["count"] == n_steps and b.averages()["count"] == n_steps
    for k, name in enumerate(AVG_NAMES):
        print(f"[{cfg} skin {skin}] {name}: device {avg[name]['avg']:.12e} "
              f"+- {avg[name]['stdv']:.3e}, host {m[k]:.12e} +- {sd[k]:.3e}")
        assert avg[name]["avg"] == pytest.approx(m[k], rel=1e-12, abs=1e-14)
        assert avg[name]["stdv"] == pytest.approx(sd[k], rel=1e-6, abs=1e-9 * (abs(m[k]) + 1))
    # the last step's totals are those the stepwise query reports
    ea, eb = a.energies(), b.energies()
    for k in ("ke", "u_lj", "w_lj", "u_fea", "w_fea", "u_bond", "w_bond", "pressure", "vv_max"):
        assert ea[k] == eb[k], k
    # Properties::accumulate(0)
    a.reset_averages()
    assert a.averages()["count"] == 0
    a.run(5)
    assert a.averages()["count"] == 5


@pytest.mark.parametrize("skin", SKINS)
def test_observe_at_run_end_only_same_trajectory(skin):
    w = configs.c2(scale=0.005)
    a = configs.build_system(w, skin=skin)
    b = configs.build_system(w, commit=False, skin=skin)
    b.set_observe_every(0)
    b.allocate_cuda()
    for _ in range(3):
        a.run(25)
        b.run(25)
    for u, v in zip(a.get_state(), b.get_state()):
        np.testing.assert_array_equal(u, v)
    ea, eb = a.energies(), b.energies()
    for k in ("ke", "u_lj", "w_lj", "u_fea", "w_fea", "pressure"):
        assert ea[k] == eb[k], k
    assert a.averages()["count"] == 75
    assert b.averages()["count"] == 3  # one sample per run


def test_velocities_set_alone_refresh_kinetic_energy():
    w = configs.c1()
    s = configs.build_system(w)
    s.run(10)
    e0 = s.energies()
    assert e0["ke"] > 0
    x, y, vx, vy = s.get_state()
    s.set_state(None, None, np.zeros_like(vx), np.zeros_like(vy))
    e = s.energies()
    assert e["ke"] == 0.0 and e["vv_max"] == 0.0 and e["temperature"] == 0.0
    # the configuration did not change: potential energy and virials stay
    for k in ("u_lj", "w_lj", "u_fea", "w_fea"):
        assert e[k] == e0[k], k
    V = float(np.float32(w.Lx)) * float(np.float32(w.Ly))
    assert e["pressure"] == pytest.approx(0.5 * (e["w_lj"] + e["w_fea"]) / V, rel=1e-12)
    # and a rescale to a temperature shows in the next query
    s.set_state(None, None, vx, vy)
    s.set_temperature(2.0)
    assert s.energies()["temperature"] == pytest.approx(2.0, rel=1e-5)


@pytest.mark.parametrize("skin", SKINS)
def test_batch_averages_equal_single_systems(skin):
    ws = configs.c4(n_replicas=6, n_max=3000)
    singles = []
    for w in ws:
        s = configs.build_system(w, skin=skin)
        s.run(30)
        singles.append(s.averages())
    b = Batch([configs.build_system(w, commit=False, skin=skin) for w in ws])
    b.run(30)
    for r, sa in enumerate(singles):
        ba = b.averages(r)
        assert ba["count"] == sa["count"] == 30
        for name in AVG_NAMES:
            assert ba[name]["avg"] == pytest.approx(sa[name]["avg"], rel=1e-12, abs=1e-14), name
