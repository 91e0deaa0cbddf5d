"""Observables layer on the CPU: the oracle's cluster properties and cluster
virial pinned by closed forms and an independent implementation, and the
host bookkeeping of femmd_old_amd.observables (running averages, cluster
RDF) on a stand-in system.

  * unfolding: a disk placed across the periodic edge (either side) has its
    unfolded centre of mass at the disk centre (femmd2D/src/feaElements.cpp:
    86-133), its circular-mean centroid there too (the disk is symmetric),
    and a uniform velocity is its centre-of-mass velocity;
  * cluster virial: with every particle its own cluster (centre = position)
    the sum of 1/2 f_ij . (R_ci - R_cj) is the pair virial W_lj, which the C
    oracle forms separately (or_forces); with one cluster holding every
    particle it is 0;
  * Prop.avg is the mean and population standard deviation; the RDF of
    uncorrelated centres is 1 within noise (2D shell normalisation).
"""
import numpy as np
import pytest

import oracle as O
from femmd_old_amd import configs
from femmd_old_amd.observables import Prop, Properties


def _disk_at(cx, cy, L=60.0, r=8.0, v=(0.3, -0.2)):
    verts, _ = configs.hex_disk(r, 1.0)
    x = np.mod(verts[:, 0].astype(np.float32) + np.float32(cx), np.float32(L))
    y = np.mod(verts[:, 1].astype(np.float32) + np.float32(cy), np.float32(L))
    n = len(x)
    vx = np.full(n, v[0], np.float32)
    vy = np.full(n, v[1], np.float32)
    return x, y, vx, vy, np.ones(n), L


@pytest.mark.parametrize("cx,cy", [(0.5, 30.0), (59.5, 30.0), (30.0, 0.25), (0.75, 59.25),
                                   (30.0, 30.0)])
def test_cluster_unfold_closed_form(cx, cy):
    x, y, vx, vy, m, L = _disk_at(cx, cy)
    out, ux, uy = O.cluster_props(x, y, vx, vy, m, [(0, len(x))], L, L)
    M, cmx, cmy, vcx, vcy, gx, gy = out[0]
    assert M == len(x)
    # unfolded centre of mass = the disk centre (mod L), the folding leaves
    # the cluster contiguous
    assert abs(cmx - cx) < 1e-4 and abs(cmy - cy) < 1e-4, (cmx, cmy)
    assert ux.max() - ux.min() < 2 * 8.0 + 1e-3 and uy.max() - uy.min() < 2 * 8.0 + 1e-3
    # circular mean of a symmetric cluster = its centre
    assert abs(gx - cx) < 1e-4 and abs(gy - cy) < 1e-4, (gx, gy)
    assert abs(vcx - 0.3) < 1e-7 and abs(vcy + 0.2) < 1e-7


def test_cluster_virial_identity_and_zero():
    w = configs.c1(with_disk=False)
    n = 600
    x, y = w.x[:n], w.y[:n]
    sig = np.ones(n)
    excl = -np.ones(n, np.int32)
    # every particle its own cluster: the pair virial
    cid = np.arange(n)
    cm = np.stack([x.astype(np.float32), y.astype(np.float32)], 1).astype(np.float64)
    v, ab = O.cluster_virial(x, y, sig, excl, cid, cm, w.Lx, w.Ly, w.eps, w.rc)
    # the same n particles in the C oracle
    sub = configs.Workload("sub", w.Lx, w.Ly, w.dt, x, y, w.vx[:n], w.vy[:n], w.mass[:n],
                           w.sigma[:n], [], eps=w.eps, rc=w.rc, ushift=w.ushift)
    o = O.from_workload(sub)
    _, _, en = O.forces(o, pred_f32=True)
    assert ab > 0
    assert abs(v - en[O.OR_W_LJ]) <= 1e-9 * ab, (v, en[O.OR_W_LJ])
    # a single cluster: every R difference is 0
    v1, _ = O.cluster_virial(x, y, sig, excl, np.zeros(n, int), cm[:1], w.Lx, w.Ly, w.eps, w.rc)
    assert v1 == 0.0


def test_prop_mean_and_stdv():
    p = Prop()
    vals = [1.0, 2.0, 4.0, 7.0]
    for v in vals:
        p.val = v
        p.accum()
    p.avg(len(vals))
    assert p.avgval == pytest.approx(np.mean(vals))
    assert p.stdv == pytest.approx(np.std(vals))


class _Stand:
    """A stand-in System: fixed energies, cluster centres drawn at random."""

    def __init__(self, nc=400, L=100.0, seed=0):
        self.rng = np.random.default_rng(seed)
        self.nc, self.L, self.steps = nc, L, 0

    def run(self, m):
        self.steps += m

    def get_box(self):
        return {"x": self.L, "y": self.L}

    def energies(self):
        return {"n": 10, "ke": 10.0, "u_lj": -20.0, "u_bond": 0.0, "u_fea": 5.0, "w_lj": 3.0,
                "w_fea": 1.0, "pressure": 0.5, "steps": self.steps}

    def cluster_count(self):
        return self.nc

    def cluster_props(self):
        cm = self.rng.uniform(0, self.L, (self.nc, 2))
        return {"mass": np.ones(self.nc), "cm": cm, "vcm": np.zeros((self.nc, 2)),
                "centroid": cm}

    def cluster_virial(self):
        return 2.0


def test_properties_bookkeeping_and_rdf():
    s = _Stand()
    p = Properties(s, step_avg=50, step_equi=0, skip_steps_rdf=1, size_hist_rdf=20,
                   avg_count_rdf=200)
    p.run(200, sample_every=1)
    # constant inputs: averages equal the values, zero spread
    assert p.uKinetic.avgval == pytest.approx(1.0)
    assert p.uPotential.avgval == pytest.approx(-1.5)
    assert p.clusterVirial.avgval == pytest.approx(2.0)
    assert p.molPressure.avgval == pytest.approx(-0.5 * 2.0 / 1e4)
    assert p.pressure.stdv == pytest.approx(0.0, abs=1e-12)
    # uncorrelated centres: g(r) = 1 (shells past the first few carry
    # enough pairs for a 10 % bar)
    rb, g = p.last_rdf
    assert rb.shape == g.shape == (20,)
    assert np.all(np.abs(g[3:] - 1.0) < 0.1), g
