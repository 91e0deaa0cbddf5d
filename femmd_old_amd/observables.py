"""Observables layer over a ``System``: the reference's per-step properties,
their running averages and the cluster radial distribution function
(femmd2D/src/properties.h:6-73, femmd2D/src/properties.cpp:15-158).

The per-configuration values come from the engine (FP64 device sums):
energies and virials (``fm_get_energies``), cluster centres of mass and
velocities (``fm_cluster_props``) and the cluster virial
(``fm_cluster_virial``).  The running averages and the histogram are host
bookkeeping, as in the reference.  The reference evaluates them after every
step; ``Properties.run`` samples every ``sample_every`` steps (1 = the
reference's cadence), because each sample is a query that synchronises with
the device.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Dict, Optional

import numpy as np

__all__ = ["Prop", "Properties"]


@dataclass
class Prop:
    """One averaged quantity (femmd2D/src/properties.h:6-13)."""
    val: float = 0.0
    sum: float = 0.0
    sum2: float = 0.0
    avgval: float = 0.0
    stdv: float = 0.0

    def zero(self):
        self.val = self.sum = self.sum2 = 0.0

    def accum(self):
        self.sum += self.val
        self.sum2 += self.val * self.val

    def avg(self, n: int):
        self.avgval = self.sum / n
        self.stdv = math.sqrt(max(self.sum2 / n - self.avgval * self.avgval, 0.0))


_NAMES = ("temperature", "uPotential", "uKinetic", "clusterUKinetic", "uTotal", "uBonded",
          "uFea", "pressure", "molPressure", "virial", "feaVirial", "clusterVirial")


@dataclass
class Properties:
    """femmd2D/src/properties.h:15-73 over an engine ``System``.

    step_avg, step_equi, skip_steps_rdf, size_hist_rdf, avg_count_rdf keep
    the reference's defaults.  The RDF range is half the box width, as
    ``evalClusterRdf`` sets it."""
    sys: object
    step_avg: int = 1000
    step_equi: int = 10000
    skip_steps_rdf: int = 10
    size_hist_rdf: int = 100
    avg_count_rdf: int = 30000
    count: int = 0
    props: Dict[str, Prop] = field(default_factory=lambda: {k: Prop() for k in _NAMES})
    hist_rdf: Optional[np.ndarray] = None
    count_rdf: int = 0
    last_rdf: Optional[tuple] = None

    def __getattr__(self, name):
        p = self.__dict__.get("props")
        if p is not None and name in p:
            return p[name]
        raise AttributeError(name)

    # -- femmd2D/src/properties.cpp:15-36 -----------------------------------
    def evaluate(self) -> Dict[str, float]:
        s = self.sys
        e = s.energies()
        n = e["n"]
        box = s.get_box()
        V = box["x"] * box["y"]
        P = self.props
        P["uKinetic"].val = e["ke"] / n
        P["uPotential"].val = (e["u_lj"] + e["u_bond"] + e["u_fea"]) / n
        P["uTotal"].val = P["uKinetic"].val + P["uPotential"].val
        P["temperature"].val = P["uKinetic"].val
        P["uFea"].val = e["u_fea"] / n
        P["uBonded"].val = e["u_bond"] / n
        P["virial"].val = e["w_lj"]
        P["feaVirial"].val = e["w_fea"]
        P["pressure"].val = e["pressure"]
        nc = s.cluster_count()
        if nc > 0:
            c = s.cluster_props()
            cke = float(np.sum(c["mass"] * (c["vcm"][:, 0] ** 2 + c["vcm"][:, 1] ** 2)))
            cvir = s.cluster_virial()
            P["clusterUKinetic"].val = 0.5 * cke / nc
            P["clusterVirial"].val = cvir
            P["molPressure"].val = 0.5 * cke / V - 0.5 * cvir / V
            self._cm = c["cm"]
        else:
            self._cm = None
        return {k: p.val for k, p in P.items()}

    # -- femmd2D/src/properties.cpp:39-86 -----------------------------------
    def accumulate(self, icode: int):
        if icode == 0:
            self.count = 0
            for p in self.props.values():
                p.zero()
        elif icode == 1:
            self.count += 1
            for p in self.props.values():
                p.accum()
        elif icode == 2:
            for p in self.props.values():
                p.avg(max(self.count, 1))

    # -- femmd2D/src/properties.cpp:88-98 -----------------------------------
    def compute(self):
        self.evaluate()
        self.accumulate(1)
        steps = self.sys.energies()["steps"]
        if steps > self.step_equi:
            self.eval_cluster_rdf(steps)
            if steps % self.step_avg == 0:
                self.accumulate(2)
                self.accumulate(0)

    def run(self, steps: int, sample_every: int = 1):
        """Advance ``steps`` steps, computing the properties every
        ``sample_every`` steps (the reference computes them every step)."""
        k = max(1, int(sample_every))
        done = 0
        while done < steps:
            m = min(k, steps - done)
            self.sys.run(m)
            done += m
            self.compute()

    # -- femmd2D/src/properties.cpp:104-158 ---------------------------------
    def eval_cluster_rdf(self, steps: int):
        """Histogram of cluster centre-of-mass separations (minimum image)
        over the range L/2, normalised after ``avg_count_rdf`` samples.  The
        2D shell normalisation uses the shell's mid radius (n + 1/2) dr; the
        reference divides by (n - 1/2), which is negative for the first
        shell."""
        if steps % self.skip_steps_rdf != 0 or self._cm is None:
            return
        cm = np.asarray(self._cm, dtype=np.float64)
        nc = len(cm)
        box = self.sys.get_box()
        Lx, Ly = box["x"], box["y"]
        rng = Lx / 2.0
        nh = self.size_hist_rdf
        if self.count_rdf == 0 or self.hist_rdf is None:
            self.hist_rdf = np.zeros(nh)
        dR = rng / nh
        if nc > 1:
            i, j = np.triu_indices(nc, 1)
            dx = cm[i, 0] - cm[j, 0]
            dy = cm[i, 1] - cm[j, 1]
            dx -= Lx * np.round(dx / Lx)
            dy -= Ly * np.round(dy / Ly)
            rr = dx * dx + dy * dy
            sel = rr < rng * rng
            idx = np.minimum((np.sqrt(rr[sel]) / dR).astype(np.int64), nh - 1)
            np.add.at(self.hist_rdf, idx, 1.0)
        self.count_rdf += 1
        if self.count_rdf == self.avg_count_rdf:
            norm = Lx * Ly / (math.pi * dR * dR * nc * nc * self.count_rdf)
            n = np.arange(nh)
            g = self.hist_rdf * norm / (n + 0.5)
            rb = (n + 0.5) * rng / nh
            self.last_rdf = (rb, g)
            self.count_rdf = 0

    def reset(self):
        self.count = 0
        self.count_rdf = 0
