"""Script front end: the globals a femmd2D run script sees (the Duktape
``System`` object and helper functions of femmd2D/src/interface.cpp:135-227,
644-696), over the engine, so a study written like femmd2D/src/test.js reads
the same in Python::

    from femmd_old_amd.script import System, read_file, FILECONTENT, make_dir
    System.init(L, L)
    System.set_dt(0.001)
    System.set_pair_potential([1.0], [2 ** (1 / 6)], [1.0])
    nodes = read_file("squareVertices.txt", FILECONTENT.FLOAT)
    cells = read_file("squareTriangles.txt", FILECONTENT.INT)
    System.add_elastomer(nodes, cells, x, y, 0, 0, 1.0, M, True)
    System.set_velocities(v0)
    System.allocate_cuda()
    System.run(5000)
    System.print_props()

As in the reference, everything before ``allocate_cuda`` is set-up: the calls
are recorded and the engine system is built from them there, so ``set_box``
and ``set_scale`` may come after ``init`` and before the particles, and
``set_velocities`` applies to the particles added so far.

Departures (each raises or is documented rather than silently different):
  * ``DEV`` and ``init_neighbor_list`` are accepted and do nothing (the
    engine is the device path and builds its cells and lists at commit);
  * ``useNN(flag)`` selects neighbour lists (true) or the per-step cell walk
    (false); the reference's function stores its argument as dt;
  * ``PBC`` accepts only the XY-periodic box; ``set_gravity`` and
    ``set_kArea`` with a non-zero value raise (walls, gravity and area forces
    are not part of this engine);
  * ``set_velocities`` draws its angles from numpy's generator (seed
    ``velocity_seed``), not the reference's RANDOM01;
  * ``set_kBond(k)`` with k > 0 switches bonds on as the reference does
    (femmd2D/src/interface.cpp:455-465): every elastomer then carries the
    bonds ``Particles::addElastomer`` creates on its triangle edges
    (femmd2D/src/particles.cpp:335-337), shared edges twice unless
    ``edge_bonds`` is set to "dedup", and every bond, edge or
    ``create_bond``, has stiffness k (the global kBond of the CPU bond loop,
    femmd2D/src/interactions.cpp:143-173; the last value set before
    ``allocate_cuda`` counts).  With k <= 0 no bond acts, and none is built.
"""
from __future__ import annotations

import json
import math
import os
from typing import Any, Dict, List, Optional

import numpy as np

from . import io as _io
from .io import FILECONTENT, read_file
from .observables import Properties

__all__ = ["System", "FILECONTENT", "DEV", "PBC", "read_file", "write_to_file",
           "append_to_file", "make_dir", "change_dir", "ScriptSystem"]


class DEV:
    CPU = 0
    GPU = 1


class PBC:
    NONE = 0
    X = 1
    XY = 2


def write_to_file(name: str, text: str):
    """femmd2D/src/interface.cpp:105-117."""
    with open(name, "w") as f:
        f.write(text)


def append_to_file(name: str, text: str):
    """femmd2D/src/interface.cpp:119-133."""
    with open(name, "a") as f:
        f.write(text)


def make_dir(name: str):
    """femmd2D/src/interface.cpp:197-213 (an existing directory is fine)."""
    os.makedirs(name, exist_ok=True)


def change_dir(name: str):
    """femmd2D/src/interface.cpp:215-227."""
    os.chdir(name)


def _node_masses(verts: np.ndarray, tris: np.ndarray, total: float, node_mass: float):
    """The masses fm_add_elastomer gives the nodes: node_mass each, or the
    total spread as rho * A / 3 per corner (femmd2D/src/particles.cpp:
    340-366)."""
    nv = verts.shape[0]
    if node_mass > 0:
        return np.full(nv, node_mass)
    p0, p1, p2 = verts[tris[:, 0]], verts[tris[:, 1]], verts[tris[:, 2]]
    a = 0.5 * ((p2[:, 1] - p0[:, 1]) * (p1[:, 0] - p0[:, 0])
               - (p1[:, 1] - p0[:, 1]) * (p2[:, 0] - p0[:, 0]))
    m = np.zeros(nv)
    rho = total / a.sum()
    for k in range(3):
        np.add.at(m, tris[:, k], rho * a / 3.0)
    return m


class ScriptSystem:
    """The ``System`` global.  ``engine`` (after allocate_cuda) is the
    femmd_old_amd.System doing the work."""

    _PROP_KEYS = ("uKinetic", "uPotential", "uFea", "uBonded", "uTotal", "pressure", "virial",
                  "feaVirial")

    def __init__(self, engine_factory=None):
        self._factory = engine_factory
        self.velocity_seed = 1234
        self.sample_every = 1
        self.edge_bonds = "all"  # or "dedup": each shared triangle edge bonded once
        self.reset()

    # -- set-up -------------------------------------------------------------
    def reset(self):
        """femmd2D/src/md3dsystem.cpp:82-90."""
        self.engine = None
        self._box = None
        self._dt = 0.005
        self._pair = None
        self._E, self._nu = None, 0.3
        self._kbond = 0.0
        self._skin = None
        self._ntypes = 1
        self._scale = 1.0
        self._parts: List[Dict[str, np.ndarray]] = []  # particle blocks, caller order
        self._elast: List[tuple] = []                  # (block index, args)
        self._bonds: List[tuple] = []
        self.props: Optional[Properties] = None
        self._step_equi, self._step_avg = 10000, 1000
        self._avg_count_rdf, self._range_rdf = 30000, None
        self._traj, self._traj_steps = False, 1

    def init(self, Lx: float, Ly: float):
        """System.init(Lx, Ly) (femmd2D/src/md3dsystem.cpp:49-80)."""
        self.reset()
        self._box = (float(Lx), float(Ly))

    def _need_setup(self):
        if self.engine is not None:
            raise RuntimeError("the system is allocated; set-up calls come before allocate_cuda")
        if self._box is None:
            raise RuntimeError("call System.init(Lx, Ly) first")

    def set_box(self, Lx: float, Ly: float):
        self._need_setup()
        self._box = (float(Lx), float(Ly))

    def get_box(self) -> dict:
        return {"x": self._box[0], "y": self._box[1]}

    def set_dt(self, dt: float):
        self._dt = float(dt)
        if self.engine is not None:
            self.engine.set_dt(dt)

    def get_dt(self) -> float:
        return self._dt

    def set_scale(self, s: float):
        self._scale = float(s)

    def set_numTypes(self, n: int):
        self._ntypes = int(n)

    def DEV(self, dev: int):
        pass

    def PBC(self, kind: int):
        if int(kind) != PBC.XY:
            raise NotImplementedError("only the XY-periodic box is supported")

    def useNN(self, flag):
        self._skin = None if flag else 0.0

    def set_gravity(self, g: float):
        if g:
            raise NotImplementedError("gravity is not part of this engine")

    def set_kArea(self, k: float):
        if k:
            raise NotImplementedError("area forces are not part of this engine")

    def set_kBond(self, k: float):
        self._kbond = float(k)

    def set_youngs_modulus(self, E: float, nu: float = 0.3):
        self._E, self._nu = float(E), float(nu)

    def set_pair_potential(self, eps, rc, us):
        self._pair = (np.ravel(eps), np.ravel(rc), np.ravel(us))

    def add_particle(self, x, y, vx, vy, m, sig, typ=0):
        self._need_setup()
        self._parts.append(dict(x=np.array([x], float), y=np.array([y], float),
                                vx=np.array([vx], float), vy=np.array([vy], float),
                                m=np.array([m], float), sig=np.array([sig], float),
                                typ=np.array([typ], np.int32)))

    def add_elastomer(self, nodes, cells, x, y, typ=0, g=0, sig=1.0, mass=0.0, exclude=True):
        """System.add_elastomer (femmd2D/src/interface.cpp:533-563): nodes
        flat (x, y) pairs relative to (x, y), cells 1-based, mass the total."""
        self._need_setup()
        v = np.asarray(nodes, np.float32).reshape(-1, 2)
        t = np.asarray(cells, np.int32).reshape(-1, 3)
        px = (v[:, 0] + np.float32(x)).astype(np.float64)
        py = (v[:, 1] + np.float32(y)).astype(np.float64)
        nv = v.shape[0]
        m = _node_masses(v.astype(np.float64), t - 1, float(mass), 0.0)
        self._parts.append(dict(x=px, y=py, vx=np.zeros(nv), vy=np.zeros(nv), m=m,
                                sig=np.full(nv, float(sig)), typ=np.full(nv, int(typ), np.int32)))
        self._elast.append((len(self._parts) - 1,
                            (v.ravel(), t.ravel(), float(x), float(y), int(typ), int(g),
                             float(sig), float(mass), bool(exclude))))

    def create_bond(self, i: int, j: int):
        self._bonds.append((int(i), int(j)))

    def set_velocities(self, v0: float):
        """Particles::setParticlesVel (femmd2D/src/particles.cpp:233-261):
        speed v0 in a random direction for every particle so far, then the
        mass-weighted mean velocity removed."""
        self._need_setup()
        rng = np.random.default_rng(self.velocity_seed)
        n = sum(len(b["x"]) for b in self._parts)
        s = 2.0 * math.pi * rng.random(n)
        vx, vy = np.cos(s) * v0, np.sin(s) * v0
        m = np.concatenate([b["m"] for b in self._parts]) if n else np.zeros(0)
        M = m.sum()
        if M > 0:
            vx -= np.sum(vx * m) / M
            vy -= np.sum(vy * m) / M
        k = 0
        for b in self._parts:
            nb = len(b["x"])
            b["vx"], b["vy"] = vx[k:k + nb].copy(), vy[k:k + nb].copy()
            k += nb

    # -- allocation ---------------------------------------------------------
    def allocate_cuda(self):
        """System.allocate_cuda: the engine system is built from the set-up."""
        self._need_setup()
        factory = self._factory
        if factory is None:
            from . import System as factory
        s = factory(self._box[0], self._box[1], dt=self._dt)
        if self._skin is not None:
            s.set_skin(self._skin)
        if self._pair is not None:
            s.set_pair_potential(*self._pair)
        if self.edge_bonds not in ("all", "dedup"):
            raise ValueError("edge_bonds must be 'all' or 'dedup'")
        kb = self._kbond
        edges = self.edge_bonds if kb > 0 else None
        el = dict(self._elast)
        for k, b in enumerate(self._parts):
            if k in el:
                a = el[k]
                s.add_elastomer(a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8],
                                edge_bonds=edges, bond_k=kb)
            else:
                s.add_particles(b["x"], b["y"], b["vx"], b["vy"], b["m"], b["sig"], b["typ"])
        if self._E is not None:
            s.set_youngs_modulus(self._E, self._nu)
        if self._bonds and kb > 0:
            i, j = zip(*self._bonds)
            s.add_bonds(list(i), list(j), [kb] * len(i))
        vx = np.concatenate([b["vx"] for b in self._parts])
        vy = np.concatenate([b["vy"] for b in self._parts])
        s.set_state(None, None, vx, vy)
        s.allocate_cuda()
        self.engine = s
        self.props = Properties(s, step_avg=self._step_avg, step_equi=self._step_equi,
                                avg_count_rdf=self._avg_count_rdf)

    def init_neighbor_list(self):
        pass

    # -- running ------------------------------------------------------------
    def _need_engine(self):
        if self.engine is None:
            raise RuntimeError("call System.allocate_cuda() first")
        return self.engine

    def run(self, steps: int):
        """System.run (femmd2D/src/integrator.cpp:31-57): the properties are
        computed every ``sample_every`` steps (1 = every step, as the
        reference), the cluster trajectory appended every
        set_trajectory_steps steps when save_trajectory is on."""
        s = self._need_engine()
        k = max(1, int(self.sample_every))
        if self._traj:
            k = math.gcd(k, max(1, self._traj_steps))
        done = 0
        while done < steps:
            m = min(k, steps - done)
            s.run(m)
            done += m
            st = s.energies()["steps"]
            if st % max(1, int(self.sample_every)) == 0 or done == steps:
                self.props.compute()
            if self._traj and st % self._traj_steps == 0:
                _io.save_cluster_xyz(s, "traj.xyz", 1)

    def save_trajectory(self, flag: bool):
        self._traj = bool(flag)

    def set_trajectory_steps(self, n: int):
        self._traj_steps = max(1, int(n))

    # -- properties (femmd2D/src/properties.cpp, interface.cpp:266-400) ------
    def set_equilibrium_steps(self, n: int):
        self._step_equi = int(n)
        if self.props:
            self.props.step_equi = int(n)

    def set_average_average_steps(self, n: int):
        self._step_avg = int(n)
        if self.props:
            self.props.step_avg = int(n)

    def set_average_steps_rdf(self, n: int):
        self._avg_count_rdf = int(n)
        if self.props:
            self.props.avg_count_rdf = int(n)

    def set_range_rdf(self, r: float):
        self._range_rdf = float(r)

    def set_reset_properties(self, *_):
        if self.props:
            self.props.reset()

    def get_props(self) -> Dict[str, Any]:
        """System.get_props (femmd2D/src/interface.cpp:627-644)."""
        s = self._need_engine()
        p = self.props
        out: Dict[str, Any] = {"steps": s.energies()["steps"]}
        for k in self._PROP_KEYS:
            q = p.props[k]
            out[k] = {"avgval": q.avgval, "curr": q.val, "stdv": q.stdv}
        return out

    def print_props(self):
        """System.print_props (femmd2D/src/interface.cpp:272-286)."""
        p = self.props.props
        st = self._need_engine().energies()["steps"]
        print("%d %f %f %f %f %f %f\n%f %f %f" % (
            st, p["uKinetic"].avgval, p["uPotential"].avgval, p["uTotal"].avgval,
            p["virial"].avgval, p["feaVirial"].avgval, p["pressure"].avgval,
            p["clusterUKinetic"].avgval, p["clusterVirial"].avgval, p["molPressure"].avgval))

    def info(self) -> Dict[str, Any]:
        """System.info (femmd2D/src/interface.cpp:565-598)."""
        n = sum(len(b["x"]) for b in self._parts)
        return {"N": n, "nelastomers": len(self._elast), "nbonds": len(self._bonds),
                "box": self.get_box(), "dt": self._dt}

    def calc_cluster_props(self):
        return self._need_engine().cluster_props()

    # -- files --------------------------------------------------------------
    def save_vtk(self, prefix: str):
        return _io.save_vtk(self._need_engine(), prefix)

    def save_vtf(self, name: str, s: int = 0):
        _io.save_vtf(self._need_engine(), name, s)

    def save_cluster_vtf(self, name: str, s: int = 0):
        _io.save_cluster_vtf(self._need_engine(), name, s, self._scale)

    def save_cluster_xyz(self, name: str, s: int = 0):
        _io.save_cluster_xyz(self._need_engine(), name, s)

    def to_json(self) -> str:
        return json.dumps(self.info(), indent=4)


System = ScriptSystem()
