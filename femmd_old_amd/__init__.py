"""femmd_old_amd -- MI355X-native engine for femmd2D's per-step
force-and-integrate loop.

The compute path is libfemmd.so: hand-written gfx950 HIP kernels behind the
C ABI of include/femmd.h.  This module is a thin ctypes front end whose
``System`` mirrors the reference's script API (the ``System.*`` table of
femmd2D/src/interface.cpp:656-696 and the c_* functions of :24-81): the same
method names, argument meaning and order, so reference-style run scripts
(femmd2D/src/test.js) translate line by line.  Errors raise ``FemmdError``
(the reference exits the process instead, femmd2D/src/CudaUtil.h:4-14).
"""
from __future__ import annotations

import ctypes as C
import math
from typing import Iterable, Optional, Sequence

import numpy as np

from . import _lib
from ._lib import (FM_LEGACY_GPU, FM_NEO_HOOKEAN, FM_SVK, FmAverages, FmEnergies, FmParams)

__all__ = ["System", "Batch", "FemmdError", "FM_NEO_HOOKEAN", "FM_SVK", "FM_LEGACY_GPU"]


class FemmdError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def _fp(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_float))


def _ip(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_int32))


def _dp(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


def _f32(a, n=None) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    if n is not None and a.shape[0] != n:
        raise ValueError(f"expected {n} values, got {a.shape[0]}")
    return a


def _i32(a, n=None) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(a, dtype=np.int32))
    if n is not None and a.shape[0] != n:
        raise ValueError(f"expected {n} values, got {a.shape[0]}")
    return a


class System:
    """One periodic 2D system (the reference's MDSystem).

    ``System(Lx, Ly)`` is ``System.init(Lx, Ly)`` of the reference
    (femmd2D/src/md3dsystem.cpp:49-80): XY-periodic box, dt = 0.005, WCA pair
    potential (eps 1, rc 2^(1/6), shift 1) until set_pair_potential.
    """

    def __init__(self, Lx: float, Ly: float, dt: float = 0.005, device: int = 0):
        self._lib = _lib.load()
        p = FmParams(Lx=float(Lx), Ly=float(Ly), dt=float(dt), device=int(device))
        h = C.c_void_p()
        self._check(self._lib.fm_create(C.byref(p), C.byref(h)), h=None)
        self._h = h
        self._box = (float(Lx), float(Ly))
        self._dt = float(dt)
        self._kbond = 0.0
        self._committed = False
        self._nelem = 0
        # caller-order records for the file writers (femmd_old_amd.io)
        self._rec_typ, self._rec_sig, self._rec_tri, self._rec_bond = [], [], [], []

    # -- plumbing -------------------------------------------------------------
    def _check(self, rc: int, h="self"):
        if rc != _lib.FM_OK:
            hh = self._h if h == "self" else h
            msg = self._lib.fm_last_error(hh).decode() if hh else "invalid argument"
            raise FemmdError(rc, msg)

    def close(self):
        if getattr(self, "_h", None):
            self._lib.fm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    @property
    def N(self) -> int:
        n = C.c_int32()
        self._check(self._lib.fm_num_particles(self._h, C.byref(n)))
        return n.value

    # -- construction (reference System.* names) -----------------------------
    def add_particle(self, x, y, vx, vy, m, sig, typ=0) -> int:
        """System.add_particle(x, y, vx, vy, m, sig, typ)
        (femmd2D/src/interface.cpp:243-254)."""
        return self.add_particles([x], [y], [vx], [vy], [m], [sig], [typ])

    def add_particles(self, x, y, vx=None, vy=None, m=1.0, sig=1.0, typ=None) -> int:
        x = _f32(x)
        n = x.shape[0]
        y = _f32(y, n)
        vx = None if vx is None else _f32(vx, n)
        vy = None if vy is None else _f32(vy, n)
        m = _f32(np.broadcast_to(np.asarray(m, np.float32), (n,)))
        sig = _f32(np.broadcast_to(np.asarray(sig, np.float32), (n,)))
        typ = None if typ is None else _i32(np.broadcast_to(np.asarray(typ, np.int32), (n,)))
        first = C.c_int32()
        self._check(self._lib.fm_add_particles(self._h, n, _fp(x), _fp(y), _fp(vx), _fp(vy),
                                               _fp(m), _fp(sig), _ip(typ), C.byref(first)))
        self._rec_typ.append(np.zeros(n, np.int32) if typ is None else typ.copy())
        self._rec_sig.append(sig.copy())
        return first.value

    def add_elastomer(self, nodes, cells, x, y, typ=0, g=0, sig=1.0, mass=0.0, exclude=True,
                      *, index_base=1, node_mass=0.0, E=None, nu=None, edge_bonds=None,
                      bond_k=0.0) -> int:
        """System.add_elastomer(nodes, cells, x, y, typ, g, sig, mass, exclude)
        (femmd2D/src/interface.cpp:533-563, femmd2D/src/particles.cpp:295-385).

        nodes: flat (2*nv) vertex coordinates relative to (x, y); cells: flat
        (3*nt) triangle indices, 1-based like the reference mesh files unless
        index_base says otherwise; mass: total mass spread as rho*A/3 per
        corner, or node_mass per node when given.  g (the reference's group
        id) only tags particles for its group observables and is ignored.
        E / nu: optional per-triangle arrays.  edge_bonds: None (no bonds),
        "all" (the reference's edge bonds, femmd2D/src/particles.cpp:335-337:
        every triangle's three edges, shared edges twice) or "dedup" (each
        edge once), stiffness bond_k, r0 the current length.  Returns the
        first node index.
        """
        if edge_bonds not in (None, "all", "dedup"):
            raise ValueError("edge_bonds must be None, 'all' or 'dedup'")
        v = _f32(np.asarray(nodes, np.float32).reshape(-1))
        t = _i32(np.asarray(cells, np.int32).reshape(-1))
        nv, nt = v.shape[0] // 2, t.shape[0] // 3
        Ea = None if E is None else _f32(np.broadcast_to(np.asarray(E, np.float32), (nt,)))
        nua = None if nu is None else _f32(np.broadcast_to(np.asarray(nu, np.float32), (nt,)))
        first = C.c_int32()
        self._check(self._lib.fm_add_elastomer(
            self._h, nv, _fp(v), nt, _ip(t), int(index_base), float(x), float(y), int(typ),
            float(sig), float(node_mass), float(mass), int(bool(exclude)), _fp(Ea), _fp(nua),
            C.byref(first)))
        self._nelem += nt
        self._rec_typ.append(np.full(nv, int(typ), np.int32))
        self._rec_sig.append(np.full(nv, float(sig), np.float32))
        self._rec_tri.append(t.reshape(-1, 3) - int(index_base) + first.value)
        if edge_bonds is not None:
            self._check(self._lib.fm_add_edge_bonds(self._h, first.value, nt, _ip(t),
                                                    int(index_base), float(bond_k),
                                                    int(edge_bonds == "dedup")))
        return first.value

    @property
    def n_elements(self) -> int:
        return self._nelem

    def particle_types(self) -> np.ndarray:
        """Type of every particle, caller order."""
        return np.concatenate(self._rec_typ) if self._rec_typ else np.zeros(0, np.int32)

    def particle_sigmas(self) -> np.ndarray:
        return np.concatenate(self._rec_sig) if self._rec_sig else np.zeros(0, np.float32)

    def triangles(self) -> np.ndarray:
        """(nt, 3) global 0-based node ids of every element, in order."""
        return np.concatenate(self._rec_tri) if self._rec_tri else np.zeros((0, 3), np.int32)

    def bond_pairs(self) -> np.ndarray:
        """(nb, 2) caller ids of every bond (c_getBondsPtr,
        femmd2D/src/interface.cpp:45)."""
        n = C.c_int32()
        self._check(self._lib.fm_num_bonds(self._h, C.byref(n)))
        out = np.empty((n.value, 2), np.int32)
        if n.value:
            self._check(self._lib.fm_get_bonds(self._h, _ip(out)))
        return out

    def n_bonds(self) -> int:
        """c_getNumberOfBonds (femmd2D/src/interface.cpp:42)."""
        n = C.c_int32()
        self._check(self._lib.fm_num_bonds(self._h, C.byref(n)))
        return n.value

    def set_kBond(self, k: float):
        """System.set_kBond (femmd2D/src/interface.cpp:455-465): the stiffness
        given to later create_bond calls that pass no k."""
        self._kbond = float(k)

    def create_bond(self, i: int, j: int, k: Optional[float] = None, r0: float = -1.0):
        """c_createBond (femmd2D/src/interface.cpp:67); r0 < 0 = current length."""
        self.add_bonds([i], [j], [self._kbond if k is None else k], [r0])

    def add_bonds(self, i, j, k, r0=None):
        i = _i32(i)
        n = i.shape[0]
        j = _i32(j, n)
        k = _f32(np.broadcast_to(np.asarray(k, np.float32), (n,)))
        r0 = None if r0 is None else _f32(np.broadcast_to(np.asarray(r0, np.float32), (n,)))
        self._check(self._lib.fm_add_bonds(self._h, n, _ip(i), _ip(j), _fp(k), _fp(r0)))

    def set_pair_potential(self, eps, rc, us):
        """System.set_pair_potential(eps, rc, us) (femmd2D/src/interface.cpp:493-531):
        row-major ntypes x ntypes tables; rc in sigma units; us additive."""
        eps, rc, us = _f32(np.ravel(eps)), _f32(np.ravel(rc)), _f32(np.ravel(us))
        nt = int(round(math.sqrt(eps.shape[0])))
        if nt * nt != eps.shape[0] or rc.shape[0] != eps.shape[0] or us.shape[0] != eps.shape[0]:
            raise ValueError("pair tables must be square and of equal size")
        self._check(self._lib.fm_set_pair_table(self._h, nt, _fp(eps), _fp(rc), _fp(us)))

    def set_youngs_modulus(self, E: float, nu: float = 0.3, model: int = FM_NEO_HOOKEAN):
        """System.set_youngs_modulus (femmd2D/src/interface.cpp:479-490) with the
        Poisson ratio and material model made explicit."""
        self._check(self._lib.fm_set_material(self._h, int(model), float(E), float(nu)))

    def set_lj(self, on: bool):
        self._check(self._lib.fm_set_lj(self._h, int(bool(on))))

    def set_exclusion(self, i: int, group: int):
        self._check(self._lib.fm_set_exclusion(self._h, int(i), int(group)))

    def set_mass(self, ids, mass):
        """Per-particle masses; zero or infinite mass pins a particle."""
        ids = _i32(np.ravel(ids))
        mass = _f32(np.broadcast_to(np.asarray(mass, np.float32), ids.shape))
        self._check(self._lib.fm_set_mass(self._h, ids.shape[0], _ip(ids), _fp(mass)))

    def set_dt(self, dt: float):
        self._check(self._lib.fm_set_dt(self._h, float(dt)))
        self._dt = float(dt)

    def get_dt(self) -> float:
        return self._dt

    def get_box(self) -> dict:
        return {"x": self._box[0], "y": self._box[1]}

    def allocate_cuda(self):
        """System.allocate_cuda + init_neighbor_list: upload and validate."""
        self._check(self._lib.fm_commit(self._h))
        self._committed = True

    commit = allocate_cuda

    def init_neighbor_list(self):
        """Kept for script compatibility: cells are built by allocate_cuda."""
        if not self._committed:
            self.allocate_cuda()

    def set_tuning(self, tile_cells: int = 0, block: int = 0, use_graph: int = -1):
        self._check(self._lib.fm_set_tuning(self._h, int(tile_cells), int(block), int(use_graph)))

    # -- stepping --------------------------------------------------------------
    def run(self, steps: int):
        """System.run / c_run (femmd2D/src/integrator.cpp:31-57)."""
        self._check(self._lib.fm_step(self._h, int(steps)))

    def do_step(self):
        """c_doStep (femmd2D/src/integrator.cpp:8-29)."""
        self.run(1)

    def calc_forces(self):
        """c_calcForces (femmd2D/src/interactions.cpp:83-109)."""
        self._check(self._lib.fm_compute_forces(self._h))

    def energies(self) -> dict:
        e = FmEnergies()
        self._check(self._lib.fm_get_energies(self._h, C.byref(e)))
        return e.as_dict()

    def set_observe_every(self, every: int):
        """1 (default): every step forms and accumulates its observables on
        the device, as Integrator::doStep -> Properties::compute does
        (femmd2D/src/integrator.cpp:18); 0: only the last step of each run."""
        self._check(self._lib.fm_set_observe_every(self._h, int(every)))

    def averages(self) -> dict:
        """Running averages (avg, stdv) of the per-step observables since the
        last reset (femmd2D/src/properties.cpp:39-86), with their count."""
        a = FmAverages()
        self._check(self._lib.fm_get_averages(self._h, C.byref(a)))
        return a.as_dict()

    def reset_averages(self):
        """Properties::accumulate(0) (femmd2D/src/properties.cpp:41-55)."""
        self._check(self._lib.fm_reset_averages(self._h))

    def get_props(self) -> dict:
        """System.get_props (femmd2D/src/interface.cpp:627-644): current values
        per particle as femmd2D/src/properties.cpp:15-36 defines them."""
        e = self.energies()
        n = e["n"]
        uk = e["ke"] / n
        up = (e["u_lj"] + e["u_bond"] + e["u_fea"]) / n
        return {
            "steps": e["steps"],
            "uKinetic": uk,
            "uPotential": up,
            "uFea": e["u_fea"] / n,
            "uBonded": e["u_bond"] / n,
            "uTotal": uk + up,
            "temperature": e["temperature"],
            "pressure": e["pressure"],
            "virial": e["w_lj"],
            "feaVirial": e["w_fea"],
        }

    def info(self) -> dict:
        return {"N": self.N, "box": self.get_box(), "dt": self._dt}

    # -- trajectory files (femmd2D/src/md3dsystem.cpp:232-377) --------------
    def save_xyz(self, name: str, s: int = 0):
        from . import io
        io.save_xyz(self, name, s)

    def save_vtk(self, prefix: str) -> str:
        """System.save_vtk (femmd2D/src/interface.cpp:664)."""
        from . import io
        return io.save_vtk(self, prefix)

    def save_vtf(self, name: str, s: int = 0):
        """System.save_vtf (femmd2D/src/interface.cpp:665)."""
        from . import io
        io.save_vtf(self, name, s)

    def save_cluster_vtf(self, name: str, s: int = 0, scale: float = 1.0):
        from . import io
        io.save_cluster_vtf(self, name, s, scale)

    def save_cluster_xyz(self, name: str, s: int = 0):
        from . import io
        io.save_cluster_xyz(self, name, s)

    # -- observables (femmd2D/src/feaElements.cpp:86-196,
    #    femmd2D/src/md3dsystem.cpp:92-179) ---------------------------------
    def cluster_count(self) -> int:
        n = C.c_int32()
        self._check(self._lib.fm_cluster_count(self._h, C.byref(n)))
        return n.value

    def cluster_props(self, unfolded: bool = False) -> dict:
        """Per elastomer cluster: mass, centre of mass ``cm`` (of the unfolded
        nodes), its velocity ``vcm`` and the circular-mean ``centroid``; with
        ``unfolded`` also the unfolded positions ``ux``, ``uy`` (caller order;
        particles outside clusters keep their wrapped positions)."""
        nc = self.cluster_count()
        out = np.zeros((max(nc, 1), 7), dtype=np.float64)
        ux = uy = None
        if unfolded:
            ux = np.zeros(self.N, dtype=np.float32)
            uy = np.zeros(self.N, dtype=np.float32)
        self._check(self._lib.fm_cluster_props(self._h, _dp(out), _fp(ux), _fp(uy)))
        out = out[:nc]
        r = {"mass": out[:, 0].copy(), "cm": out[:, 1:3].copy(), "vcm": out[:, 3:5].copy(),
             "centroid": out[:, 5:7].copy()}
        if unfolded:
            r["ux"], r["uy"] = ux, uy
        return r

    def calc_cluster_props(self) -> dict:
        """System.calc_cluster_props (femmd2D/src/interface.cpp:333)."""
        return self.cluster_props()

    def cluster_virial(self) -> float:
        """sum over interacting pairs of cluster particles of
        1/2 f_ij . (R_ci - R_cj) (femmd2D/src/interactions.cpp:370-382)."""
        v = C.c_double()
        self._check(self._lib.fm_cluster_virial(self._h, C.byref(v)))
        return v.value

    def set_temperature(self, temp: float, ids=None):
        """MDSystem::setTemperature (femmd2D/src/md3dsystem.cpp:92-136):
        all particles (T = KE / N), or a group of caller ids with the
        reference's group formula T = 2 KE_group / (3 N)."""
        if ids is None:
            self._check(self._lib.fm_set_temperature(self._h, float(temp), 0, None))
        else:
            a = _i32(ids)
            self._check(self._lib.fm_set_temperature(self._h, float(temp), len(a), _ip(a)))

    def set_temperature_noalloc(self, temp: float, ids=None):
        """set_temperature without device allocation (fm_set_temperature_noalloc:
        scratch in a fixed device array; at most 2^20 ids)."""
        if ids is None:
            self._check(self._lib.fm_set_temperature_noalloc(self._h, float(temp), 0, None))
        else:
            a = _i32(ids)
            self._check(self._lib.fm_set_temperature_noalloc(self._h, float(temp), len(a), _ip(a)))

    def get_state(self):
        n = self.N
        out = [np.empty(n, np.float32) for _ in range(4)]
        self._check(self._lib.fm_get_state(self._h, *map(_fp, out)))
        return tuple(out)

    def set_state(self, x=None, y=None, vx=None, vy=None):
        n = self.N
        arrs = [None if a is None else _f32(a, n) for a in (x, y, vx, vy)]
        self._check(self._lib.fm_set_state(self._h, *map(_fp, arrs)))

    def get_forces(self):
        n = self.N
        fx, fy = np.empty(n, np.float32), np.empty(n, np.float32)
        self._check(self._lib.fm_get_forces(self._h, _fp(fx), _fp(fy)))
        return fx, fy

    # -- diagnostics -----------------------------------------------------------
    def cells(self):
        """(ncx, ncy, key[n], start[ncells+1], perm[n]) of the last binning."""
        n = self.N
        ncx, ncy = C.c_int32(), C.c_int32()
        self._check(self._lib.fm_get_cells(self._h, C.byref(ncx), C.byref(ncy), None, None, None))
        key = np.empty(n, np.int32)
        start = np.empty(ncx.value * ncy.value + 1, np.int32)
        perm = np.empty(n, np.int32)
        self._check(self._lib.fm_get_cells(self._h, C.byref(ncx), C.byref(ncy), _ip(key),
                                           _ip(start), _ip(perm)))
        return ncx.value, ncy.value, key, start, perm

    def pair_sets(self):
        n = self.N
        cnt = np.empty(n, np.int32)
        s1 = np.empty(n, np.int64)
        s2 = np.empty(n, np.int64)
        lp = C.POINTER(C.c_int64)
        self._check(self._lib.fm_pair_sets(self._h, _ip(cnt), s1.ctypes.data_as(lp),
                                           s2.ctypes.data_as(lp)))
        return cnt, s1, s2

    def device_errors(self) -> int:
        return int(self._lib.fm_device_errors(self._h))

    def clear_errors(self):
        self._check(self._lib.fm_clear_errors(self._h))

    def time_forces(self, reps: int = 10) -> float:
        ms = C.c_double()
        self._check(self._lib.fm_time_forces(self._h, int(reps), C.byref(ms)))
        return ms.value

    def profile(self, nsteps: int = 10) -> dict:
        """Per-step device ms of each kernel (its dispatch's start and end
        timestamps) and of the whole step body."""
        ms = (C.c_double * 6)()
        self._check(self._lib.fm_profile_step(self._h, int(nsteps), ms))
        return dict(zip(("kick", "bin", "sort", "list_build", "forces", "step"), list(ms)))

    def set_skin(self, skin: float):
        """Neighbour-list skin (default 0.6; the reference's CPU list uses 0.4,
        femmd2D/src/nn.h:63); 0 walks the cells every step.  Before
        allocate_cuda."""
        self._check(self._lib.fm_set_skin(self._h, float(skin)))

    def list_stats(self) -> dict:
        """{"mean", "max", "rebuilds"} of the neighbour lists (zeros when
        the cells are walked every step)."""
        m, mx, nb = C.c_double(), C.c_int32(), C.c_int64()
        self._check(self._lib.fm_list_stats(self._h, C.byref(m), C.byref(mx), C.byref(nb)))
        return {"mean": m.value, "max": mx.value, "rebuilds": nb.value}

    def set_reorder_interval(self, steps: int):
        """Re-sort the internal order of the particles that no element, bond
        or cluster refers to into the current cell order every `steps` steps
        (default 1000; 0 = never).  Keeps the step's memory accesses local as
        particles diffuse; caller ids never change."""
        self._check(self._lib.fm_set_reorder_interval(self._h, int(steps)))

    def reorder(self):
        """Re-sort the internal order now (see set_reorder_interval)."""
        self._check(self._lib.fm_reorder(self._h))

    def reorder_info(self) -> dict:
        """{"interval", "count"}: the re-sort interval and the re-sorts so far."""
        iv, n = C.c_int32(), C.c_int64()
        self._check(self._lib.fm_reorder_info(self._h, C.byref(iv), C.byref(n)))
        return {"interval": iv.value, "count": n.value}

    # -- adjoint (no reference counterpart) -----------------------------------
    def force_jvp(self, wx, wy):
        """(jx, jy, dE, deps, dsig): j = (df/dx) w at the current positions and
        the parameter terms w . df/dE_e (per element), w . df/deps,
        w . df/dsigma (common pair eps and sigma)."""
        n = self.N
        wx, wy = _f32(wx, n), _f32(wy, n)
        jx, jy = np.empty(n, np.float32), np.empty(n, np.float32)
        dE = np.empty(self._nelem, np.float64)
        deps, dsig = C.c_double(), C.c_double()
        self._check(self._lib.fm_force_jvp(self._h, _fp(wx), _fp(wy), _fp(jx), _fp(jy),
                                           _dp(dE), C.byref(deps), C.byref(dsig)))
        return jx, jy, dE, deps.value, dsig.value

    def adjoint_forward(self, nsteps: int, ckpt: int = 0):
        """fm_adjoint_forward: nsteps steps (the state advances), recorded
        for adjoint_backward (checkpoints every ckpt steps; 0: the engine's
        choice, the whole trajectory when it fits)."""
        self._check(self._lib.fm_adjoint_forward(self._h, int(nsteps), int(ckpt)))

    def adjoint_backward(self, dx=None, dy=None) -> dict:
        """fm_adjoint_backward: the gradient {"dE", "deps", "dsigma"} of a
        loss whose derivative with respect to the final positions is
        (dx, dy) (caller order), through the recorded forward pass."""
        n = self.N
        dx = None if dx is None else _f32(dx, n)
        dy = None if dy is None else _f32(dy, n)
        deps, dsig = C.c_double(), C.c_double()
        dE = np.empty(self._nelem, np.float64)
        self._check(self._lib.fm_adjoint_backward(self._h, _fp(dx), _fp(dy), _dp(dE),
                                                  C.byref(deps), C.byref(dsig)))
        return {"dE": dE, "deps": deps.value, "dsigma": dsig.value}

    def adjoint(self, nsteps: int, ids, targets, ckpt: int = 0) -> dict:
        """Runs nsteps forward and returns the loss
        (1/K) sum_k |x_{ids[k]}(T) - targets[k]|^2 with its gradient
        {"loss", "dE" (per element), "deps", "dsigma"} from the discrete
        adjoint (checkpoints every ckpt steps; 0 lets the engine store as
        much of the trajectory as a quarter of the free HBM holds)."""
        ids = _i32(ids)
        t = np.asarray(targets, np.float64).reshape(-1, 2)
        if t.shape[0] != ids.shape[0]:
            raise ValueError("one target per loss particle")
        tx, ty = _f32(t[:, 0]), _f32(t[:, 1])
        loss, deps, dsig = C.c_double(), C.c_double(), C.c_double()
        dE = np.empty(self._nelem, np.float64)
        self._check(self._lib.fm_adjoint(self._h, int(nsteps), int(ckpt), ids.shape[0], _ip(ids),
                                         _fp(tx), _fp(ty), C.byref(loss), _dp(dE),
                                         C.byref(deps), C.byref(dsig)))
        return {"loss": loss.value, "dE": dE, "deps": deps.value, "dsigma": dsig.value}


class Batch:
    """Many independent replicas stepped together (one launch per kernel for
    all replicas, graph-captured step).  No reference counterpart."""

    def __init__(self, systems: Sequence[System]):
        self._lib = _lib.load()
        self._systems = list(systems)
        arr = (C.c_void_p * len(systems))(*[s.handle for s in systems])
        h = C.c_void_p()
        rc = self._lib.fm_batch_create(len(systems), arr, C.byref(h))
        self._h = h
        if rc != _lib.FM_OK:
            msg = self._lib.fm_batch_last_error(h).decode() if h else "invalid argument"
            raise FemmdError(rc, msg)
        self._n = [s.N for s in systems]

    def _check(self, rc):
        if rc != _lib.FM_OK:
            raise FemmdError(rc, self._lib.fm_batch_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            self._lib.fm_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return len(self._n)

    def run(self, steps: int):
        self._check(self._lib.fm_batch_step(self._h, int(steps)))

    def get_state(self, r: int):
        n = self._n[r]
        out = [np.empty(n, np.float32) for _ in range(4)]
        self._check(self._lib.fm_batch_get_state(self._h, int(r), *map(_fp, out)))
        return tuple(out)

    def energies(self, r: int) -> dict:
        e = FmEnergies()
        self._check(self._lib.fm_batch_get_energies(self._h, int(r), C.byref(e)))
        return e.as_dict()

    def averages(self, r: int) -> dict:
        a = FmAverages()
        self._check(self._lib.fm_batch_get_averages(self._h, int(r), C.byref(a)))
        return a.as_dict()

    def set_tuning(self, tile_cells: int = 0, block: int = 0, use_graph: int = -1):
        self._check(self._lib.fm_batch_set_tuning(self._h, int(tile_cells), int(block),
                                                  int(use_graph)))

    def set_reorder_interval(self, steps: int):
        """As System.set_reorder_interval, for every replica of the batch."""
        self._check(self._lib.fm_batch_set_reorder_interval(self._h, int(steps)))

    def tune(self, cache_path: Optional[str] = None):
        p = None if cache_path is None else str(cache_path).encode()
        self._check(self._lib.fm_batch_tune(self._h, p))

    def device_errors(self) -> int:
        return int(self._lib.fm_batch_device_errors(self._h))
