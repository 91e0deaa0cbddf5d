"""Python access to the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by femmd_old_amd/.  See oracle.h for the
reference lines each function restates.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

OR_U_LJ, OR_W_LJ, OR_U_BOND, OR_W_BOND, OR_U_FEA, OR_W_FEA, OR_NPAIRS, OR_NINV = range(8)
# sums of the magnitudes of the bond / element virial terms (oracle.h)
OR_WS_BOND, OR_WS_FEA, OR_NEN = 8, 9, 10
# or_run history columns beyond the energies
OR_H_KE, OR_H_VVMAX = OR_NEN, OR_NEN + 1

_d = C.POINTER(C.c_double)
_i = C.POINTER(C.c_int32)
_l = C.POINTER(C.c_int64)
_f = C.POINTER(C.c_float)


class OrSystem(C.Structure):
    _fields_ = [
        ("n", C.c_int32),
        ("Lx", C.c_double), ("Ly", C.c_double),
        ("x", _d), ("y", _d), ("sigma", _d),
        ("type", _i), ("excl", _i),
        ("ntypes", C.c_int32),
        ("eps", _d), ("rc", _d), ("ushift", _d),
        ("lj_on", C.c_int32),
        ("nbond", C.c_int32),
        ("bi", _i), ("bj", _i), ("bk", _d), ("br0", _d),
        ("ntri", C.c_int32),
        ("tri", _i),
        ("rx", _d), ("ry", _d),
        ("tE", _d), ("tnu", _d),
        ("model", C.c_int32),
    ]


def build():
    """Compile liboracle.so with its Makefile (g++)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


_LIB = None


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        L.or_wrap_f32.argtypes = [C.c_int32, _f, _f, C.c_float, C.c_float]
        L.or_cell_keys_f32.argtypes = [C.c_int32, _f, _f, C.c_float, C.c_float, C.c_int32,
                                       C.c_int32, _i]
        L.or_cell_count.argtypes = [C.c_double, C.c_double]
        L.or_cell_count.restype = C.c_int32
        L.or_counting_sort.argtypes = [C.c_int32, _i, C.c_int32, _i, _i]
        L.or_pair_sets_f32.argtypes = [C.POINTER(OrSystem), C.c_int32, _i, _l, _l]
        L.or_pair_sets_f32.restype = C.c_int32
        L.or_forces.argtypes = [C.POINTER(OrSystem), C.c_int32, C.c_int32, _d, _d, _d, _d]
        L.or_forces.restype = C.c_int32
        L.or_run.argtypes = [C.POINTER(OrSystem), _d, _d, _d, _d, _d, _d, C.c_int32, C.c_double,
                             C.c_int32, _d]
        L.or_run.restype = C.c_int32
        L.or_tangent.argtypes = [C.POINTER(OrSystem), _d, _d, C.c_int32, C.c_double, _d, _d,
                                 C.c_double, C.c_double, C.c_double, _d, _d]
        L.or_tangent.restype = C.c_int32
        L.or_bl_create.argtypes = [C.POINTER(OrSystem), _f, _f, _f, C.c_float, C.c_float,
                                   C.c_int32]
        L.or_bl_create.restype = C.c_void_p
        L.or_bl_run.argtypes = [C.c_void_p, C.c_int32]
        L.or_bl_run.restype = C.c_double
        L.or_bl_get.argtypes = [C.c_void_p, _f, _f, _f, _f, _d]
        L.or_bl_destroy.argtypes = [C.c_void_p]
        _LIB = L
    return _LIB


def _dp(a):
    return None if a is None else a.ctypes.data_as(_d)


def _ip(a):
    return None if a is None else a.ctypes.data_as(_i)


class OracleSystem:
    """Holds the numpy arrays behind an OrSystem (keeps them alive)."""

    def __init__(self, Lx, Ly, x, y, sigma=None, type=None, excl=None, eps=(1.0,), rc=(2.5,),
                 ushift=(0.0,), lj_on=True, bonds=None, tri=None, refx=None, refy=None, E=None,
                 nu=None, model=0):
        n = len(x)
        f64 = lambda a: np.ascontiguousarray(np.asarray(a, np.float64))
        i32 = lambda a: np.ascontiguousarray(np.asarray(a, np.int32))
        self.keep = {}
        k = self.keep
        k["x"], k["y"] = f64(x), f64(y)
        k["sigma"] = f64(np.ones(n) if sigma is None else sigma)
        k["type"] = i32(np.zeros(n) if type is None else type)
        k["excl"] = i32(-np.ones(n) if excl is None else excl)
        k["eps"], k["rc"], k["ushift"] = f64(np.ravel(eps)), f64(np.ravel(rc)), f64(np.ravel(ushift))
        nt = int(round(np.sqrt(k["eps"].shape[0])))
        s = OrSystem()
        s.n = n
        s.Lx, s.Ly = float(Lx), float(Ly)
        s.x, s.y, s.sigma = _dp(k["x"]), _dp(k["y"]), _dp(k["sigma"])
        s.type, s.excl = _ip(k["type"]), _ip(k["excl"])
        s.ntypes = nt
        s.eps, s.rc, s.ushift = _dp(k["eps"]), _dp(k["rc"]), _dp(k["ushift"])
        s.lj_on = int(bool(lj_on))
        if bonds is not None and len(bonds[0]):
            k["bi"], k["bj"] = i32(bonds[0]), i32(bonds[1])
            k["bk"], k["br0"] = f64(bonds[2]), f64(bonds[3])
            s.nbond = k["bi"].shape[0]
            s.bi, s.bj, s.bk, s.br0 = _ip(k["bi"]), _ip(k["bj"]), _dp(k["bk"]), _dp(k["br0"])
        if tri is not None and len(tri):
            k["tri"] = i32(np.ravel(tri))
            k["rx"] = f64(x if refx is None else refx)
            k["ry"] = f64(y if refy is None else refy)
            ntri = k["tri"].shape[0] // 3
            k["tE"] = f64(np.broadcast_to(np.asarray(E, float), (ntri,)))
            k["tnu"] = f64(np.broadcast_to(np.asarray(nu, float), (ntri,)))
            s.ntri = ntri
            s.tri, s.rx, s.ry = _ip(k["tri"]), _dp(k["rx"]), _dp(k["ry"])
            s.tE, s.tnu = _dp(k["tE"]), _dp(k["tnu"])
        s.model = int(model)
        self.s = s
        self.n = n

    def with_positions(self, x, y):
        """Same system at other positions (FP64 arrays)."""
        self.keep["x"] = np.ascontiguousarray(np.asarray(x, np.float64))
        self.keep["y"] = np.ascontiguousarray(np.asarray(y, np.float64))
        self.s.x, self.s.y = _dp(self.keep["x"]), _dp(self.keep["y"])
        return self


def box_f32(L: float) -> float:
    """The box length as the reference holds it (Box is a real=float vector,
    femmd2D/src/def.h:86): every image shift uses this FP32 value."""
    return float(np.float32(L))


def from_workload(w, x=None, y=None, *, E=None, eps=None, sigma=None) -> OracleSystem:
    """Oracle system of a configs.Workload at caller-order positions (x, y),
    default the generated ones (wrapped in FP32 as the engine stores them).
    The box is the FP32 one the engine and the reference compute with.
    E (per element), eps and sigma (per particle) override the workload's
    parameters."""
    ga = w.global_arrays()
    Lx, Ly = box_f32(w.Lx), box_f32(w.Ly)
    if x is None:
        xf = ga["x"].astype(np.float32)
        yf = ga["y"].astype(np.float32)
        wrap_f32(xf, yf, Lx, Ly)
        x, y = xf, yf
    return OracleSystem(Lx, Ly, x, y, sigma=ga["sigma"] if sigma is None else sigma,
                        excl=ga["excl"], eps=[w.eps if eps is None else eps],
                        rc=[w.rc], ushift=[w.ushift], bonds=w.bonds, tri=ga["tri"],
                        refx=ga["refx"].astype(np.float32), refy=ga["refy"].astype(np.float32),
                        E=ga["E"] if E is None else E, nu=ga["nu"], model=w.model)


# ---------------------------------------------------------------------------
def wrap_f32(x: np.ndarray, y: np.ndarray, Lx: float, Ly: float):
    assert x.dtype == np.float32 and y.dtype == np.float32
    lib().or_wrap_f32(x.shape[0], x.ctypes.data_as(_f), y.ctypes.data_as(_f), Lx, Ly)


def cell_keys(x, y, Lx, Ly, ncx, ncy) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(y, np.float32)
    key = np.empty(x.shape[0], np.int32)
    lib().or_cell_keys_f32(x.shape[0], x.ctypes.data_as(_f), y.ctypes.data_as(_f), Lx, Ly, ncx,
                           ncy, key.ctypes.data_as(_i))
    return key


def counting_sort(key: np.ndarray, ncells: int):
    key = np.ascontiguousarray(key, np.int32)
    start = np.empty(ncells + 1, np.int32)
    perm = np.empty(key.shape[0], np.int32)
    lib().or_counting_sort(key.shape[0], _ip(key), ncells, _ip(start), _ip(perm))
    return start, perm


def pair_sets(osys: OracleSystem, mode: int = 1):
    n = osys.n
    cnt = np.empty(n, np.int32)
    s1 = np.empty(n, np.int64)
    s2 = np.empty(n, np.int64)
    rc = lib().or_pair_sets_f32(C.byref(osys.s), mode, _ip(cnt), s1.ctypes.data_as(_l),
                                s2.ctypes.data_as(_l))
    assert rc == 0
    return cnt, s1, s2


def forces(osys: OracleSystem, pred_f32: bool = True, all_pairs: bool = False,
           scale: bool = False):
    """FP64 forces (fx, fy), energies[OR_NEN] and, with scale=True, an (2, n) array:
    row 0 the per-particle sum of force-term magnitudes, row 1 the sum of
    r |df/dr| over those terms (see oracle.h)."""
    n = osys.n
    fx = np.empty(n)
    fy = np.empty(n)
    en = np.empty(OR_NEN)
    fs = np.empty(2 * n) if scale else None
    rc = lib().or_forces(C.byref(osys.s), int(pred_f32), int(all_pairs), _dp(fx), _dp(fy), _dp(en),
                         _dp(fs))
    assert rc == 0, rc
    return (fx, fy, en, fs.reshape(2, n)) if scale else (fx, fy, en)


def force_error(fx, fy, rx, ry, fs):
    """Per-particle error of an FP32 force against the FP64 oracle:
    |f - f_ref| / max(|f_ref|, sum_j |f_ij|, 0.1 * sum_j r_ij |df_ij/dr_ij|).

    The second term is the scale of the cancellation in the net force; the
    third is the force change that a relative change of 1e-6 (16 FP32 ulps)
    in each separation produces, times 1e-5: below it a force is not
    resolved by FP32 inputs at all (an unstrained element, a bond at rest, a
    pair sitting at the zero of the LJ force)."""
    d = np.hypot(np.asarray(fx, float) - rx, np.asarray(fy, float) - ry)
    den = np.maximum(np.maximum(np.hypot(rx, ry), fs[0]), 0.1 * fs[1])
    return d / np.where(den > 0, den, 1.0)


def force_error_terms(fx, fy, rx, ry, fs, tol=1e-5):
    """Which denominator term of force_error binds, per particle: the
    fraction of particles where each of |f_ref|, sum_j |f_ij| and
    0.1 * sum_j r_ij |df_ij/dr_ij| is the largest, and the largest error
    and the number of particles above tol under the two-term rule
    max(|f_ref|, sum_j |f_ij|) alone."""
    d = np.hypot(np.asarray(fx, float) - rx, np.asarray(fy, float) - ry)
    terms = np.stack([np.hypot(rx, ry), np.asarray(fs[0], float), 0.1 * np.asarray(fs[1], float)])
    which = np.argmax(terms, axis=0)
    den2 = np.maximum(terms[0], terms[1])
    e2 = d / np.where(den2 > 0, den2, 1.0)
    n = max(1, d.shape[0])
    return {"net": float(np.mean(which == 0)), "cancel": float(np.mean(which == 1)),
            "resolution": float(np.mean(which == 2)), "max_two_term": float(e2.max(initial=0.0)),
            "above_tol_two_term": int(np.count_nonzero(e2 > tol)), "n": n}


def run(osys: OracleSystem, x, y, vx, vy, mass, nsteps, dt, pred_f32=False, history=False):
    """FP64 KDK rollout; returns (x, y, vx, vy, hist or None)."""
    x, y, vx, vy = (np.array(a, np.float64) for a in (x, y, vx, vy))
    mass = np.asarray(mass, np.float64)
    inv_m = np.where(np.isfinite(mass) & (mass > 0), 1.0 / np.where(mass > 0, mass, 1.0), 0.0)
    m_ke = np.where(np.isfinite(mass), mass, 0.0)
    hist = np.empty((nsteps, OR_NEN + 2)) if history else None
    rc = lib().or_run(C.byref(osys.s), _dp(x), _dp(y), _dp(vx), _dp(vy), _dp(m_ke), _dp(inv_m),
                      int(nsteps), float(dt), int(pred_f32), _dp(hist))
    assert rc == 0, rc
    return x, y, vx, vy, hist


def _inv_mass(mass):
    mass = np.asarray(mass, np.float64)
    return np.where(np.isfinite(mass) & (mass > 0), 1.0 / np.where(mass > 0, mass, 1.0), 0.0)


def trajectory(osys: OracleSystem, x, y, vx, vy, mass, nsteps, dt):
    """The FP64 or_run rollout's positions at steps 0..nsteps, each an
    (nsteps + 1, n) array."""
    x, y, vx, vy = (np.array(a, np.float64) for a in (x, y, vx, vy))
    tx = np.empty((nsteps + 1, x.shape[0]))
    ty = np.empty_like(tx)
    tx[0], ty[0] = x, y
    for k in range(nsteps):
        x, y, vx, vy, _ = run(osys, x, y, vx, vy, mass, 1, dt)
        tx[k + 1], ty[k + 1] = x, y
    return tx, ty


def tangent(osys: OracleSystem, tx, ty, dt, mass, dE=None, deps=0.0, dsig=0.0, h=1e-6):
    """Forward-mode derivative (dx, dy) of the final positions of a rollout
    along the recorded trajectory (tx, ty), (nsteps + 1, n) each, for the
    parameter direction (dE per element, deps on eps, dsig on every sigma);
    force derivatives by FP64 central differences of step h (or_tangent)."""
    tx = np.ascontiguousarray(tx, np.float64)
    ty = np.ascontiguousarray(ty, np.float64)
    n = tx.shape[1]
    inv_m = _inv_mass(mass)
    dEa = None if dE is None else np.ascontiguousarray(dE, np.float64)
    dx, dy = np.empty(n), np.empty(n)
    rc = lib().or_tangent(C.byref(osys.s), _dp(tx), _dp(ty), tx.shape[0] - 1, float(dt),
                          _dp(inv_m), _dp(dEa), float(deps), float(dsig), float(h), _dp(dx),
                          _dp(dy))
    assert rc == 0, rc
    return dx, dy


class Baseline:
    """The reference DEV=CPU path restated in FP32 (Verlet list, skin 0.4)."""

    def __init__(self, osys: OracleSystem, vx, vy, mass, dt, skin=0.4, threads=1):
        self._vx = np.ascontiguousarray(vx, np.float32)
        self._vy = np.ascontiguousarray(vy, np.float32)
        self._m = np.ascontiguousarray(mass, np.float32)
        self.osys = osys
        self.h = lib().or_bl_create(C.byref(osys.s), self._vx.ctypes.data_as(_f),
                                    self._vy.ctypes.data_as(_f), self._m.ctypes.data_as(_f),
                                    float(dt), float(skin), int(threads))
        if not self.h:
            raise RuntimeError("baseline create failed (box too small?)")
        self.threads = threads

    def run(self, nsteps: int) -> float:
        return lib().or_bl_run(self.h, int(nsteps))

    def state(self):
        n = self.osys.n
        out = [np.empty(n, np.float32) for _ in range(4)]
        props = np.empty(5)
        lib().or_bl_get(self.h, *(a.ctypes.data_as(_f) for a in out), _dp(props))
        return out, props

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_bl_destroy(self.h)
            self.h = None


# ---------------------------------------------------------------------------
# Observables (numpy restatements; test infrastructure only)
# ---------------------------------------------------------------------------
def cluster_props(x, y, vx, vy, mass, ranges, Lx, Ly):
    """Per cluster (caller-id ranges [first, first + n)): the unfolded node
    positions and (mass, cm x, cm y, vcm x, vcm y, centroid x, centroid y).

    femmd2D/src/feaElements.cpp:86-149 (FeaElements::unfoldPos): positions
    are FP32 as the reference's real; a cluster wider than L/2 has nodes
    beyond L/2 moved down by L and, if the mean is then negative, all moved
    up by L.  The centre of mass and its velocity are mass weighted (sums in
    FP64).  :161-193 (updateClustersCentroid): circular mean of the wrapped
    positions, x and y each from their own cos/sin sums (the reference reuses
    the x sums for y)."""
    Lxf, Lyf = np.float32(Lx), np.float32(Ly)
    hx, hy = np.float32(0.5) * Lxf, np.float32(0.5) * Lyf
    x32 = np.asarray(x, np.float32)
    y32 = np.asarray(y, np.float32)
    ux = x32.copy()
    uy = y32.copy()
    out = []
    for first, n in ranges:
        sl = slice(first, first + n)
        for u, h, L in ((ux, hx, Lxf), (uy, hy, Lyf)):
            v = u[sl].copy()
            if v.max() - v.min() > h:
                v = np.where(v > h, v - L, v).astype(np.float32)
                if v.astype(np.float64).sum() / n < 0:
                    v = (v + L).astype(np.float32)
            u[sl] = v
        m = np.asarray(mass[sl], np.float64)
        M = m.sum()
        cmx = (m * ux[sl]).sum() / M
        cmy = (m * uy[sl]).sum() / M
        vcx = (m * np.asarray(vx[sl], np.float32)).sum() / M
        vcy = (m * np.asarray(vy[sl], np.float32)).sum() / M
        tx = 2 * np.pi * x32[sl].astype(np.float64) / np.float64(Lxf)
        ty = 2 * np.pi * y32[sl].astype(np.float64) / np.float64(Lyf)
        cx = np.float64(Lxf) * (np.arctan2(-np.sin(tx).sum(), -np.cos(tx).sum()) + np.pi) / (2 * np.pi)
        cy = np.float64(Lyf) * (np.arctan2(-np.sin(ty).sum(), -np.cos(ty).sum()) + np.pi) / (2 * np.pi)
        out.append((M, cmx, cmy, vcx, vcy, cx, cy))
    return np.array(out, np.float64).reshape(-1, 7), ux, uy


def cluster_virial(x, y, sigma, excl, cid, cm, Lx, Ly, eps=1.0, rc=2.5, sigma_mode=False):
    """sum over ordered interacting pairs (i, j), both in clusters, of
    1/2 f_ij . (R_ci - R_cj) with the minimum image
    (femmd2D/src/interactions.cpp:370-382; pairs with a particle outside
    every cluster add nothing).  Pair set: i != j by index and the FP32
    predicate of pred32 (femmd2D/src/interactionsCuda.cu:79-90); force in
    FP64 at the image-exact separation.  O(n^2): small systems only."""
    Lxf, Lyf = np.float32(Lx), np.float32(Ly)
    x32 = np.asarray(x, np.float32)
    y32 = np.asarray(y, np.float32)
    s32 = np.asarray(sigma, np.float32)
    cid = np.asarray(cid)
    excl = np.asarray(excl)
    rcf = np.float32(rc)
    tot = 0.0
    absum = 0.0
    idx = np.nonzero(cid >= 0)[0]
    for i in idx:
        j = idx[idx != i]
        dx = x32[i] - x32[j]
        sx = np.where(dx >= Lxf * np.float32(0.5), 1, np.where(dx < -Lxf * np.float32(0.5), -1, 0))
        dx = (dx - sx * Lxf).astype(np.float32)
        dy = y32[i] - y32[j]
        sy = np.where(dy >= Lyf * np.float32(0.5), 1, np.where(dy < -Lyf * np.float32(0.5), -1, 0))
        dy = (dy - sy * Lyf).astype(np.float32)
        dr2 = (dx * dx + dy * dy).astype(np.float32)
        sij = ((s32[i] + s32[j]) / np.float32(2)).astype(np.float32)
        if not sigma_mode:
            sij = np.full_like(sij, s32[i])
        rc2 = (rcf * rcf * sij * sij).astype(np.float32)
        gi, gj = excl[i], excl[j]
        ok = (dr2 < rc2) & ~((gi >= 0) & (gj == gi))
        j = j[ok]
        if j.size == 0:
            continue
        ddx = np.float64(x32[i]) - (x32[j].astype(np.float64) + sx[ok] * np.float64(Lxf))
        ddy = np.float64(y32[i]) - (y32[j].astype(np.float64) + sy[ok] * np.float64(Lyf))
        r2 = ddx * ddx + ddy * ddy
        s2 = sij[ok].astype(np.float64) ** 2
        s6 = (s2 / r2) ** 3
        g = 48.0 * eps * (s6 * s6 - 0.5 * s6) / r2
        dRx = cm[cid[i], 0] - cm[cid[j], 0]
        dRy = cm[cid[i], 1] - cm[cid[j], 1]
        dRx -= Lx * np.round(dRx / Lx)
        dRy -= Ly * np.round(dRy / Ly)
        t = 0.5 * (g * ddx * dRx + g * ddy * dRy)
        tot += t.sum()
        absum += np.abs(t).sum()
    return tot, absum
