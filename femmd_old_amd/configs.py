"""Seeded synthetic workloads of the benchmark configurations.

The reference ships no inputs beyond its small unit meshes
(femmd2D/meshes/*.txt), so every workload is generated here:

  c1  small hybrid: L = 74, 4,000 LJ solvent on a jittered square lattice,
      one elastomer disk of radius 10 (hex lattice, spacing 1), E = 50,
      nu = 0.3, node mass 1, rc = 2.5, T = 1, dt = 0.005
  c2  1,048,576 LJ solvent at density 0.8 over the free area plus 1,024
      disks of radius 5 (E ~ U[20, 100] per disk), rc = 2.5, T = 1
  c3  600k solvent at 0.75, 4,096 disks r ~ logU[4, 8], per-element
      E ~ logU[10, 200], nu ~ U[0.25, 0.45]; 8 anchor solvent particles per
      disk bonded (k = 100, r0 = 1) to the nearest rim node, dt = 0.002
  c4  512 replicas, N ~ logU[256, 32768], density ~ U[0.6, 0.9],
      T ~ U[0.5, 2.0], even replicas carry one disk r ~ U[6, 15]
  c5  128 x 256 hex sheet (bottom row pinned) in 200k solvent at 0.7,
      per-element E from 16 Fourier modes in [20, 120]; WCA pair potential

Positions are generated in FP64 and stored as FP32 by the engine.  A
``scale`` argument shrinks the particle counts of c2-c5 for tests.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

RC_LJ = 2.5
U_SHIFT_25 = float(-4.0 * (RC_LJ ** -12 - RC_LJ ** -6))  # ~ +0.016316
RC_WCA = float(2.0 ** (1.0 / 6.0))


# ---------------------------------------------------------------------------
# meshes
# ---------------------------------------------------------------------------
def hex_disk(radius: float, spacing: float = 1.0):
    """Regular hexagonal-lattice disk: vertices (nv, 2) centred on 0 and CCW
    triangles (nt, 3), 0-based; only vertices used by a triangle are kept."""
    a = spacing
    m = int(math.ceil(radius / a)) + 2
    ii, jj = np.meshgrid(np.arange(-2 * m, 2 * m + 1), np.arange(-2 * m, 2 * m + 1), indexing="ij")
    px = a * (ii + 0.5 * jj)
    py = a * (jj * math.sqrt(3.0) / 2.0)
    inside = px * px + py * py <= radius * radius * (1 + 1e-9)
    idx = -np.ones(ii.shape, np.int64)
    idx[inside] = np.arange(int(inside.sum()))
    tris = []
    n0 = idx[:-1, :-1]
    n1 = idx[1:, :-1]
    n2 = idx[:-1, 1:]
    n3 = idx[1:, 1:]
    up = (n0 >= 0) & (n1 >= 0) & (n2 >= 0)
    tris.append(np.stack([n0[up], n1[up], n2[up]], 1))
    dn = (n1 >= 0) & (n3 >= 0) & (n2 >= 0)
    tris.append(np.stack([n1[dn], n3[dn], n2[dn]], 1))
    tri = np.concatenate(tris, 0)
    verts = np.stack([px[inside], py[inside]], 1)
    used = np.zeros(verts.shape[0], bool)
    used[tri.ravel()] = True
    remap = -np.ones(verts.shape[0], np.int64)
    remap[used] = np.arange(int(used.sum()))
    return verts[used], remap[tri].astype(np.int32)


def hex_sheet(nx: int, ny: int, spacing: float = 1.0):
    """nx x ny hexagonal-lattice sheet (rows offset by half a spacing on odd
    rows), CCW triangles.  Returns verts (nx*ny, 2) with the sheet's lower
    left corner at 0, tris (nt, 3), and the bottom-row vertex ids."""
    a = spacing
    h = a * math.sqrt(3.0) / 2.0
    j, i = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    px = a * (i + 0.5 * (j % 2))
    py = h * j
    verts = np.stack([px.ravel(), py.ravel()], 1)
    vid = (j * nx + i)
    tris = []
    for r in range(ny - 1):
        row, nxt = vid[r], vid[r + 1]
        if r % 2 == 0:
            # even row: upper neighbours of i are i-1+.. shifted right
            t1 = np.stack([row[:-1], row[1:], nxt[:-1]], 1)
            t2 = np.stack([row[1:], nxt[1:], nxt[:-1]], 1)
        else:
            t1 = np.stack([row[:-1], row[1:], nxt[1:]], 1)
            t2 = np.stack([row[:-1], nxt[1:], nxt[:-1]], 1)
        tris += [t1, t2]
    tri = np.concatenate(tris, 0).astype(np.int32)
    return verts, tri, vid[0].astype(np.int32)


def signed_areas(verts, tri):
    p0, p1, p2 = verts[tri[:, 0]], verts[tri[:, 1]], verts[tri[:, 2]]
    return 0.5 * ((p1[:, 0] - p0[:, 0]) * (p2[:, 1] - p0[:, 1]) -
                  (p1[:, 1] - p0[:, 1]) * (p2[:, 0] - p0[:, 0]))


# ---------------------------------------------------------------------------
# workload description
# ---------------------------------------------------------------------------
@dataclass
class Elastomer:
    verts: np.ndarray           # (nv, 2) relative to centre
    tris: np.ndarray            # (nt, 3) 0-based
    cx: float
    cy: float
    E: np.ndarray               # (nt,)
    nu: np.ndarray              # (nt,)
    node_mass: float = 1.0
    sigma: float = 1.0
    exclude: bool = True
    pinned: Optional[np.ndarray] = None  # local vertex ids with infinite mass


@dataclass
class Workload:
    name: str
    Lx: float
    Ly: float
    dt: float
    x: np.ndarray
    y: np.ndarray
    vx: np.ndarray
    vy: np.ndarray
    mass: np.ndarray
    sigma: np.ndarray
    elastomers: List[Elastomer] = field(default_factory=list)
    bonds: Optional[tuple] = None       # (i, j, k, r0) over global ids
    eps: float = 1.0
    rc: float = RC_LJ
    ushift: float = U_SHIFT_25
    model: int = 0
    seed: int = 0
    T: float = 1.0

    @property
    def n_solvent(self) -> int:
        return int(self.x.shape[0])

    @property
    def n_nodes(self) -> int:
        return sum(e.verts.shape[0] for e in self.elastomers)

    @property
    def n_total(self) -> int:
        return self.n_solvent + self.n_nodes

    @property
    def n_tri(self) -> int:
        return sum(e.tris.shape[0] for e in self.elastomers)

    @property
    def n_bonds(self) -> int:
        return 0 if self.bonds is None else int(len(self.bonds[0]))

    def global_arrays(self):
        """All particles in caller order (solvent, then each disk's nodes):
        x, y, vx, vy, mass, sigma, excl, tri (global), E, nu, refx, refy."""
        xs, ys, vxs, vys, ms, ss, ex = [self.x], [self.y], [self.vx], [self.vy], [self.mass], \
            [self.sigma], [-np.ones(self.n_solvent, np.int32)]
        tri, E, nu = [], [], []
        off = self.n_solvent
        g = 0
        for el in self.elastomers:
            nv = el.verts.shape[0]
            # FP32 sum, exactly as fm_add_elastomer forms vertex + centre
            px = (el.verts[:, 0].astype(np.float32) + np.float32(el.cx)).astype(np.float64)
            py = (el.verts[:, 1].astype(np.float32) + np.float32(el.cy)).astype(np.float64)
            xs.append(px)
            ys.append(py)
            vxs.append(np.zeros(nv))
            vys.append(np.zeros(nv))
            m = np.full(nv, el.node_mass)
            if el.pinned is not None:
                m[el.pinned] = np.inf
            ms.append(m)
            ss.append(np.full(nv, el.sigma))
            ex.append(np.full(nv, g if el.exclude else -1, np.int32))
            if el.exclude:
                g += 1
            tri.append(el.tris + off)
            E.append(el.E)
            nu.append(el.nu)
            off += nv
        cat = np.concatenate
        tri = cat(tri) if tri else np.zeros((0, 3), np.int32)
        E = cat(E) if E else np.zeros(0)
        nu = cat(nu) if nu else np.zeros(0)
        x, y = cat(xs), cat(ys)
        return dict(x=x, y=y, vx=cat(vxs), vy=cat(vys), mass=cat(ms), sigma=cat(ss),
                    excl=cat(ex).astype(np.int32), tri=tri.astype(np.int32), E=E, nu=nu,
                    refx=x.copy(), refy=y.copy())


def maxwell(rng, n, T, mass):
    """Maxwell velocities at temperature T (kB = 1) with zero net momentum
    (femmd2D/src/particles.cpp:246-261 removes the centre-of-mass velocity)."""
    m = np.where(np.isfinite(mass), mass, 0.0)
    s = np.sqrt(T / np.where(m > 0, m, 1.0))
    vx = rng.normal(0.0, 1.0, n) * s
    vy = rng.normal(0.0, 1.0, n) * s
    vx[m == 0] = 0.0
    vy[m == 0] = 0.0
    M = m.sum()
    vx -= (m * vx).sum() / M * (m > 0)
    vy -= (m * vy).sum() / M * (m > 0)
    return vx, vy


def _rsa(rng, L, radii, gap, max_tries=200000):
    """Random sequential addition of disks in a periodic square box."""
    centres = np.zeros((len(radii), 2))
    placed = 0
    tries = 0
    while placed < len(radii):
        tries += 1
        if tries > max_tries * max(1, len(radii)):
            raise RuntimeError("random sequential addition did not converge")
        c = rng.uniform(0, L, 2)
        if placed:
            d = centres[:placed] - c
            d -= L * np.round(d / L)
            dist = np.sqrt((d * d).sum(1))
            if np.any(dist < radii[:placed] + radii[placed] + gap):
                continue
        centres[placed] = c
        placed += 1
    return centres


def _hex_sites(rng, L, n, radii, gap):
    """n disk centres on a periodic hexagonal lattice of sites (rows offset
    by half a spacing), each jittered inside the slack its radius leaves."""
    rmax = float(np.max(radii))
    smin = 2 * rmax + gap
    ncol = int(math.floor(L / smin))
    nrow = int(math.floor(L / (smin * math.sqrt(3) / 2))) // 2 * 2
    if ncol * nrow < n or ncol < 1 or nrow < 2:
        raise RuntimeError("disks do not fit on a hexagonal lattice of sites")
    sx, sy = L / ncol, L / nrow
    nn = min(sx, math.hypot(sx / 2, sy))  # nearest-neighbour site distance
    j, i = np.meshgrid(np.arange(nrow), np.arange(ncol), indexing="ij")
    px = (i + 0.5 * (j % 2)) * sx
    py = j * sy
    sites = np.stack([px.ravel(), py.ravel()], 1)
    pick = rng.choice(sites.shape[0], n, replace=False)
    out = sites[pick].copy()
    slack = np.maximum(0.0, (nn - gap) / 2 - np.asarray(radii))
    ang = rng.uniform(0, 2 * math.pi, n)
    rad = rng.uniform(0, 1, n) * slack
    out[:, 0] = np.mod(out[:, 0] + rad * np.cos(ang), L)
    out[:, 1] = np.mod(out[:, 1] + rad * np.sin(ang), L)
    return out


def _lattice_solvent(rng, L, n, centres, radii, margin, jitter):
    """n solvent particles on a jittered square lattice over the box, with the
    lattice sites inside any disk (radius + margin) removed; surplus free
    sites are dropped at random.  Returned in raster order."""
    free_area = L * L - np.sum(np.pi * (np.asarray(radii) + margin) ** 2)
    a = math.sqrt(free_area / n) * 0.985
    m = int(math.floor(L / a))
    a = L / m
    while True:
        gx = (np.arange(m) + 0.5) * a
        X, Y = np.meshgrid(gx, gx, indexing="xy")
        X, Y = X.ravel(), Y.ravel()
        keep = np.ones(X.shape[0], bool)
        for (cx, cy), r in zip(centres, radii):
            R = r + margin
            i0, i1 = int(math.floor((cx - R) / a)) - 1, int(math.ceil((cx + R) / a)) + 1
            j0, j1 = int(math.floor((cy - R) / a)) - 1, int(math.ceil((cy + R) / a)) + 1
            ii = np.arange(i0, i1 + 1) % m
            jj = np.arange(j0, j1 + 1) % m
            I, J = np.meshgrid(ii, jj, indexing="xy")
            idx = (J * m + I).ravel()
            dx = X[idx] - cx
            dy = Y[idx] - cy
            dx -= L * np.round(dx / L)
            dy -= L * np.round(dy / L)
            keep[idx[dx * dx + dy * dy < R * R]] = False
        free = np.flatnonzero(keep)
        if free.shape[0] >= n:
            break
        m += 1
        a = L / m
    if free.shape[0] > n:
        drop = rng.choice(free.shape[0], free.shape[0] - n, replace=False)
        mask = np.ones(free.shape[0], bool)
        mask[drop] = False
        free = free[mask]
    x = X[free] + rng.uniform(-jitter, jitter, n)
    y = Y[free] + rng.uniform(-jitter, jitter, n)
    return np.mod(x, L), np.mod(y, L)


def _disk_elastomer(r, cx, cy, E_val, nu_val, node_mass=1.0):
    v, t = hex_disk(r, 1.0)
    nt = t.shape[0]
    E = np.broadcast_to(np.asarray(E_val, float), (nt,)).copy()
    nu = np.broadcast_to(np.asarray(nu_val, float), (nt,)).copy()
    return Elastomer(v, t, float(cx), float(cy), E, nu, node_mass)


# ---------------------------------------------------------------------------
# configurations
# ---------------------------------------------------------------------------
def c1(seed: int = 1, with_disk: bool = True) -> Workload:
    rng = np.random.default_rng(seed)
    L = 74.0
    m = 66
    a = L / m
    g = (np.arange(m) + 0.5) * a
    X, Y = np.meshgrid(g, g, indexing="xy")
    X = X.ravel() + rng.uniform(-0.1, 0.1, m * m)
    Y = Y.ravel() + rng.uniform(-0.1, 0.1, m * m)
    d2 = (X - L / 2) ** 2 + (Y - L / 2) ** 2
    keep = np.flatnonzero(d2 >= 11.0 ** 2)
    assert keep.shape[0] >= 4000
    keep = keep[:4000]
    x, y = np.mod(X[keep], L), np.mod(Y[keep], L)
    els = [_disk_elastomer(10.0, L / 2, L / 2, 50.0, 0.3)] if with_disk else []
    n = 4000 + sum(e.verts.shape[0] for e in els)
    mass = np.ones(4000)
    vx, vy = maxwell(rng, n, 1.0, np.ones(n))
    w = Workload("c1", L, L, 0.005, x, y, vx[:4000], vy[:4000], mass, np.ones(4000), els,
                 seed=seed)
    w._node_v = (vx[4000:], vy[4000:])
    return w


def c2(seed: int = 2, scale: float = 1.0, n_solvent: int = 1048576, n_disks: int = 1024,
       radius: float = 5.0, density: float = 0.8) -> Workload:
    rng = np.random.default_rng(seed)
    ns = max(16, int(round(n_solvent * scale)))
    nd = int(round(n_disks * scale))
    radii = np.full(nd, radius)
    disk_area = nd * math.pi * radius * radius
    L = math.sqrt(ns / density + disk_area)
    centres = _rsa(rng, L, radii, 1.0) if nd else np.zeros((0, 2))
    x, y = _lattice_solvent(rng, L, ns, centres, radii, 1.0, 0.1)
    Es = rng.uniform(20.0, 100.0, nd)
    els = [_disk_elastomer(radius, c[0], c[1], Es[k], 0.3) for k, c in enumerate(centres)]
    n = ns + sum(e.verts.shape[0] for e in els)
    vx, vy = maxwell(rng, n, 1.0, np.ones(n))
    w = Workload("c2", L, L, 0.005, x, y, vx[:ns], vy[:ns], np.ones(ns), np.ones(ns), els,
                 seed=seed)
    w._node_v = (vx[ns:], vy[ns:])
    return w


def c3(seed: int = 3, scale: float = 1.0) -> Workload:
    rng = np.random.default_rng(seed)
    ns = max(64, int(round(600000 * scale)))
    nd = max(1, int(round(4096 * scale)))
    radii = np.exp(rng.uniform(math.log(4.0), math.log(8.0), nd))
    L = math.sqrt(ns / 0.75 + np.sum(np.pi * radii ** 2))
    # disk + anchor shell (rim + 1) occupies 36-60 % of the box: random
    # sequential addition jams there, so disks sit on a jittered hexagonal
    # lattice of sites whose spacing clears the largest disk and its shell
    while True:  # small test scales may need a slightly larger box
        try:
            centres = _hex_sites(rng, L, nd, radii + 1.5, 0.5)
            break
        except RuntimeError:
            L *= 1.02
    els = []
    anchors = []
    for k, (c, r) in enumerate(zip(centres, radii)):
        v, t = hex_disk(r, 1.0)
        nt = t.shape[0]
        E = np.exp(rng.uniform(math.log(10.0), math.log(200.0), nt))
        nu = rng.uniform(0.25, 0.45, nt)
        el = Elastomer(v, t, float(c[0]), float(c[1]), E, nu, 1.0)
        els.append(el)
        rv = np.sqrt((v * v).sum(1))
        rim = rv.max()
        rot = rng.uniform(0, 2 * math.pi)
        for q in range(8):
            th = 2 * math.pi * q / 8 + rot
            ax, ay = (rim + 1.0) * math.cos(th), (rim + 1.0) * math.sin(th)
            node = int(np.argmin((v[:, 0] - ax) ** 2 + (v[:, 1] - ay) ** 2))
            # anchor at distance 1 from its rim node, radially outward
            nx, ny = v[node]
            nr = math.hypot(nx, ny) or 1.0
            anchors.append((k, node, c[0] + nx + nx / nr, c[1] + ny + ny / nr))
    na = len(anchors)
    # solvent keeps >= 1 sigma from the anchor shell (rim + 1)
    xb, yb = _lattice_solvent(rng, L, ns - na, centres, radii + 2.0, 0.0, 0.1)
    ax = np.array([a[2] for a in anchors]) % L
    ay = np.array([a[3] for a in anchors]) % L
    x = np.concatenate([xb, ax])
    y = np.concatenate([yb, ay])
    node_off = np.cumsum([0] + [e.verts.shape[0] for e in els])[:-1] + ns
    bi = np.arange(ns - na, ns, dtype=np.int32)
    bj = np.array([node_off[a[0]] + a[1] for a in anchors], np.int32)
    n = ns + sum(e.verts.shape[0] for e in els)
    vx, vy = maxwell(rng, n, 1.0, np.ones(n))
    w = Workload("c3", L, L, 0.002, x, y, vx[:ns], vy[:ns], np.ones(ns), np.ones(ns), els,
                 bonds=(bi, bj, np.full(na, 100.0), np.full(na, 1.0)), seed=seed)
    w._node_v = (vx[ns:], vy[ns:])
    return w


def c4(seed: int = 4, n_replicas: int = 512, n_min: int = 256, n_max: int = 32768,
       scale: float = 1.0) -> List[Workload]:
    rng = np.random.default_rng(seed)
    out = []
    nr = max(1, int(round(n_replicas * scale)))
    for r in range(nr):
        n = int(round(math.exp(rng.uniform(math.log(n_min), math.log(n_max)))))
        rho = rng.uniform(0.6, 0.9)
        T = rng.uniform(0.5, 2.0)
        rdisk = rng.uniform(6.0, 15.0) if r % 2 == 0 else 0.0
        radii = [rdisk] if rdisk > 0 else []
        L = math.sqrt(n / rho + (math.pi * rdisk * rdisk if rdisk else 0.0))
        if rdisk and L < 2 * rdisk + 8:
            L = 2 * rdisk + 8
        centres = np.array([[L / 2, L / 2]]) if rdisk else np.zeros((0, 2))
        sub = np.random.default_rng(seed * 100003 + r)
        x, y = _lattice_solvent(sub, L, n, centres, radii, 1.0, 0.1)
        els = [_disk_elastomer(rdisk, L / 2, L / 2, 50.0, 0.3)] if rdisk else []
        ntot = n + sum(e.verts.shape[0] for e in els)
        vx, vy = maxwell(sub, ntot, T, np.ones(ntot))
        w = Workload(f"c4_{r}", L, L, 0.005, x, y, vx[:n], vy[:n], np.ones(n), np.ones(n), els,
                     seed=seed, T=T)
        w._node_v = (vx[n:], vy[n:])
        out.append(w)
    return out


def fourier_field(rng, x, y, Lx, Ly, modes=16):
    """Smooth random field: sum of `modes` random Fourier modes, normalised
    to [-1, 1] over the given points."""
    f = np.zeros_like(x)
    for _ in range(modes):
        kx, ky = rng.integers(1, 5), rng.integers(1, 5)
        ph = rng.uniform(0, 2 * math.pi)
        amp = rng.normal()
        f += amp * np.cos(2 * math.pi * (kx * x / Lx + ky * y / Ly) + ph)
    f -= f.min()
    f /= max(f.max(), 1e-12)
    return 2 * f - 1


def c5(seed: int = 5, scale: float = 1.0, nx: int = 128, ny: int = 256) -> Workload:
    rng = np.random.default_rng(seed)
    if scale != 1.0:
        nx = max(4, int(round(nx * math.sqrt(scale))))
        ny = max(4, int(round(ny * math.sqrt(scale))))
    ns = max(64, int(round(200000 * scale)))
    verts, tri, bottom = hex_sheet(nx, ny, 1.0)
    w_sheet = verts[:, 0].max() + 1.0
    h_sheet = verts[:, 1].max() + 1.0
    L = math.sqrt(ns / 0.7 + w_sheet * h_sheet)
    L = max(L, max(w_sheet, h_sheet) + 6.0)
    ox, oy = (L - w_sheet) / 2 + 0.5, (L - h_sheet) / 2 + 0.5
    verts = verts + np.array([ox, oy])
    c = np.array([L / 2, L / 2])
    cent = verts[tri].mean(1)
    field1 = fourier_field(rng, cent[:, 0], cent[:, 1], L, L)
    E = np.clip(70.0 + 50.0 * field1, 20.0, 120.0)
    # solvent: lattice sites outside the sheet rectangle (+1 margin)
    m = int(math.ceil(math.sqrt(ns * 1.6)))
    while True:
        g = (np.arange(m) + 0.5) * (L / m)
        X, Y = np.meshgrid(g, g, indexing="xy")
        X, Y = X.ravel(), Y.ravel()
        inside = (X > ox - 1.5) & (X < ox + w_sheet + 0.5) & (Y > oy - 1.5) & (Y < oy + h_sheet + 0.5)
        free = np.flatnonzero(~inside)
        if free.shape[0] >= ns:
            break
        m += 1
    free = np.sort(rng.choice(free, ns, replace=False))
    x = X[free] + rng.uniform(-0.1, 0.1, ns)
    y = Y[free] + rng.uniform(-0.1, 0.1, ns)
    el = Elastomer(verts - c, tri, float(c[0]), float(c[1]), E, np.full(tri.shape[0], 0.3), 1.0,
                   pinned=bottom)
    n = ns + verts.shape[0]
    mass = np.ones(n)
    mass[ns + bottom] = np.inf
    vx, vy = maxwell(rng, n, 1.0, mass)
    w = Workload("c5", L, L, 0.005, np.mod(x, L), np.mod(y, L), vx[:ns], vy[:ns], np.ones(ns),
                 np.ones(ns), [el], eps=1.0, rc=RC_WCA, ushift=1.0, seed=seed)
    w._node_v = (vx[ns:], vy[ns:])
    w._field_rng_state = rng.bit_generator.state
    return w


def c5_target_E(w: Workload) -> np.ndarray:
    """The per-element Young's moduli of C5's target rollout: the workload's
    field E times (1 + 0.1 field2), field2 a second smooth random field of
    16 Fourier modes drawn from the same seed (after the first), at the
    element centroids."""
    el = w.elastomers[0]
    rng = np.random.default_rng()
    rng.bit_generator.state = w._field_rng_state
    cent = (el.verts[el.tris] + np.array([el.cx, el.cy])).mean(1)
    field2 = fourier_field(rng, cent[:, 0], cent[:, 1], w.Lx, w.Ly)
    return np.asarray(el.E, np.float64) * (1.0 + 0.1 * field2)


def with_params(w: Workload, E=None, eps=None, sigma=None) -> Workload:
    """A copy of workload w with other per-element Young's moduli E (of its
    single elastomer), pair eps and a common sigma for every particle."""
    import copy
    v = copy.copy(w)
    v.elastomers = [copy.copy(el) for el in w.elastomers]
    if E is not None:
        v.elastomers[0].E = np.asarray(E, np.float64).copy()
    if eps is not None:
        v.eps = float(eps)
    if sigma is not None:
        v.sigma = np.full(w.n_solvent, float(sigma))
        for el in v.elastomers:
            el.sigma = float(sigma)
    return v


# ---------------------------------------------------------------------------
# engine construction
# ---------------------------------------------------------------------------
def build_system(w: Workload, device: int = 0, commit: bool = True, skin=None):
    """A femmd_old_amd.System holding workload w (caller order: solvent, then
    each elastomer's nodes, as global_arrays() lists them).  skin: the
    neighbour-list skin (None keeps the engine default 0.6; 0 walks the
    cells every step)."""
    from . import System
    s = System(w.Lx, w.Ly, dt=w.dt, device=device)
    if skin is not None:
        s.set_skin(skin)
    s.set_pair_potential([w.eps], [w.rc], [w.ushift])
    s.add_particles(w.x, w.y, w.vx, w.vy, w.mass, w.sigma)
    for el in w.elastomers:
        first = s.add_elastomer(el.verts.ravel(), el.tris.ravel(), el.cx, el.cy, 0, 0, el.sigma,
                                0.0, el.exclude, index_base=0, node_mass=el.node_mass,
                                E=el.E, nu=el.nu)
    if w.elastomers:
        s.set_youngs_modulus(float(w.elastomers[0].E[0]), 0.3, w.model)
    if w.bonds is not None:
        s.add_bonds(*w.bonds)
    n = w.n_total
    if w.elastomers:
        vx, vy = node_velocities(w)
        s.set_state(None, None, vx, vy)
        m = w.global_arrays()["mass"]
        pinned = np.flatnonzero(~np.isfinite(m))
        if pinned.size:
            s.set_mass(pinned, np.inf)
    assert s.N == n
    if commit:
        s.allocate_cuda()
    return s


def node_velocities(w: Workload):
    """Caller-order initial velocities including elastomer nodes."""
    ga = w.global_arrays()
    vx, vy = ga["vx"].copy(), ga["vy"].copy()
    nv = getattr(w, "_node_v", None)
    if nv is not None:
        vx[w.n_solvent:] = nv[0]
        vy[w.n_solvent:] = nv[1]
    m = ga["mass"]
    vx[~np.isfinite(m)] = 0
    vy[~np.isfinite(m)] = 0
    return vx, vy
