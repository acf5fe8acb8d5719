"""Synthetic scenes for the benchmark configurations and for the checks that
follow the reference's own test scenes.

* :func:`haengende_roehre` — the plain tube: an open cylinder hanging from a
  held ring, gravity only (closed form for its static elongation).
* :func:`roehre_auf_zylinder` — the reference's tube acceptance scene
  (roehre.rs:49-239): a reduced rectangle sewn into a tube around a cylinder.
* :func:`streifen` — the 20 x 4 cm strip of tests/kontinuum.rs:343-371.
* :func:`drape` — random convex-ish panels sewn into strips and dropped onto an
  ellipsoid-cylinder body given as a grid SDF.
* :func:`koerper_sdf` — that body's signed distance sampled on a regular grid.
* :func:`netz_zylinder`, :func:`netz_ellipsoid`, :func:`koerper_netz` — closed
  triangle meshes as bodies (the reference's mesh body, netzkoerper.rs).

Meshes are built here (structured grids, or a jittered hexagonal lattice
triangulated by Delaunay); the reference meshes its patterns with
klon3d-mesh, which this build does not rebuild.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List, Optional, Tuple

import numpy as np

from .material import RADICAL_RUBBER_04, Material, gewebe, latex
from .szene import GitterSdf, Netz, Simparameter, Szene, Teil, Teilnaht, Zylinder, bauen

# ---------------------------------------------------------------------------
# hanging tube
# ---------------------------------------------------------------------------


@dataclass
class HaengendeRoehre:
    radius_m: float = 0.10
    hoehe_m: float = 0.30
    segmente: int = 63
    ringe: int = 31
    material: Material = latex(1.0, 0.5e-3, 950.0, 1e-6)

    def gummi_dehnung_soll(self, g: float = 9.81) -> float:
        """Static elongation of the hanging tube (m).

        Uniaxial membrane with the Green strain: first Piola stress
        P = lambda * E * (lambda**2 - 1) / 2 per rest width must carry the weight
        below, rho_A * g * Z (Z = rest distance from the free end); nu = 0 keeps
        the hoop stretch at one. The elongation is the integral of lambda - 1
        over the rest height, plus the stretch of the ring springs, which carry
        the whole weight in parallel.
        """
        e = self.material.kette_n_m
        rho = self.material.flaechengewicht_kg_m2
        h = self.hoehe_m
        nq = 4000
        tot = 0.0
        for q in range(nq):
            z = (q + 0.5) / nq * h
            rhs = rho * g * z
            lam = 1.0
            for _ in range(60):  # Newton on lam*E*(lam^2-1)/2 - rhs
                fval = lam * e * (lam * lam - 1.0) / 2.0 - rhs
                dval = e * (3.0 * lam * lam - 1.0) / 2.0
                lam -= fval / dval
            tot += (lam - 1.0) * (h / nq)
        return tot

    def ring_dehnung(self, naht_k: float, g: float = 9.81) -> float:
        gewicht = self.material.flaechengewicht_kg_m2 * g * 2.0 * math.pi * self.radius_m * self.hoehe_m
        return gewicht / (self.segmente * naht_k)


def haengende_roehre(cfg: Optional[HaengendeRoehre] = None, bild_dt: float = 1.0 / 240.0,
                     schwerkraft=(0.0, 0.0, -9.81)) -> Szene:
    """Open cylinder, ``segmente`` around x ``ringe`` along the axis, periodic
    around (no seam), warp along the axis, top ring held by seam springs of rest
    length 0 to a ring of held points (mass 0)."""
    c = cfg or HaengendeRoehre()
    ns, nr = c.segmente, c.ringe
    # hoop spacing = the chord, so the rolled tube starts free of in-plane stress
    du = 2.0 * c.radius_m * math.sin(math.pi / ns) * 1e3   # mm around
    dz = c.hoehe_m / (nr - 1) * 1e3                    # mm along
    idx = lambda r, s: r * ns + s
    tris, uv = [], []
    for r in range(nr - 1):
        for s in range(ns):
            s1 = (s + 1) % ns
            a, b, cc, d = idx(r, s), idx(r, s1), idx(r + 1, s1), idx(r + 1, s)
            ua, ub = s * du, (s + 1) * du
            va, vb = -r * dz, -(r + 1) * dz
            tris.append((a, b, cc)); uv.append(((ua, va), (ub, va), (ub, vb)))
            tris.append((a, cc, d)); uv.append(((ua, va), (ub, vb), (ua, vb)))
    pts = np.array([(s * du, -r * dz) for r in range(nr) for s in range(ns)], dtype=np.float64)
    teil = Teil("Rohr", pts, np.array(tris), fadenlauf=(0.0, 1.0), material=c.material, reduktion=0.0,
                dreieck_uv_mm=np.array(uv))

    def lage(_ti, p_mm):
        out = np.empty((p_mm.shape[0], 3))
        for k in range(p_mm.shape[0]):
            s = int(round(p_mm[k, 0] / du)) % ns
            r = int(round(-p_mm[k, 1] / dz))
            w = 2.0 * math.pi * s / ns
            out[k] = (c.radius_m * math.cos(w), c.radius_m * math.sin(w), -r * dz * 1e-3)
        return out

    ring = np.array([(c.radius_m * math.cos(2 * math.pi * s / ns), c.radius_m * math.sin(2 * math.pi * s / ns), 0.0)
                     for s in range(ns)])
    extra = {"x": ring, "naht": [(idx(0, s), s) for s in range(ns)]}
    param = Simparameter(bild_dt=bild_dt, schwerkraft=tuple(schwerkraft), daempfung=0.0)
    return bauen([teil], [], lage, [], param, extra_punkte=extra)


# ---------------------------------------------------------------------------
# structured rectangle (the reference meshes it with klon3d-mesh)
# ---------------------------------------------------------------------------


def rechteck_netz(breite_mm: float, hoehe_mm: float, kante_mm: float) -> Tuple[np.ndarray, np.ndarray, int, int]:
    """Rectangle meshed as a grid of right triangles with alternating diagonals,
    ~``kante_mm`` edges. Returns (points (n,2) mm, triangles, columns, rows)."""
    nx = max(1, int(round(breite_mm / kante_mm)))
    ny = max(1, int(round(hoehe_mm / kante_mm)))
    xs = np.linspace(0.0, breite_mm, nx + 1)
    ys = np.linspace(0.0, hoehe_mm, ny + 1)
    pts = np.array([(x, y) for y in ys for x in xs], dtype=np.float64)
    k = lambda i, j: j * (nx + 1) + i
    tris = []
    for j in range(ny):
        for i in range(nx):
            a, b, c, d = k(i, j), k(i + 1, j), k(i + 1, j + 1), k(i, j + 1)
            if (i + j) % 2 == 0:
                tris += [(a, b, c), (a, c, d)]
            else:
                tris += [(a, b, d), (b, c, d)]
    return pts, np.array(tris), nx, ny


def streifen(material: Material = RADICAL_RUBBER_04, kante_mm: float = 10.0) -> Szene:
    """The 20 x 4 cm strip of tests/kontinuum.rs:343-371: flat in z = 0, warp along x."""
    pts, tris, _, _ = rechteck_netz(200.0, 40.0, kante_mm)
    teil = Teil("Streifen", pts, tris, fadenlauf=(1.0, 0.0), material=material, reduktion=0.0)
    lage = lambda _ti, p: np.column_stack([p[:, 0] * 1e-3, p[:, 1] * 1e-3, np.zeros(p.shape[0])])
    param = Simparameter(bild_dt=1.0 / 30.0, schwerkraft=(0.0, 0.0, 0.0), daempfung=5.0)
    return bauen([teil], [], lage, [], param)


@dataclass
class Roehre:
    """roehre.rs:49-84 (defaults roehre.rs:72-84)."""
    zylinder_radius_m: float = 0.09
    hoehe_m: float = 0.20
    anfangsabstand_m: float = 0.005
    kante_mm: float = 10.0
    material: Material = RADICAL_RUBBER_04
    reduktion: float = 0.055

    def mittelradius_m(self) -> float:
        return self.zylinder_radius_m + self.material.dicke_m / 2.0

    def breite_m(self) -> float:
        return 2.0 * math.pi * self.mittelradius_m()

    def dehnung_soll(self) -> float:
        """roehre.rs:128-130."""
        return 1.0 / (1.0 - self.reduktion) - 1.0


def param_ruhe() -> Simparameter:
    """roehre.rs:291-293."""
    return Simparameter(schwerkraft=(0.0, 0.0, 0.0), daempfung=5.0)


def roehre_auf_zylinder(r: Optional[Roehre] = None, param: Optional[Simparameter] = None) -> Szene:
    """roehre.rs:221-239: the rectangle rolled on a cylinder 5 mm larger than its
    seat, long sides sewn (rest 0), the cylinder as the only body."""
    r = r or Roehre()
    param = param or param_ruhe()
    b_mm, h_mm = r.breite_m() * 1e3, r.hoehe_m * 1e3
    pts, tris, nx, ny = rechteck_netz(b_mm, h_mm, r.kante_mm)
    r0 = r.mittelradius_m() + r.anfangsabstand_m
    je_m = 2.0 * math.pi / r.breite_m()

    def lage(_ti, p):
        u, y = p[:, 0] * 1e-3, p[:, 1] * 1e-3
        w = u * je_m
        return np.column_stack([r0 * np.cos(w), r0 * np.sin(w), -y])

    teil = Teil("Rohr", pts, tris, fadenlauf=(1.0, 0.0), material=r.material, reduktion=r.reduktion)
    # right edge B->C (x = b) against left edge A->D (x = 0), bottom to top
    paare = [(j * (nx + 1) + nx, j * (nx + 1)) for j in range(ny + 1)]
    koerper = [Zylinder((0.0, 0.0, 0.0), (0.0, 0.0, 1.0), r.zylinder_radius_m)]
    return bauen([teil], [[Teilnaht(0, 0, paare)]], lage, koerper, param)


# ---------------------------------------------------------------------------
# drape: panels and body
# ---------------------------------------------------------------------------


@dataclass
class Koerperform:
    """Ellipsoid (torso) united with a capped vertical cylinder (neck)."""
    ell_c: Tuple[float, float, float] = (0.0, 0.0, -0.08)
    ell_r: Tuple[float, float, float] = (0.20, 0.14, 0.26)
    cyl_c: Tuple[float, float] = (0.0, 0.0)
    cyl_r: float = 0.06
    cyl_z0: float = 0.10
    cyl_z1: float = 0.32


def _ellipsoid_sdf_np(p: np.ndarray, c, r) -> np.ndarray:
    """Exact signed distance to an axis-aligned ellipsoid by bisection on the
    Lagrange multiplier (Eberly, 'Distance from a point to an ellipse, an
    ellipsoid, or a hyperellipsoid'), vectorised over points."""
    y = np.abs(p - np.asarray(c))
    e = np.asarray(r, dtype=np.float64)
    order = np.argsort(-e)
    e = e[order]
    y = y[:, order]
    z = y / e
    g = np.sum(z * z, axis=1) - 1.0
    r0 = (e[0] / e[2]) ** 2
    r1 = (e[1] / e[2]) ** 2
    n0, n1 = r0 * z[:, 0], r1 * z[:, 1]
    s0 = z[:, 2] - 1.0
    s1 = np.where(g < 0.0, 0.0, np.sqrt(n0 * n0 + n1 * n1 + z[:, 2] ** 2) - 1.0)
    # also valid for points on axis planes: F(s) stays monotone for s > -1
    for _ in range(80):
        s = 0.5 * (s0 + s1)
        ra, rb, rc = n0 / (s + r0), n1 / (s + r1), z[:, 2] / (s + 1.0)
        gg = ra * ra + rb * rb + rc * rc - 1.0
        s0 = np.where(gg > 0.0, s, s0)
        s1 = np.where(gg > 0.0, s1, s)
    s = 0.5 * (s0 + s1)
    x0 = r0 * y[:, 0] / (s + r0)
    x1 = r1 * y[:, 1] / (s + r1)
    x2 = y[:, 2] / (s + 1.0)
    d = np.sqrt((x0 - y[:, 0]) ** 2 + (x1 - y[:, 1]) ** 2 + (x2 - y[:, 2]) ** 2)
    return np.where(g < 0.0, -d, d)


def _zylinder_sdf_np(p: np.ndarray, k: Koerperform) -> np.ndarray:
    dr = np.sqrt((p[:, 0] - k.cyl_c[0]) ** 2 + (p[:, 1] - k.cyl_c[1]) ** 2) - k.cyl_r
    dz = np.maximum(k.cyl_z0 - p[:, 2], p[:, 2] - k.cyl_z1)
    a, b = np.maximum(dr, 0.0), np.maximum(dz, 0.0)
    return np.minimum(np.maximum(dr, dz), 0.0) + np.sqrt(a * a + b * b)


def koerper_sdf(aufloesung: int = 256, spacing: float = 0.003, form: Optional[Koerperform] = None,
                band: float = 0.03) -> GitterSdf:
    """The body's signed distance on an ``aufloesung``^3 grid of ``spacing``
    (node-centred, centred on the origin). Exact within ``band`` of the
    surface; farther out a lower bound of the same sign (contacts never query
    there)."""
    k = form or Koerperform()
    n = aufloesung
    o = -0.5 * (n - 1) * spacing
    ax = o + spacing * np.arange(n)
    werte = np.empty((n, n, n), dtype=np.float32)
    gx, gy = np.meshgrid(ax, ax, indexing="xy")
    flat = np.column_stack([gx.ravel(), gy.ravel(), np.zeros(gx.size)])
    e = np.asarray(k.ell_r)
    for iz in range(n):
        flat[:, 2] = ax[iz]
        # cheap bound: scaled algebraic distance times the smallest semi-axis
        q = (flat - np.asarray(k.ell_c)) / e
        approx = (np.sqrt(np.sum(q * q, axis=1)) - 1.0) * e.min()
        d_ell = approx.copy()
        nah = np.abs(approx) < band
        if np.any(nah):
            d_ell[nah] = _ellipsoid_sdf_np(flat[nah], k.ell_c, k.ell_r)
        d = np.minimum(d_ell, _zylinder_sdf_np(flat, k))
        werte[iz] = d.reshape(n, n).astype(np.float32)
    return GitterSdf(werte=werte, origin=(o, o, o), spacing=spacing)


def _panel_mesh(ecken: np.ndarray, kante: float, rng: np.random.Generator, sagitta: float,
                seiten_n: List[int]):
    """Mesh one convex panel: boundary points along slightly bulged edges
    (``seiten_n[k]`` segments on edge k), a jittered hexagonal lattice inside,
    Delaunay, then Laplacian smoothing of the interior. Returns points (mm),
    triangles and, per edge, the boundary point indices from corner k to k+1."""
    from scipy.spatial import Delaunay

    m = ecken.shape[0]
    centroid = ecken.mean(axis=0)
    rand, seiten = [], []
    for k in range(m):
        a, b = ecken[k], ecken[(k + 1) % m]
        nseg = seiten_n[k]
        t = np.linspace(0.0, 1.0, nseg + 1)[:-1]
        d = b - a
        nrm = np.array([d[1], -d[0]]) / np.linalg.norm(d)
        if np.dot(nrm, (a + b) / 2 - centroid) < 0:
            nrm = -nrm
        pts = a[None, :] + t[:, None] * d[None, :] + (4.0 * sagitta * t * (1 - t))[:, None] * nrm[None, :]
        start = len(rand)
        rand.extend(pts.tolist())
        idx = list(range(start, start + nseg)) + [((start + nseg) if k < m - 1 else 0)]
        seiten.append(idx)
    rand = np.array(rand)
    nrand = rand.shape[0]
    # interior lattice
    lo, hi = ecken.min(axis=0), ecken.max(axis=0)
    ang = rng.uniform(0, math.pi)
    ca, sa = math.cos(ang), math.sin(ang)
    span = np.linalg.norm(hi - lo)
    nn = int(span / kante) + 3
    ii, jj = np.meshgrid(np.arange(-nn, nn + 1), np.arange(-nn, nn + 1))
    lx = (ii + 0.5 * (jj % 2)) * kante
    ly = jj * kante * math.sqrt(3) / 2
    px = centroid[0] + ca * lx - sa * ly + rng.uniform(-0.5, 0.5) * kante
    py = centroid[1] + sa * lx + ca * ly + rng.uniform(-0.5, 0.5) * kante
    cand = np.column_stack([px.ravel(), py.ravel()])
    keep = np.ones(cand.shape[0], dtype=bool)
    for k in range(m):
        a, b = ecken[k], ecken[(k + 1) % m]
        d = b - a
        nrm = np.array([-d[1], d[0]]) / np.linalg.norm(d)
        if np.dot(nrm, centroid - a) < 0:
            nrm = -nrm
        keep &= (cand - a) @ nrm > 0.6 * kante
    inner = cand[keep]
    pts = np.vstack([rand, inner])
    for _ in range(4):
        tri = Delaunay(pts).simplices
        # smoothing of interior points towards the mean of their neighbours
        nb_sum = np.zeros_like(pts)
        nb_cnt = np.zeros(pts.shape[0])
        for c in range(3):
            for c2 in range(3):
                if c != c2:
                    np.add.at(nb_sum, tri[:, c], pts[tri[:, c2]])
                    np.add.at(nb_cnt, tri[:, c], 1.0)
        mean = nb_sum / np.maximum(nb_cnt, 1.0)[:, None]
        pts[nrand:] = 0.5 * pts[nrand:] + 0.5 * mean[nrand:]
    tri = Delaunay(pts).simplices
    a = pts[tri]
    area = 0.5 * ((a[:, 1, 0] - a[:, 0, 0]) * (a[:, 2, 1] - a[:, 0, 1]) - (a[:, 2, 0] - a[:, 0, 0]) * (a[:, 1, 1] - a[:, 0, 1]))
    tri = tri[np.abs(area) > 1e-6 * kante * kante]
    # orient counter-clockwise
    a = pts[tri]
    area = 0.5 * ((a[:, 1, 0] - a[:, 0, 0]) * (a[:, 2, 1] - a[:, 0, 1]) - (a[:, 2, 0] - a[:, 0, 0]) * (a[:, 1, 1] - a[:, 0, 1]))
    neg = area < 0
    tri[neg] = tri[neg][:, [0, 2, 1]]
    return pts, tri.astype(np.int64), seiten


def min_winkel_grad(pts: np.ndarray, tri: np.ndarray) -> np.ndarray:
    a = pts[tri]
    out = []
    for k in range(3):
        u = a[:, (k + 1) % 3] - a[:, k]
        v = a[:, (k + 2) % 3] - a[:, k]
        c = np.sum(u * v, axis=1) / (np.linalg.norm(u, axis=1) * np.linalg.norm(v, axis=1))
        out.append(np.degrees(np.arccos(np.clip(c, -1, 1))))
    return np.min(np.array(out), axis=0)


@dataclass
class DrapeKonfig:
    seed: int = 20261203
    spalten: int = 7
    zeilen: int = 4
    zelle_mm: float = 352.0
    kante_mm: float = 10.0
    luecke_mm: float = 3.0
    jitter_mm: float = 30.0
    sagitta_mm: float = 1.5
    hoehe_m: float = 0.34
    material: Optional[Material] = None
    bild_dt: float = 1.0 / 120.0
    # cloth shift in the plane: the 7 x 4 grid centred on the body would put
    # the neck under the unsewn border between rows 1 and 2 (a 6 mm gap that
    # the neck pushes through and the cloth catches on); half a cell in -y
    # puts it under the middle of a panel
    versatz_mm: Tuple[float, float] = (0.0, -176.0)


def drape_material() -> Material:
    """Woven orthotropic fabric: warp/weft E 5/3 MPa, shear 0.5 MPa at 0.5 mm
    (2500/1500/250 N/m), 0.2 kg/m^2, bending warp/weft/bias 1.2/0.8/0.9 uN*m,
    friction 0.3."""
    return gewebe(5.0, 3.0, 0.5, 0.5e-3, 0.20, 1.2e-6, 0.8e-6, 0.9e-6, 0.3)


def netz_zylinder(radius: float, z_unten: float, z_oben: float, segmente: int, ringe: int) -> Tuple[np.ndarray, np.ndarray]:
    """Closed cylinder (axis z) as a triangle mesh with outward normals:
    ``segmente`` points per ring, ``ringe`` sections between ``z_unten`` and
    ``z_oben``, caps as fans about a centre point (netzkoerper.rs:934-964; the
    same points and triangles in the same order)."""
    s, r = max(segmente, 3), max(ringe, 1)
    p = []
    for i in range(r + 1):
        z = z_unten + (z_oben - z_unten) * i / r
        for j in range(s):
            w = 2.0 * math.pi * j / s
            p.append((radius * math.cos(w), radius * math.sin(w), z))
    unten = len(p)
    p.append((0.0, 0.0, z_unten))
    oben = len(p)
    p.append((0.0, 0.0, z_oben))

    def i_(i, j):
        return i * s + j % s

    d = []
    for i in range(r):
        for j in range(s):
            a, b, c, e = i_(i, j), i_(i, j + 1), i_(i + 1, j + 1), i_(i + 1, j)
            d.append((a, b, c))
            d.append((a, c, e))
    for j in range(s):
        d.append((unten, i_(0, j + 1), i_(0, j)))
        d.append((oben, i_(r, j), i_(r, j + 1)))
    return np.array(p, dtype=np.float64), np.array(d, dtype=np.uint32)


def netz_ellipsoid(c, r, segmente: int, ringe: int) -> Tuple[np.ndarray, np.ndarray]:
    """Axis-aligned ellipsoid (centre ``c``, semi-axes ``r``) as a closed
    latitude-longitude mesh with outward normals: ``ringe`` - 1 rings of
    ``segmente`` points plus the two poles."""
    s, n = max(segmente, 3), max(ringe, 2)
    p = [(c[0], c[1], c[2] - r[2])]
    for i in range(1, n):
        th = math.pi * i / n - 0.5 * math.pi
        for j in range(s):
            w = 2.0 * math.pi * j / s
            p.append((c[0] + r[0] * math.cos(th) * math.cos(w), c[1] + r[1] * math.cos(th) * math.sin(w),
                      c[2] + r[2] * math.sin(th)))
    p.append((c[0], c[1], c[2] + r[2]))
    sued, nord = 0, len(p) - 1

    def i_(i, j):  # ring i in 1..n-1
        return 1 + (i - 1) * s + j % s

    d = []
    for j in range(s):
        d.append((sued, i_(1, j + 1), i_(1, j)))
    for i in range(1, n - 1):
        for j in range(s):
            a, b, cc, e = i_(i, j), i_(i, j + 1), i_(i + 1, j + 1), i_(i + 1, j)
            d.append((a, b, cc))
            d.append((a, cc, e))
    for j in range(s):
        d.append((nord, i_(n - 1, j), i_(n - 1, j + 1)))
    return np.array(p, dtype=np.float64), np.array(d, dtype=np.uint32)


def koerper_netz(form: Optional[Koerperform] = None, kante: float = 0.01, suchtiefe: float = 0.01) -> List[Netz]:
    """The drape body as two closed triangle meshes of about ``kante`` edge
    length: the torso ellipsoid and the neck cylinder, each its own body (the
    reference steps a list of bodies, the last contact found wins)."""
    k = form or Koerperform()
    umfang = 2.0 * math.pi * max(k.ell_r[0], k.ell_r[1])
    seg = max(8, int(round(umfang / kante)))
    ringe = max(4, int(round(math.pi * k.ell_r[2] / kante)))
    ep, ed = netz_ellipsoid(k.ell_c, k.ell_r, seg, ringe)
    zseg = max(8, int(round(2.0 * math.pi * k.cyl_r / kante)))
    zr = max(1, int(round((k.cyl_z1 - k.cyl_z0) / kante)))
    zp, zd = netz_zylinder(k.cyl_r, k.cyl_z0, k.cyl_z1, zseg, zr)
    zp[:, 0] += k.cyl_c[0]
    zp[:, 1] += k.cyl_c[1]
    return [Netz(ep, ed, suchtiefe), Netz(zp, zd, suchtiefe)]


def drape(cfg: Optional[DrapeKonfig] = None, sdf: Optional[GitterSdf] = None, mit_koerper: bool = True) -> Tuple[Szene, dict]:
    """Panels from a jittered grid of convex cells, sewn along every vertical
    cell border (one seam per horizontally adjacent pair), laid flat above the
    body and dropped. Returns the scene and a report (counts, min angle)."""
    cfg = cfg or DrapeKonfig()
    mat = cfg.material or drape_material()
    rng = np.random.default_rng(cfg.seed)
    nc, nr, cw = cfg.spalten, cfg.zeilen, cfg.zelle_mm
    x0 = -0.5 * nc * cw + cfg.versatz_mm[0]
    y0 = -0.5 * nr * cw + cfg.versatz_mm[1]
    ecken = np.zeros((nc + 1, nr + 1, 2))
    for i in range(nc + 1):
        for j in range(nr + 1):
            jit = rng.uniform(-cfg.jitter_mm, cfg.jitter_mm, 2)
            ecken[i, j] = (x0 + i * cw + jit[0], y0 + j * cw + jit[1])
    # segments per cell border, shared by both sides of a sewn border
    vseg = np.zeros((nc + 1, nr), dtype=int)
    hseg = np.zeros((nc, nr + 1), dtype=int)
    for i in range(nc + 1):
        for j in range(nr):
            vseg[i, j] = max(2, int(round(np.linalg.norm(ecken[i, j + 1] - ecken[i, j]) / cfg.kante_mm)))
    for i in range(nc):
        for j in range(nr + 1):
            hseg[i, j] = max(2, int(round(np.linalg.norm(ecken[i + 1, j] - ecken[i, j]) / cfg.kante_mm)))
    teile, seiten_je, min_w = [], [], []
    for j in range(nr):
        for i in range(nc):
            c = np.array([ecken[i, j], ecken[i + 1, j], ecken[i + 1, j + 1], ecken[i, j + 1]])
            cen = c.mean(axis=0)
            c = cen + (c - cen) * (1.0 - cfg.luecke_mm / (0.5 * cw))
            n_s = [hseg[i, j], vseg[i + 1, j], hseg[i, j + 1], vseg[i, j]]
            pts, tri, seiten = _panel_mesh(c, cfg.kante_mm, rng, cfg.sagitta_mm, n_s)
            theta = rng.uniform(0.0, math.pi)
            teile.append(Teil(f"Teil {j}/{i}", pts, tri, fadenlauf=(math.cos(theta), math.sin(theta)),
                              material=mat, reduktion=0.0))
            seiten_je.append(seiten)
            min_w.append(min_winkel_grad(pts, tri))
    naehte = []
    for j in range(nr):
        for i in range(nc - 1):
            a, b = j * nc + i, j * nc + i + 1
            # edge 1 of a runs bottom->top, edge 3 of b runs top->bottom
            sa = seiten_je[a][1]
            sb = list(reversed(seiten_je[b][3]))
            naehte.append([Teilnaht(a, b, list(zip(sa, sb)))])
    z = cfg.hoehe_m
    lage = lambda _ti, p: np.column_stack([p[:, 0] * 1e-3, p[:, 1] * 1e-3, np.full(p.shape[0], z)])
    koerper = []
    if mit_koerper:
        koerper = [sdf if sdf is not None else koerper_sdf()]
    param = Simparameter(bild_dt=cfg.bild_dt, schwerkraft=(0.0, 0.0, -9.81), daempfung=0.0)
    s = bauen(teile, naehte, lage, koerper, param)
    mw = np.concatenate(min_w)
    bericht = {
        "punkte": s.n, "dreiecke": int(s.dreiecke.shape[0]), "teile": len(teile), "naehte": len(naehte),
        "naht_paare": int(s.naht_punkte.shape[0]), "biegungen": int(s.biegung_punkte.shape[0]),
        "min_winkel_grad": float(mw.min()), "anteil_unter_25_grad": float(np.mean(mw < 25.0)),
    }
    return s, bericht


@dataclass
class StapelKonfig:
    """Batched variants: independent garments of uneven sizes on one body."""
    anzahl: int = 64
    v_min: int = 2000
    v_max: int = 60000
    seed: int = 20261204
    kante_mm: float = 10.0


# points per mm^2 of the drape's panels at 10 mm edges (40,758 points on 28
# cells of 352 mm)
_DRAPE_DICHTE = 40758.0 / (28 * 352.0 * 352.0)


def stapel_variante(cfg: StapelKonfig, k: int) -> Tuple[DrapeKonfig, dict]:
    """Instance k of the batch: the drape's 7 x 4 panel grid scaled to a point
    count log-uniform in [v_min, v_max], latex or woven at random, stiffnesses
    (stretch, shear, bending) scaled by one factor log-uniform over one decade,
    friction uniform in [0.1, 0.5]. Deterministic in (seed, k)."""
    rng = np.random.default_rng([cfg.seed, k])
    v_ziel = math.exp(rng.uniform(math.log(cfg.v_min), math.log(cfg.v_max)))
    zelle = math.sqrt(v_ziel / (28.0 * _DRAPE_DICHTE)) * cfg.kante_mm / 10.0
    faktor = 10.0 ** rng.uniform(-0.5, 0.5)
    mu = rng.uniform(0.1, 0.5)
    if rng.random() < 0.5:
        mat = latex(1.0 * faktor, 0.5e-3, 950.0, 1e-6 * faktor, reibung=mu)
        art = "latex"
    else:
        mat = gewebe(5.0 * faktor, 3.0 * faktor, 0.5 * faktor, 0.5e-3, 0.20, 1.2e-6 * faktor, 0.8e-6 * faktor,
                     0.9e-6 * faktor, mu)
        art = "gewebe"
    d = DrapeKonfig(seed=cfg.seed + 7919 * (k + 1), zelle_mm=zelle, kante_mm=cfg.kante_mm,
                    jitter_mm=30.0 * zelle / 352.0, versatz_mm=(0.0, -0.5 * zelle), material=mat)
    return d, {"v_ziel": v_ziel, "material": art, "faktor": faktor, "reibung": mu, "zelle_mm": zelle}


def _stapel_eins(args):
    cfg, k = args
    d, info = stapel_variante(cfg, k)
    s, bericht = drape(d, mit_koerper=False)
    bericht.update(info)
    return s, bericht


def stapel(cfg: Optional[StapelKonfig] = None, sdf: Optional[GitterSdf] = None, prozesse: int = 0) -> List[Tuple[Szene, dict]]:
    """The batch's scenes, each on the same body ``sdf`` (built once here if
    not given). prozesse > 1 builds the meshes in that many worker processes."""
    cfg = cfg or StapelKonfig()
    sdf = sdf if sdf is not None else koerper_sdf()
    auftraege = [(cfg, k) for k in range(cfg.anzahl)]
    if prozesse > 1:
        from concurrent.futures import ProcessPoolExecutor
        with ProcessPoolExecutor(max_workers=prozesse) as ex:
            out = list(ex.map(_stapel_eins, auftraege))
    else:
        out = [_stapel_eins(a) for a in auftraege]
    for s, _ in out:
        s.koerper = [sdf]
    return out
