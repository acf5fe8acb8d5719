"""The scene: points, constraints and bodies as flat arrays (szene.rs:197-269),
and the builder that turns meshed pattern pieces into one (szene.rs:320-715).

The builder is host-side set-up work that runs once per scene. It restates the
reference's construction so that the constraint records, stiffnesses and the
colour order (which fixes the summation order of every per-point sum) are the
ones the reference would build from the same mesh:

* mass lumping ``A * skala**2 * rho_A / 3`` per triangle corner (szene.rs:386-390);
* one distance constraint per mesh edge with ``k = E * sum(A) / L**2`` and
  rest length ``L * skala`` (szene.rs:425-432);
* one bending constraint per interior edge with ``k = D(theta) * L**2 / (A1+A2)``,
  ``D`` from the directional rule of biegung.rs (szene.rs:433-461);
* bending across sewn boundary edges (szene.rs:494-529);
* one seam spring per seam point pair, rest 0 (szene.rs:532-544);
* one membrane per triangle from the reduced rest edge lengths (szene.rs:602-647),
  whose three edges are then flagged ``membran_kante`` (k = 0 in the solver);
* greedy colouring and a stable sort by colour of every group (szene.rs:683-686,
  faerbung.rs:11-57).

Unlike the reference, a piece may give its rest geometry per triangle corner
(``dreieck_uv_mm``) instead of per point. That is what a closed ring needs: the
periodic hanging tube has no planar chart in which every triangle is flat at
once, but each triangle has one.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Callable, List, Optional, Sequence, Tuple

import numpy as np

from .material import Material

# ---------------------------------------------------------------------------
# data model
# ---------------------------------------------------------------------------


@dataclass
class Simparameter:
    """szene.rs:95-126 (without the seam ease switch)."""
    bild_dt: float = 1.0 / 60.0
    schwerkraft: Tuple[float, float, float] = (0.0, 0.0, -9.81)
    daempfung: float = 0.0
    naht_nachgiebigkeit: float = 0.0
    # substeps per frame of the XPBD path (one constraint round each)
    unterschritte: int = 20


# szene.rs:83: an XPBD edge is "soft" when alpha / h^2 >= 9 (w_i + w_j)
WEICH_FAKTOR = 9.0


@dataclass
class Zylinder:
    """Infinite cylinder body (koerper.rs:19)."""
    punkt: Tuple[float, float, float]
    achse: Tuple[float, float, float]
    radius: float
    geschwindigkeit: Tuple[float, float, float] = (0.0, 0.0, 0.0)

    def verschoben(self, t: float) -> "Zylinder":
        """koerper.rs:122-133."""
        p = tuple(self.punkt[k] + self.geschwindigkeit[k] * t for k in range(3))
        return Zylinder(p, self.achse, self.radius, self.geschwindigkeit)


@dataclass
class GitterSdf:
    """A body given as a signed distance field sampled on a regular grid
    (float32, x fastest, shape (nz, ny, nx)), rigid, moved by ``versatz``.

    It stands where the reference has a triangle-mesh body (koerper.rs:25) and
    answers the same contact query (koerper.rs:69-91): a point closer than
    ``abstand`` to the surface, or inside, is in contact with depth
    ``abstand - d`` and the normalised gradient as normal. Trilinear sampling.
    """
    werte: np.ndarray
    origin: Tuple[float, float, float]
    spacing: float
    versatz: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    geschwindigkeit: Tuple[float, float, float] = (0.0, 0.0, 0.0)

    @property
    def dims(self) -> Tuple[int, int, int]:
        nz, ny, nx = self.werte.shape
        return (nx, ny, nz)

    def verschoben(self, t: float) -> "GitterSdf":
        v = tuple(self.versatz[k] + self.geschwindigkeit[k] * t for k in range(3))
        return GitterSdf(self.werte, self.origin, self.spacing, v, self.geschwindigkeit)


@dataclass
class Netz:
    """A closed, consistently oriented triangle mesh as body (koerper.rs:25,
    netzkoerper.rs:145-177), rigid, moved by ``versatz``.

    ``punkte`` (n, 3) in m, ``dreiecke`` (m, 3) with outward normals;
    ``suchtiefe`` is how deep inside a cloth point is still found
    (netzkoerper.rs:44-46, default 10 mm). ``punktgeschwindigkeit`` (n, 3) is
    the velocity field of a deforming body (netzkoerper.rs:454-489), or None
    for one velocity of the whole body.
    """
    punkte: np.ndarray
    dreiecke: np.ndarray
    suchtiefe: float = 0.01
    versatz: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    geschwindigkeit: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    punktgeschwindigkeit: Optional[np.ndarray] = None

    def verschoben(self, t: float) -> "Netz":
        """koerper.rs:122-133."""
        v = tuple(self.versatz[k] + self.geschwindigkeit[k] * t for k in range(3))
        return Netz(self.punkte, self.dreiecke, self.suchtiefe, v, self.geschwindigkeit, self.punktgeschwindigkeit)


@dataclass
class Szene:
    """szene.rs:197-269 as numpy arrays. Constraint groups are colour-ordered."""
    x: np.ndarray                      # (n, 3) f64
    v: np.ndarray                      # (n, 3) f64
    masse: np.ndarray                  # (n,)
    w: np.ndarray                      # (n,)
    kollisionsabstand: np.ndarray      # (n,)
    reibung: np.ndarray                # (n,)
    abstand_punkte: np.ndarray         # (E, 2) u32
    abstand_ruhe: np.ndarray
    abstand_nachgiebigkeit: np.ndarray
    membran_kante: np.ndarray          # (E,) bool
    naht_punkte: np.ndarray            # (S, 2) u32
    naht_ruhe: np.ndarray
    biegung_punkte: np.ndarray         # (B, 4) u32
    biegung_ruhe: np.ndarray
    biegung_nachgiebigkeit: np.ndarray
    membran_punkte: np.ndarray         # (T, 3) u32
    membran_daten: np.ndarray          # (T, 12) f64
    dreiecke: np.ndarray               # (T_all, 3) u32
    teil_von_punkt: np.ndarray         # (n,) u32
    koerper: list = field(default_factory=list)
    param: Simparameter = field(default_factory=Simparameter)
    zeit: float = 0.0
    membran_entartet: int = 0
    naht_biegung_ohne: int = 0
    # self-collision inputs (selbstkollision.rs:660-680): stretch stiffness
    # (N/m) of each point's and each triangle's piece, rest area per triangle
    selbst_e_punkt: Optional[np.ndarray] = None
    selbst_e_dreieck: Optional[np.ndarray] = None
    selbst_a_dreieck: Optional[np.ndarray] = None

    @property
    def n(self) -> int:
        return int(self.x.shape[0])

    def gesamtmasse(self) -> float:
        return float(np.sum(self.masse))

    def unterschritte_weich(self, faktor: float = WEICH_FAKTOR) -> int:
        """szene.rs:794-807: XPBD substeps per frame so that every distance
        constraint lies in XPBD's soft range, alpha / h^2 >= faktor (w_i + w_j)."""
        h_max = math.inf
        for c in range(self.abstand_punkte.shape[0]):
            i, j = int(self.abstand_punkte[c, 0]), int(self.abstand_punkte[c, 1])
            wsum = self.w[i] + self.w[j]
            if wsum > 0.0:
                h_max = min(h_max, math.sqrt(self.abstand_nachgiebigkeit[c] / (faktor * wsum)))
        if not math.isfinite(h_max):
            return self.param.unterschritte
        return int(max(math.ceil(self.param.bild_dt / h_max), 1.0))

    def kopie(self) -> "Szene":
        import copy
        s = copy.copy(self)
        for name in ("x", "v", "w", "naht_ruhe", "masse"):
            setattr(s, name, getattr(self, name).copy())
        s.koerper = list(self.koerper)
        return s

    def naht_offen(self) -> None:
        """szene.rs:778-782."""
        self.naht_ruhe[:] = np.inf

    def naht_kuerzen(self, schritt: float) -> None:
        """szene.rs:768-774."""
        for c in range(self.naht_punkte.shape[0]):
            i, j = self.naht_punkte[c]
            d = self.x[i] - self.x[j]
            l = math.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])
            self.naht_ruhe[c] = max(min(self.naht_ruhe[c], l) - schritt, 0.0)


# ---------------------------------------------------------------------------
# restated helpers
# ---------------------------------------------------------------------------


def _l2(a, b) -> float:
    """szene.rs:272-274."""
    dx, dy = a[0] - b[0], a[1] - b[1]
    return math.sqrt(dx * dx + dy * dy)


def _flaeche(a, b, c) -> float:
    """szene.rs:277-279."""
    return 0.5 * ((b[0] - a[0]) * (c[1] - a[1]) - (c[0] - a[0]) * (b[1] - a[1]))


def _norm2(x: float, y: float) -> float:
    return math.sqrt(x * x + y * y + 0.0 * 0.0)


def membran_bauen(ruhe, kanten, dir2, material) -> Optional[np.ndarray]:
    """kontinuum.rs:193-286: rest triangle from the three (reduced) edge lengths,
    warp/weft basis turned into the frame of edge 0->1. Returns the 12-value
    element record {A0, ruhe_inv[4], kette[2], schuss[2], E_k, E_s, G}."""
    l01, l12, l20 = kanten
    if not (l01 > 0.0 and l12 > 0.0 and l20 > 0.0):
        return None
    x = (l01 * l01 + l20 * l20 - l12 * l12) / (2.0 * l01)
    y2 = l20 * l20 - x * x
    if not (y2 > 0.0):
        return None
    y = math.sqrt(y2)
    det = l01 * y
    if not (det > 0.0):
        return None
    inv = (1.0 / l01, -x / det, 0.0, 1.0 / y)
    fl = 0.5 * det
    d0, d1 = ruhe[1][0] - ruhe[0][0], ruhe[1][1] - ruhe[0][1]
    ld = _norm2(d0, d1)
    if not (ld > 0.0):
        return None
    c, s = d0 / ld, d1 / ld

    def nach_ruhe(v0, v1):
        l = _norm2(v0, v1)
        if l > 0.0:
            a, b = v0 / l, v1 / l
            return (c * a + s * b, -s * a + c * b)
        return None

    kette = nach_ruhe(dir2[0], dir2[1])
    if kette is None:
        return None
    q = nach_ruhe(-dir2[1], dir2[0])
    if q is None:
        return None
    along = q[0] * kette[0] + q[1] * kette[1]
    w0, w1 = q[0] - along * kette[0], q[1] - along * kette[1]
    lw = _norm2(w0, w1)
    if not (lw > 1e-12):
        return None
    schuss = (w0 / lw, w1 / lw)
    return np.array([fl, *inv, *kette, *schuss, material[0], material[1], material[2]], dtype=np.float64)


def cos2_sin2(e, warp):
    """biegung.rs:156-173."""
    le = math.sqrt(e[0] * e[0] + e[1] * e[1])
    lw = math.sqrt(warp[0] * warp[0] + warp[1] * warp[1])
    if not (le > 1e-12) or not (lw > 1e-12):
        return None
    ex, ey = e[0] / le, e[1] / le
    wx, wy = warp[0] / lw, warp[1] / lw
    c_edge = ex * wx + ey * wy
    s_edge = ex * wy - ey * wx
    c, s = -s_edge, c_edge
    return (c * c, s * s)


def steifigkeit_aus(c2, s2, weft, warp, bias) -> float:
    """biegung.rs:111-142 (with the positivity floor)."""
    if weft == warp and warp == bias:
        return weft
    d = warp * (c2 * c2) + weft * (s2 * s2) + (4.0 * bias - warp - weft) * (c2 * s2)
    if d > 0.0:
        return d
    return 1.0e-3 * abs(min(min(weft, warp), bias))


def faerbe(punkte: np.ndarray, npunkte: int):
    """faerbung.rs:11-43: greedy colouring in input order."""
    maske = [0] * npunkte
    mehr: dict = {}
    farbe = np.zeros(punkte.shape[0], dtype=np.uint32)
    zahl = 0
    voll = (1 << 64) - 1
    rows = punkte.tolist()
    for c, p in enumerate(rows):
        belegt = 0
        for i in p:
            belegt |= maske[i]
        if belegt != voll:
            frei = (~belegt) & voll
            f = (frei & -frei).bit_length() - 1
        else:
            f = 64
            while any(f in mehr.get(i, ()) for i in p):
                f += 1
        for i in p:
            if f < 64:
                maske[i] |= 1 << f
            else:
                mehr.setdefault(i, []).append(f)
        farbe[c] = f
        zahl = max(zahl, f + 1)
    return farbe, zahl


def ordnung(farbe: np.ndarray) -> np.ndarray:
    """faerbung.rs:46-57: stable order by colour."""
    return np.argsort(farbe, kind="stable")


# ---------------------------------------------------------------------------
# builder input
# ---------------------------------------------------------------------------


@dataclass
class Teil:
    """One meshed pattern piece: points in pattern coordinates (mm), triangles,
    optional per-corner rest charts (mm), grainline, material and reduction."""
    name: str
    punkte_mm: np.ndarray             # (n, 2)
    dreiecke: np.ndarray              # (t, 3) local indices
    fadenlauf: Tuple[float, float] = (1.0, 0.0)
    material: Optional[Material] = None
    reduktion: float = 0.0
    dreieck_uv_mm: Optional[np.ndarray] = None   # (t, 3, 2)


@dataclass
class Teilnaht:
    """A sewn run between two pieces: point pairs (local indices) in order
    along the seam, as klon3d-mesh emits them (crates/klon3d-mesh/src/lib.rs:351-381)."""
    teil_a: int
    teil_b: int
    paare: List[Tuple[int, int]]


def bauen(teile: Sequence[Teil], naehte: Sequence[Sequence[Teilnaht]],
          lage: Callable[[int, np.ndarray], np.ndarray], koerper: list,
          param: Simparameter, extra_punkte: Optional[dict] = None) -> Szene:
    """Szene::bauen without allowance band and ease (szene.rs:320-715).

    ``lage(teil, punkte_mm)`` maps a piece's points (n, 2) to 3D (n, 3).
    ``extra_punkte`` optionally appends held points (mass 0) and seam springs
    to them: {"x": (m, 3), "naht": [(cloth point, extra index), ...]}.
    """
    xs, masse_l, koll, reib, teilnr = [], [], [], [], []
    a_p, a_r, a_a = [], [], []
    b_p, b_r, b_a = [], [], []
    versatz = []
    randkanten = []           # per piece: (i, j) -> (o, area)
    punkte_m_l = []
    biegesteif = []
    dreiecke_all = []
    e_punkt_l, e_dreieck_l, a_dreieck_l = [], [], []
    membran_records = []      # (piece, local tri, global tri, uv)
    kanten_index_je_teil = []
    off = 0
    for ti, t in enumerate(teile):
        mat = t.material
        skala = 1.0 - t.reduktion
        p_mm = np.asarray(t.punkte_mm, dtype=np.float64)
        p_m = p_mm * 1e-3
        n_t = p_m.shape[0]
        versatz.append(off)
        xs.append(np.asarray(lage(ti, p_mm), dtype=np.float64))
        masse_t = [0.0] * n_t
        koll.append(np.full(n_t, mat.dicke_m / 2.0))
        reib.append(np.full(n_t, mat.reibung))
        teilnr.append(np.full(n_t, ti, dtype=np.uint32))
        tris = np.asarray(t.dreiecke, dtype=np.int64)
        if t.dreieck_uv_mm is not None:
            uv = np.asarray(t.dreieck_uv_mm, dtype=np.float64) * 1e-3
        else:
            uv = p_m[tris]
        # edge -> [asum, [(opposite, tri)], uv of (i, j) from the first triangle]
        kanten = {}
        fg = mat.flaechengewicht_kg_m2
        for k_t in range(tris.shape[0]):
            d = tris[k_t]
            q = uv[k_t]
            a = abs(_flaeche(q[0], q[1], q[2]))
            for c in range(3):
                masse_t[d[c]] += a * skala * skala * fg / 3.0
            for k in range(3):
                i, j, o = int(d[k]), int(d[(k + 1) % 3]), int(d[(k + 2) % 3])
                key = (min(i, j), max(i, j))
                e = kanten.get(key)
                if e is None:
                    e = [0.0, [], {}]
                    kanten[key] = e
                e[0] += a
                e[1].append((o, k_t))
                # chart coordinates of the edge's end points in this triangle
                e[2].setdefault(k_t, {i: q[k], j: q[(k + 1) % 3], o: q[(k + 2) % 3]})
            dreiecke_all.append([int(d[0]) + off, int(d[1]) + off, int(d[2]) + off])
            e_dreieck_l.append(mat.dehnung_n_m)
            a_dreieck_l.append(a * skala * skala)
            membran_records.append((ti, k_t, [int(d[0]), int(d[1]), int(d[2])], q))
        masse_l.append(np.array(masse_t))
        e_punkt_l.append(np.full(n_t, mat.dehnung_n_m))
        rand = {}
        kanten_e = {}
        dir2 = t.fadenlauf
        for key in sorted(kanten.keys()):
            asum, gegen, charts = kanten[key]
            i, j = key
            first_tri = gegen[0][1]
            ci = charts[first_tri]
            pi, pj = ci[i], ci[j]
            l = _l2(pi, pj)
            if l <= 0.0:
                continue
            kk = mat.dehnung_n_m * asum / (l * l)
            a_p.append((i + off, j + off))
            a_r.append(l * skala)
            a_a.append(1.0 / kk)
            kanten_e[key] = len(a_r) - 1
            if len(gegen) == 1:
                rand[key] = (gegen[0][0], asum)
            if len(gegen) == 2:
                (o1, t1), (o2, t2) = gegen
                c1, c2 = charts[t1], charts[t2]
                if _flaeche(c1[i], c1[j], c1[o1]) > 0.0:
                    (o_a, ca), (o_b, cb) = (o1, c1), (o2, c2)
                else:
                    (o_a, ca), (o_b, cb) = (o2, c2), (o1, c1)
                a12 = abs(_flaeche(ca[i], ca[j], ca[o_a])) + abs(_flaeche(cb[j], cb[i], cb[o_b]))
                cs = cos2_sin2((pj[0] - pi[0], pj[1] - pi[1]), dir2)
                if cs is None:
                    dd = mat.biegung_nm
                else:
                    dd = steifigkeit_aus(cs[0], cs[1], mat.biegung_nm, mat.biegung_kette_nm, mat.biegung_schraege_nm)
                kb = dd * l * l / a12
                b_p.append((i + off, j + off, o_a + off, o_b + off))
                b_r.append(0.0)
                b_a.append(1.0 / kb)
        randkanten.append(rand)
        punkte_m_l.append(p_m)
        biegesteif.append(mat.biegung_nm)
        kanten_index_je_teil.append(kanten_e)
        off += n_t

    # bending across the seams (szene.rs:494-529)
    naht_biegung_ohne = 0
    naht_kanten = set()
    for naht in naehte:
        for tn in naht:
            ta, tb = tn.teil_a, tn.teil_b
            for w0, w1 in zip(tn.paare[:-1], tn.paare[1:]):
                a0, a1, b0, b1 = w0[0], w1[0], w0[1], w1[1]
                ra = randkanten[ta].get((min(a0, a1), max(a0, a1)))
                rb = randkanten[tb].get((min(b0, b1), max(b0, b1)))
                if ra is None or rb is None:
                    continue
                oa, fa = ra
                ob, fb = rb
                key = (versatz[ta] + min(a0, a1), versatz[ta] + max(a0, a1))
                if key in naht_kanten:
                    continue
                naht_kanten.add(key)
                pa, pb = punkte_m_l[ta], punkte_m_l[tb]
                if _flaeche(pa[a0], pa[a1], pa[oa]) > 0.0:
                    (x0a, x1a), (x0b, x1b) = (a0, a1), (b0, b1)
                else:
                    (x0a, x1a), (x0b, x1b) = (a1, a0), (b1, b0)
                if _flaeche(pb[x1b], pb[x0b], pb[ob]) <= 0.0:
                    naht_biegung_ohne += 1
                    continue
                l = _l2(pa[a0], pa[a1])
                kb = 0.5 * (biegesteif[ta] + biegesteif[tb]) * l * l / (fa + fb)
                b_p.append((x0a + versatz[ta], x1a + versatz[ta], oa + versatz[ta], ob + versatz[tb]))
                b_r.append(0.0)
                b_a.append(1.0 / kb)

    # seams: point pairs, duplicates and self pairs dropped (szene.rs:532-544)
    paare = set()
    for naht in naehte:
        for tn in naht:
            for (a, b) in tn.paare:
                ga, gb = versatz[tn.teil_a] + a, versatz[tn.teil_b] + b
                if ga != gb:
                    paare.add((min(ga, gb), max(ga, gb)))
    n_p = sorted(paare)

    x = np.concatenate(xs, axis=0)
    masse = np.concatenate(masse_l)
    kollisionsabstand = np.concatenate(koll)
    reibung = np.concatenate(reib)
    teil_von_punkt = np.concatenate(teilnr)
    e_punkt = np.concatenate(e_punkt_l)

    # held extra points joined by seam springs (used by the hanging tube's ring)
    if extra_punkte is not None:
        ex = np.asarray(extra_punkte["x"], dtype=np.float64)
        base = x.shape[0]
        x = np.concatenate([x, ex], axis=0)
        masse = np.concatenate([masse, np.zeros(ex.shape[0])])
        kollisionsabstand = np.concatenate([kollisionsabstand, np.zeros(ex.shape[0])])
        reibung = np.concatenate([reibung, np.zeros(ex.shape[0])])
        teil_von_punkt = np.concatenate([teil_von_punkt, np.full(ex.shape[0], len(teile), dtype=np.uint32)])
        e_punkt = np.concatenate([e_punkt, np.zeros(ex.shape[0])])
        for (c, e) in extra_punkte["naht"]:
            n_p.append((min(c, base + e), max(c, base + e)))

    # membranes from the reduced rest lengths (szene.rs:602-647)
    membran_kante = np.zeros(len(a_r), dtype=bool)
    mem_p, mem_d = [], []
    membran_entartet = 0
    for (ti, k_t, d, q) in membran_records:
        t = teile[ti]
        idx = kanten_index_je_teil[ti]
        ks = []
        ok = True
        for (i, j) in ((d[0], d[1]), (d[1], d[2]), (d[2], d[0])):
            e = idx.get((min(i, j), max(i, j)))
            if e is None:
                ok = False
                break
            ks.append(a_r[e])
        el = None
        if ok:
            mat = t.material
            el = membran_bauen(q, ks, t.fadenlauf, (mat.kette_n_m, mat.dehnung_n_m, mat.schub_n_m))
        if el is None:
            membran_entartet += 1
            continue
        for (i, j) in ((d[0], d[1]), (d[1], d[2]), (d[2], d[0])):
            e = idx.get((min(i, j), max(i, j)))
            if e is not None:
                membran_kante[e] = True
        o = versatz[ti]
        mem_p.append((d[0] + o, d[1] + o, d[2] + o))
        mem_d.append(el)

    npkt = x.shape[0]
    w = np.where(masse > 0.0, 1.0 / np.where(masse > 0.0, masse, 1.0), 0.0)

    a_p = np.asarray(a_p, dtype=np.uint32).reshape(-1, 2)
    fa, _ = faerbe(a_p, npkt)
    oa = ordnung(fa)
    b_p = np.asarray(b_p, dtype=np.uint32).reshape(-1, 4)
    fb, _ = faerbe(b_p, npkt)
    ob = ordnung(fb)
    n_pa = np.asarray(n_p, dtype=np.uint32).reshape(-1, 2)
    fn, _ = faerbe(n_pa, npkt)
    on = ordnung(fn)

    return Szene(
        x=x, v=np.zeros_like(x), masse=masse, w=w,
        kollisionsabstand=kollisionsabstand, reibung=reibung,
        abstand_punkte=np.ascontiguousarray(a_p[oa]),
        abstand_ruhe=np.asarray(a_r, dtype=np.float64)[oa],
        abstand_nachgiebigkeit=np.asarray(a_a, dtype=np.float64)[oa],
        membran_kante=membran_kante[oa],
        naht_punkte=np.ascontiguousarray(n_pa[on]),
        naht_ruhe=np.zeros(n_pa.shape[0]),
        biegung_punkte=np.ascontiguousarray(b_p[ob]),
        biegung_ruhe=np.asarray(b_r, dtype=np.float64)[ob],
        biegung_nachgiebigkeit=np.asarray(b_a, dtype=np.float64)[ob],
        membran_punkte=np.asarray(mem_p, dtype=np.uint32).reshape(-1, 3),
        membran_daten=np.asarray(mem_d, dtype=np.float64).reshape(-1, 12),
        dreiecke=np.asarray(dreiecke_all, dtype=np.uint32).reshape(-1, 3),
        teil_von_punkt=teil_von_punkt,
        koerper=list(koerper), param=param,
        membran_entartet=membran_entartet, naht_biegung_ohne=naht_biegung_ohne,
        selbst_e_punkt=e_punkt, selbst_e_dreieck=np.asarray(e_dreieck_l, dtype=np.float64),
        selbst_a_dreieck=np.asarray(a_dreieck_l, dtype=np.float64),
    )
