"""Restatement of klon3d-mesh's seam pairing by length fraction, in plain
Python FP64 (the checker of gc_naht_planen).

TEST INFRASTRUCTURE ONLY: only tests/ load it; the product never does.

Follows crates/klon3d-mesh/src/rand.rs (Kantenzug 56-130, unterteile
329-351, abbilden 353-359, orte 373-384, stellenname 386-398, naht_planen
400-501, Randplan::neu 503-533, schluessel 536-545, randindex 547-583) and
the pair emission of crates/klon3d-mesh/src/lib.rs:351-381. Distances use the
C library's hypot (numpy.hypot), as Rust's f64::hypot does.
"""
from __future__ import annotations

import struct
from bisect import bisect_right

import numpy as np

ENDE, GEMEINSAM, NUR_A, NUR_B = 0, 1, 2, 3


class PlanFehler(ValueError):
    """The pattern cannot be planned (the reference's Err)."""


def abstand(a, b) -> float:
    return float(np.hypot(a[0] - b[0], a[1] - b[1]))


def punkt_strecke(p, a, b) -> float:
    """rand.rs:46-54."""
    dx, dy = b[0] - a[0], b[1] - a[1]
    l2 = dx * dx + dy * dy
    if l2 <= 0.0:
        return abstand(p, a)
    t = ((p[0] - a[0]) * dx + (p[1] - a[1]) * dy) / l2
    t = min(max(t, 0.0), 1.0)
    return abstand(p, (a[0] + t * dx, a[1] + t * dy))


class Kantenzug:
    """rand.rs:56-130: polyline with arc length; points within 1e-9 mm of the
    previous one are dropped."""

    def __init__(self, roh):
        self.punkte, self.kum = [], []
        for p in roh:
            p = (float(p[0]), float(p[1]))
            if not self.punkte:
                self.kum.append(0.0)
            else:
                d = abstand(self.punkte[-1], p)
                if d <= 1e-9:
                    continue
                self.kum.append(self.kum[-1] + d)
            self.punkte.append(p)
        self.laenge = self.kum[-1] if self.kum else 0.0

    def ort(self, s: float):
        if s <= 0.0:
            return self.punkte[0]
        if s >= self.laenge:
            return self.punkte[-1]
        i = bisect_right(self.kum, s) - 1
        a, b = self.punkte[i], self.punkte[i + 1]
        t = (s - self.kum[i]) / (self.kum[i + 1] - self.kum[i])
        return (a[0] + t * (b[0] - a[0]), a[1] + t * (b[1] - a[1]))

    def sehne_ecke(self, s0: float, s1: float):
        s0, s1 = min(s0, s1), max(s0, s1)
        a, b = self.ort(s0), self.ort(s1)
        hoehe, ecke = 0.0, None
        for i in range(bisect_right(self.kum, s0), len(self.kum)):
            if self.kum[i] >= s1:
                break
            h = punkt_strecke(self.punkte[i], a, b)
            if ecke is None or h > hoehe:
                hoehe, ecke = h, self.kum[i]
        return hoehe, (s1 - s0) - abstand(a, b), ecke

    def sehne(self, s0: float, s1: float):
        h, v, _ = self.sehne_ecke(s0, s1)
        return h, v


class Glied:
    """rand.rs:194-218."""

    def __init__(self, teil, kante, s0, s1):
        self.teil, self.kante, self.s0, self.s1 = teil, kante, float(s0), float(s1)

    def s(self, f: float) -> float:
        if f <= 0.0:
            return self.s0
        if f >= 1.0:
            return self.s1
        return self.s0 + f * (self.s1 - self.s0)

    def laenge(self) -> float:
        return abs(self.s1 - self.s0)


class Gruppe:
    def __init__(self, glieder):
        self.glieder = glieder
        self.f = []


def unterteile(g: Gruppe, teile, p) -> None:
    """rand.rs:329-351."""
    lmax = 0.0
    for gl in g.glieder:
        lmax = max(lmax, gl.laenge())
    m = max(int(np.ceil(lmax / p["kante_mm"])), 1)
    g.f = [j / m for j in range(m + 1)]
    j = 0
    while j + 1 < len(g.f):
        f0, f1 = g.f[j], g.f[j + 1]
        grob, ecke = False, None
        for gl in g.glieder:
            kz = teile[gl.teil][gl.kante]
            s0, s1 = gl.s(f0), gl.s(f1)
            hoehe, verlust, e = kz.sehne_ecke(s0, s1)
            if hoehe > p["sehne_mm"] or verlust > p["laenge_mm"] * abs(s1 - s0) / kz.laenge:
                grob = True
                if ecke is None and e is not None:
                    ecke = (e - gl.s0) / (gl.s1 - gl.s0)
        teilung = None
        if grob:
            if ecke is not None and (ecke - f0) * lmax >= p["min_rand_mm"] and (f1 - ecke) * lmax >= p["min_rand_mm"]:
                teilung = ecke
            elif (f1 - f0) * lmax >= 2.0 * p["min_rand_mm"]:
                teilung = 0.5 * (f0 + f1)
        if teilung is None:
            j += 1
        else:
            g.f.insert(j + 1, teilung)


def abbilden(anker, x: float, rueck: bool) -> float:
    """rand.rs:353-359."""
    paar = (lambda a: (a[1], a[0])) if rueck else (lambda a: (a[0], a[1]))
    k = 0
    for i in range(len(anker) - 1, -1, -1):
        if paar(anker[i])[0] <= x:
            k = i
            break
    k = min(k, len(anker) - 2)
    (x0, y0), (x1, y1) = paar(anker[k]), paar(anker[k + 1])
    return y0 + (x - x0) / (x1 - x0) * (y1 - y0)


def _orte(s, kum, l, rho, grenze):
    """rand.rs:373-384: (piece before, s), (piece after, s)."""
    if grenze is not None:
        i = grenze
        vor = (i - 1, s[i - 1]["ende"]) if i > 0 else None
        nach = (i, s[i]["anfang"]) if i < len(s) else None
        return vor, nach
    sigma = rho * l
    k = bisect_right(kum, sigma)
    k = min(max(k - 1, 0), len(s) - 1)
    d = min(max(sigma - kum[k], 0.0), s[k]["laenge"])
    pos = s[k]["laenge"] - d if s[k]["rueck"] else d
    return (k, pos), (k, pos)


def _knoten(s, teile, i):
    """rand.rs:386-398: the piece and edge whose start node the boundary i is
    (the start of run piece i, or the end of the last)."""
    if i == len(s):
        x = s[i - 1]
        k = x["kante"] if x["rueck"] else x["kante"] + 1
    else:
        x = s[i]
        k = x["kante"] + 1 if x["rueck"] else x["kante"]
    return (x["teil"], k % len(teile[x["teil"]]))


def naht_planen(a, b, teile, p, gruppen):
    """rand.rs:400-501."""
    def kum(s):
        v = [0.0]
        for x in s:
            v.append(v[-1] + x["laenge"])
        return v
    ka, kb = kum(a), kum(b)
    la, lb = ka[-1], kb[-1]
    lm = 0.5 * (la + lb)
    ra = [x / la for x in ka[1:len(a)]]
    rb = [x / lb for x in kb[1:len(b)]]

    def naechst(x, v):
        best, bd = None, 0.0
        for i, y in enumerate(v):
            d = abs(y - x)
            if best is None or d < bd:
                best, bd = i, d
        return best
    fang = max(p["fang_mm"], 1e-3)
    anker = [(0.0, 0.0)]
    gemeinsam = []
    for i, x in enumerate(ra):
        j = naechst(x, rb)
        if j is None:
            continue
        letzt = anker[-1]
        if abs(rb[j] - x) * lm <= fang and naechst(rb[j], ra) == i and x > letzt[0] and rb[j] > letzt[1]:
            anker.append((x, rb[j]))
            gemeinsam.append((i, j))
    anker.append((1.0, 1.0))
    st = [dict(ra=0.0, rb=0.0, ga=0, gb=0, art=ENDE)]
    ia = {g[0] for g in gemeinsam}
    jb = {g[1] for g in gemeinsam}
    for i, j in gemeinsam:
        st.append(dict(ra=ra[i], rb=rb[j], ga=i + 1, gb=j + 1, art=GEMEINSAM))
    for i, x in enumerate(ra):
        if i not in ia:
            st.append(dict(ra=x, rb=abbilden(anker, x, False), ga=i + 1, gb=None, art=NUR_A))
    for j, y in enumerate(rb):
        if j not in jb:
            st.append(dict(ra=abbilden(anker, y, True), rb=y, ga=None, gb=j + 1, art=NUR_B))
    st.append(dict(ra=1.0, rb=1.0, ga=len(a), gb=len(b), art=ENDE))
    st.sort(key=lambda x: x["ra"])  # stable
    for w0, w1 in zip(st, st[1:]):
        if (w1["ra"] - w0["ra"]) * la < 1e-3 or (w1["rb"] - w0["rb"]) * lb < 1e-3:
            raise PlanFehler(f"seam stations at {w0['ra'] * la:.3f} mm coincide")
    abschnitte = []
    for w0, w1 in zip(st, st[1:]):
        _, (sa, s0a) = _orte(a, ka, la, w0["ra"], w0["ga"])
        (sa1, s1a), _ = _orte(a, ka, la, w1["ra"], w1["ga"])
        _, (sb, s0b) = _orte(b, kb, lb, w0["rb"], w0["gb"])
        (sb1, s1b), _ = _orte(b, kb, lb, w1["rb"], w1["gb"])
        if sa != sa1 or sb != sb1:
            raise PlanFehler(f"the section from {w0['ra'] * la:.3f} mm runs across an edge boundary")
        g = Gruppe([Glied(a[sa]["teil"], a[sa]["kante"], s0a, s1a), Glied(b[sb]["teil"], b[sb]["kante"], s0b, s1b)])
        unterteile(g, teile, p)
        gruppen.append(g)
        abschnitte.append(((a[sa]["teilstueck"], b[sb]["teilstueck"]), (a[sa]["teil"], b[sb]["teil"]),
                           dict(gruppe=len(gruppen) - 1, sigma_a=(w0["ra"] * la, w1["ra"] * la),
                                sigma_b=(w0["rb"] * lb, w1["rb"] * lb))))
    teilnaehte, letzt = [], None
    for schl, (ta, tb), ab in abschnitte:
        if letzt != schl:
            teilnaehte.append(dict(teil_a=ta, teil_b=tb, abschnitte=[]))
            letzt = schl
        teilnaehte[-1]["abschnitte"].append(ab)
    stellen = []
    for s in st:
        stellen.append(dict(art=s["art"], knoten_a=None if s["ga"] is None else _knoten(a, teile, s["ga"]),
                            knoten_b=None if s["gb"] is None else _knoten(b, teile, s["gb"]),
                            rho_a=s["ra"], rho_b=s["rb"], versatz_mm=(s["ra"] - s["rb"]) * lm))
    return dict(la=la, lb=lb, stellen=stellen, teilnaehte=teilnaehte)


def _bits(s: float) -> int:
    return struct.unpack("<Q", struct.pack("<d", s))[0]


def schluessel(teile, teil, kante, s):
    """rand.rs:536-545: the end of an edge is the start of the next."""
    t = teile[teil]
    if s >= t[kante].laenge:
        return ((kante + 1) % len(t), 0)
    if s <= 0.0:
        return (kante, 0)
    return (kante, _bits(s))


def planen(teile_roh, naehte, p=None):
    """Randplan::neu (rand.rs:503-533) and the pairs of lib.rs:351-381.
    teile_roh: per piece a list of edges (polylines); naehte: list of (side A,
    side B), each side a list of (piece, [edges], backwards)."""
    p = {**dict(kante_mm=10.0, sehne_mm=0.1, laenge_mm=0.1, fang_mm=2.0, min_rand_mm=0.5), **(p or {})}
    teile = []
    for ti, t in enumerate(teile_roh):
        zuege = [Kantenzug(k) for k in t]
        for k, z in enumerate(zuege):
            if z.laenge <= 1e-6:
                raise PlanFehler(f"piece {ti}: boundary edge {k} has no length")
        teile.append(zuege)
    gruppen, belegt, plaene = [], set(), []
    for n, (sa, sb) in enumerate(naehte):
        seiten = []
        for seite in (sa, sb):
            out = []
            for nr, (teil, kanten, rueck) in enumerate(seite):
                for k in kanten:
                    out.append(dict(teil=teil, kante=k, rueck=bool(rueck), laenge=teile[teil][k].laenge,
                                    teilstueck=nr, anfang=teile[teil][k].laenge if rueck else 0.0,
                                    ende=0.0 if rueck else teile[teil][k].laenge))
            if not out:
                raise PlanFehler(f"seam {n}: a side without edges")
            seiten.append(out)
        for x in seiten[0] + seiten[1]:
            if (x["teil"], x["kante"]) in belegt:
                raise PlanFehler(f"seam {n}: piece {x['teil']} edge {x['kante']} is in two seams")
            belegt.add((x["teil"], x["kante"]))
        plaene.append(naht_planen(seiten[0], seiten[1], teile, p, gruppen))
    for ti, t in enumerate(teile):
        for k, z in enumerate(t):
            if (ti, k) not in belegt:
                g = Gruppe([Glied(ti, k, 0.0, z.laenge)])
                unterteile(g, teile, p)
                gruppen.append(g)
    # boundary index per piece (rand.rs:547-583)
    rand = []
    for ti, t in enumerate(teile):
        menge = {}
        for g in gruppen:
            for gl in g.glieder:
                if gl.teil != ti:
                    continue
                for f in g.f:
                    s = gl.s(f)
                    k = schluessel(teile, ti, gl.kante, s)
                    if k not in menge:
                        menge[k] = (k[0], 0.0) if k[1] == 0 else (gl.kante, s)
        schl = sorted(menge)
        index = {k: i for i, k in enumerate(schl)}
        orte = [menge[k] for k in schl]
        rand.append(dict(kante=[o[0] for o in orte], s=[o[1] for o in orte],
                         xy=[t[o[0]].ort(o[1]) for o in orte], knoten=[index[(k, 0)] for k in range(len(t))],
                         index=index))
    # pairs (lib.rs:351-381)
    for np_ in plaene:
        for tn in np_["teilnaehte"]:
            paare, rho = [], []
            for k, ab in enumerate(tn["abschnitte"]):
                g = gruppen[ab["gruppe"]]
                for j, f in enumerate(g.f):
                    if k > 0 and j == 0:
                        continue
                    ga, gb = g.glieder
                    paare.append((rand[tn["teil_a"]]["index"][schluessel(teile, ga.teil, ga.kante, ga.s(f))],
                                  rand[tn["teil_b"]]["index"][schluessel(teile, gb.teil, gb.kante, gb.s(f))]))
                    rho.append(((ab["sigma_a"][0] + f * (ab["sigma_a"][1] - ab["sigma_a"][0])) / np_["la"],
                                (ab["sigma_b"][0] + f * (ab["sigma_b"][1] - ab["sigma_b"][0])) / np_["lb"]))
            tn["paare"], tn["rho"] = paare, rho
    return dict(teile=rand, naehte=plaene, gruppen=gruppen)
