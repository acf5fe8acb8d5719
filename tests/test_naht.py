"""Seam pairing by length fraction (gc_naht_planen) against the oracle's
restatement of klon3d-mesh rand.rs, and both against the reference's own
known-answer tests:

- rand.rs:608-656 (kantenzug_ort_und_sehne, abbildung_ueber_anker,
  unterteilung_nach_kruemmung);
- tests/naht_synthetisch.rs (two rectangles: notches 1 mm apart merged within
  fang_mm, a node only on one side; with fang 0.5 mm each side gets its own
  points).

The plan's decisions are integers (merged or not, split or not, point
indices), so the library and the oracle must agree to the bit on random
patterns too. Host code only: no device is used."""
import math

import numpy as np
import pytest

import oracle.naht as O
from garmentcraft_amd import naehte as N


# --- rand.rs unit tests, on the oracle ------------------------------------
def test_kantenzug_ort_und_sehne():
    k = O.Kantenzug([[0.0, 0.0], [10.0, 0.0], [10.0, 0.0], [10.0, 10.0]])
    assert len(k.punkte) == 3 and k.laenge == 20.0
    assert k.ort(15.0) == (10.0, 5.0) and k.ort(25.0) == (10.0, 10.0)
    h, v = k.sehne(5.0, 15.0)
    assert abs(h - math.hypot(5.0, 5.0) / 2.0) < 1e-12
    assert abs(v - (10.0 - math.sqrt(50.0))) < 1e-12
    assert k.sehne(0.0, 10.0) == (0.0, 0.0)


def test_abbildung_ueber_anker():
    anker = [(0.0, 0.0), (0.5, 0.6), (1.0, 1.0)]
    assert abs(O.abbilden(anker, 0.25, False) - 0.3) < 1e-12
    assert abs(O.abbilden(anker, 0.8, True) - 0.75) < 1e-12
    assert O.abbilden(anker, 1.0, False) == 1.0


def _viertelkreis_teil():
    bogen = [[20.0 * math.cos(w), 20.0 * math.sin(w)] for w in (np.arange(401) / 400.0 * math.pi / 2)]
    return [bogen, [[0.0, 20.0], [0.0, 0.0]], [[0.0, 0.0], [20.0, 0.0]]]


def test_unterteilung_nach_kruemmung():
    teile = [[O.Kantenzug(k) for k in _viertelkreis_teil()]]
    p = dict(kante_mm=10.0, sehne_mm=0.1, laenge_mm=0.1, fang_mm=2.0, min_rand_mm=0.5)
    g = O.Gruppe([O.Glied(0, 0, 0.0, teile[0][0].laenge)])
    O.unterteile(g, teile, p)
    # chord height s^2/8R <= 0.1 mm -> s <= 4 mm: 31.4 mm of arc in at least 8 pieces
    assert len(g.f) - 1 >= 8
    kz = teile[0][0]
    for f0, f1 in zip(g.f, g.f[1:]):
        h, v = kz.sehne(f0 * kz.laenge, f1 * kz.laenge)
        assert h <= p["sehne_mm"] and v <= p["laenge_mm"] * (f1 - f0)
    gerade = O.Gruppe([O.Glied(0, 1, 20.0, 0.0)])
    O.unterteile(gerade, teile, p)
    assert gerade.f == [0.0, 0.5, 1.0]


def test_unterteilung_der_bibliothek_wie_orakel():
    """The same piece planned by the library: its boundary points are the
    oracle's to the bit, and the arc is split as finely as above."""
    teil = _viertelkreis_teil()
    plan = N.planen([teil], [])
    ref = O.planen([teil], [])
    r = plan.teile[0]
    assert np.array_equal(r.kante, ref["teile"][0]["kante"])
    assert np.array_equal(r.s, ref["teile"][0]["s"])
    assert np.array_equal(r.xy, np.array(ref["teile"][0]["xy"]))
    assert np.array_equal(r.knoten, ref["teile"][0]["knoten"])
    assert int(np.count_nonzero(r.kante == 0)) >= 8
    assert int(np.count_nonzero(r.kante == 1)) == 2 and int(np.count_nonzero(r.kante == 2)) == 2


# --- tests/naht_synthetisch.rs ----------------------------------------------
def _rechtecke():
    """A: (0,0)-K(40,0)-(100,0)-(100,50)-(0,50); B: (0,100)-L(41,100)-M(70,100)-
    (100,100)-(100,150)-(0,150); nodes A1 K A2 A3 A4 and B1 L M B2 B3 B4."""
    def umriss(p):
        return [[p[i], p[(i + 1) % len(p)]] for i in range(len(p))]
    a = umriss([(0, 0), (40, 0), (100, 0), (100, 50), (0, 50)])
    b = umriss([(0, 100), (41, 100), (70, 100), (100, 100), (100, 150), (0, 150)])
    return [a, b]


def _naht_n1():
    return N.Naht(a=[N.Lauf(0, [0, 1])], b=[N.Lauf(1, [0, 1, 2])])


KNOTEN = {(0, 0): "A1", (0, 1): "K", (0, 2): "A2", (0, 3): "A3", (0, 4): "A4",
          (1, 0): "B1", (1, 1): "L", (1, 2): "M", (1, 3): "B2", (1, 4): "B3", (1, 5): "B4"}


def _beide(fang):
    teile = _rechtecke()
    plan = N.planen(teile, [_naht_n1()], N.Parameter(fang_mm=fang))
    ref = O.planen(teile, [([(0, [0, 1], False)], [(1, [0, 1, 2], False)])], dict(fang_mm=fang))
    return teile, plan, ref


def _gleich(plan, ref):
    """Every output of the library equals the oracle's, bit for bit."""
    assert len(plan.teile) == len(ref["teile"])
    for r, q in zip(plan.teile, ref["teile"]):
        assert np.array_equal(r.kante, q["kante"]) and np.array_equal(r.s, q["s"])
        assert np.array_equal(r.xy, np.array(q["xy"]).reshape(-1, 2)) and np.array_equal(r.knoten, q["knoten"])
    assert len(plan.naehte) == len(ref["naehte"])
    for n, q in zip(plan.naehte, ref["naehte"]):
        assert n.laenge_a_mm == q["la"] and n.laenge_b_mm == q["lb"]
        assert len(n.stellen) == len(q["stellen"])
        for s, t in zip(n.stellen, q["stellen"]):
            assert (s.art, s.knoten_a, s.knoten_b, s.rho_a, s.rho_b, s.versatz_mm) == \
                (t["art"], t["knoten_a"], t["knoten_b"], t["rho_a"], t["rho_b"], t["versatz_mm"])
        assert len(n.teilnaehte) == len(q["teilnaehte"])
        for tn, tq in zip(n.teilnaehte, q["teilnaehte"]):
            assert (tn.teil_a, tn.teil_b) == (tq["teil_a"], tq["teil_b"])
            assert np.array_equal(tn.paare, np.array(tq["paare"], dtype=np.int32).reshape(-1, 2))
            assert np.array_equal(tn.rho, np.array(tq["rho"]).reshape(-1, 2))


def test_kerben_zusammengelegt_und_eigener_punkt():
    teile, plan, ref = _beide(2.0)
    _gleich(plan, ref)
    n = plan.naehte[0]
    arten = [(N.ARTEN[s.art], KNOTEN.get(s.knoten_a, ""), KNOTEN.get(s.knoten_b, "")) for s in n.stellen]
    assert arten == [("Ende", "A1", "B1"), ("Gemeinsam", "K", "L"), ("NurB", "", "M"), ("Ende", "A2", "B2")]
    # the merged notch is 1 mm off
    assert abs(n.stellen[1].versatz_mm - (-1.0)) < 1e-9
    # M (70 mm) lies between the anchors (40|41) and (100|100): on A at 40 + 29/59 60
    soll = 40.0 + 29.0 / 59.0 * 60.0
    assert abs(n.stellen[2].rho_a * 100.0 - soll) < 1e-9
    assert len(n.teilnaehte) == 1
    paare = n.teilnaehte[0].paare
    # sections 40|41 mm (5 pieces), 29.49|29 mm (3), 30.51|30 mm (4): 13 pairs
    assert len(paare) == 13
    ka, kb = plan.teile[0].knoten, plan.teile[1].knoten
    assert any(a == ka[1] and b == kb[1] for a, b in paare)  # K with L
    m = [a for a, b in paare if b == kb[2]]
    assert len(m) == 1
    pa = plan.teile[0].xy[m[0]]
    assert abs(pa[0] - soll) < 1e-6 and abs(pa[1]) < 1e-9
    # free straight edges stay at <= 10 mm with no chord error: A's 100 mm edge in 10
    assert int(np.count_nonzero(plan.teile[0].kante == 3)) == 10


def test_enger_fang_gibt_beiden_seiten_eigene_punkte():
    teile, plan, ref = _beide(0.5)
    _gleich(plan, ref)
    n = plan.naehte[0]
    assert [s.art for s in n.stellen] == [N.ENDE, N.NUR_A, N.NUR_B, N.NUR_B, N.ENDE]
    # with no anchor the pure fraction rule holds: K at 40 mm on B too
    assert abs(n.stellen[1].rho_b - 0.4) < 1e-12 and abs(n.stellen[1].versatz_mm) < 1e-9
    ka, kb = plan.teile[0].knoten, plan.teile[1].knoten
    paare = n.teilnaehte[0].paare
    zu_k = [b for a, b in paare if a == ka[1]]
    zu_l = [a for a, b in paare if b == kb[1]]
    gb, ga = plan.teile[1].xy[zu_k[0]], plan.teile[0].xy[zu_l[0]]
    assert abs(gb[0] - 40.0) < 1e-9 and abs(gb[1] - 100.0) < 1e-9
    assert abs(ga[0] - 41.0) < 1e-9 and abs(ga[1]) < 1e-9


def test_fehler_wie_orakel():
    """An edge in two seams is refused by both, as the reference's Err."""
    teile = _rechtecke()
    with pytest.raises(Exception, match="two seams"):
        N.planen(teile, [_naht_n1(), N.Naht(a=[N.Lauf(0, [1])], b=[N.Lauf(1, [4])])])
    with pytest.raises(O.PlanFehler, match="two seams"):
        O.planen(teile, [([(0, [0, 1], False)], [(1, [0, 1, 2], False)]), ([(0, [1], False)], [(1, [4], False)])])


# --- random patterns: library == oracle, and what the plan must hold --------
def _zufallsteil(rng, mitte, n_knoten):
    """A star-shaped outline around `mitte` with n_knoten nodes; each edge a
    polyline, some straight, some bulging (curves)."""
    w = np.sort(rng.uniform(0, 2 * np.pi, n_knoten))
    r = rng.uniform(40.0, 120.0, n_knoten)
    knoten = [(mitte[0] + r[i] * np.cos(w[i]), mitte[1] + r[i] * np.sin(w[i])) for i in range(n_knoten)]
    kanten = []
    for i in range(n_knoten):
        a, b = np.array(knoten[i]), np.array(knoten[(i + 1) % n_knoten])
        m = int(rng.integers(1, 30))
        t = np.linspace(0.0, 1.0, m + 1)
        d = b - a
        nrm = np.array([-d[1], d[0]]) / max(np.linalg.norm(d), 1e-9)
        wolbung = rng.choice([0.0, rng.uniform(-8.0, 8.0)])
        pts = a[None, :] + t[:, None] * d[None, :] + (wolbung * 4 * t * (1 - t))[:, None] * nrm[None, :]
        kanten.append(pts.tolist())
    return kanten


@pytest.mark.parametrize("seed", range(24))
def test_zufallsmuster_wie_orakel(seed):
    rng = np.random.default_rng(seed)
    n_teile = int(rng.integers(2, 5))
    teile = [_zufallsteil(rng, (300.0 * i, 0.0), int(rng.integers(4, 9))) for i in range(n_teile)]
    # seams between consecutive pieces: a run of 1-3 consecutive edges on each
    # side, sometimes backwards, sometimes two runs on one side
    naehte, naehte_o, frei = [], [], [list(range(len(t))) for t in teile]
    for i in range(n_teile - 1):
        seiten, seiten_o = [], []
        for t in (i, i + 1):
            k0 = int(rng.integers(0, len(teile[t])))
            lauf = [k for k in range(k0, min(k0 + int(rng.integers(1, 4)), len(teile[t]))) if k in frei[t]]
            if not lauf:
                break
            for k in lauf:
                frei[t].remove(k)
            rueck = bool(rng.integers(0, 2))
            if rueck:
                lauf = lauf[::-1]
            seiten.append([N.Lauf(t, lauf, rueck)])
            seiten_o.append([(t, lauf, rueck)])
        if len(seiten) == 2 and i + 2 < n_teile and rng.integers(0, 2):
            # a second run on side A from another piece: two sub-seams
            t = i + 2
            k0 = int(rng.integers(0, len(teile[t])))
            if k0 in frei[t]:
                frei[t].remove(k0)
                seiten[0].append(N.Lauf(t, [k0], False))
                seiten_o[0].append((t, [k0], False))
        if len(seiten) == 2:
            naehte.append(N.Naht(a=seiten[0], b=seiten[1]))
            naehte_o.append((seiten_o[0], seiten_o[1]))
    p = N.Parameter(fang_mm=float(rng.choice([0.5, 2.0, 8.0])), kante_mm=float(rng.choice([5.0, 10.0])))
    po = dict(fang_mm=p.fang_mm, kante_mm=p.kante_mm)
    try:
        ref = O.planen(teile, naehte_o, po)
    except O.PlanFehler as e:
        with pytest.raises(Exception):
            N.planen(teile, naehte, p)
        print("both refuse:", e)
        return
    plan = N.planen(teile, naehte, p)
    _gleich(plan, ref)
    # what the plan must hold, checked from the outputs
    for n in plan.naehte:
        lm = 0.5 * (n.laenge_a_mm + n.laenge_b_mm)
        rho = [(s.rho_a, s.rho_b) for s in n.stellen]
        assert rho[0] == (0.0, 0.0) and rho[-1] == (1.0, 1.0)
        assert all(x0 < x1 and y0 < y1 for (x0, y0), (x1, y1) in zip(rho, rho[1:]))
        for s in n.stellen:
            if s.art == N.GEMEINSAM:
                assert abs(s.versatz_mm) <= max(p.fang_mm, 1e-3) + 1e-12
        for tn in n.teilnaehte:
            # pairs advance along both sides; each pair's points sit at the
            # stated length fraction of their side
            assert np.all(np.diff(tn.rho[:, 0]) > 0) and np.all(np.diff(tn.rho[:, 1]) > 0)
    for r in plan.teile:
        # boundary points in key order: edge, then arc length
        key = r.kante.astype(np.float64) * 1e9 + r.s
        assert np.all(np.diff(key) > 0)
