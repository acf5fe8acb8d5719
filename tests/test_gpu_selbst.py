"""The device's cloth self-collision against the FP64 oracle (selbstkollision.rs):
candidate lists of the spatial hash, springs and the approach damping, cuts
with their regions, and whole steps with impacts.

Lists are discrete results of FP64 arithmetic in the reference's order and
must agree element for element; the FP64 build (forces, right-hand side) to
the bit or FP64 rounding; steps to the bounds stated per test.
"""
import numpy as np
import pytest

import oracle as O
from garmentcraft_amd import szenen
from garmentcraft_amd.geraet import GeraetSzene, ImplizitParameter
from garmentcraft_amd.material import RADICAL_RUBBER_04
from garmentcraft_amd.szene import Simparameter, Teil, bauen

pytestmark = pytest.mark.gpu

H = 1.0 / 150.0


def zwei_lagen(lage_oben, versatz=(0.0, 0.0)):
    """Two 20 x 20 cm squares at 10 mm, the lower one held (tests/selbstkollision.rs:14-31)."""
    pts, tris, _, _ = szenen.rechteck_netz(200.0, 200.0, 10.0)
    teile = [Teil("unten", pts, tris, fadenlauf=(1.0, 0.0), material=RADICAL_RUBBER_04, reduktion=0.0),
             Teil("oben", pts, tris, fadenlauf=(1.0, 0.0), material=RADICAL_RUBBER_04, reduktion=0.0)]

    def lage(ti, p):
        x, y = p[:, 0] * 1e-3, p[:, 1] * 1e-3
        if ti == 0:
            return np.column_stack([x, y, np.zeros_like(x)])
        x, y = x + versatz[0], y + versatz[1]
        return np.column_stack([x, y, lage_oben(x, y)])
    s = bauen(teile, [], lage, [], Simparameter(bild_dt=1.0 / 30.0, daempfung=1.0))
    w = s.w.copy()
    w[s.teil_von_punkt == 0] = 0.0
    s.w = w
    return s


def zerknittert():
    rng = np.random.default_rng(12)
    s = zwei_lagen(lambda x, y: 0.0021 + 0.0015 * np.sin(x * 60.0) * np.cos(y * 45.0))
    s.x = s.x + rng.normal(scale=3e-4, size=s.x.shape)
    s.v = rng.normal(scale=0.02, size=s.x.shape)
    s.v[s.w <= 0.0] = 0.0
    return s


P = dict(schritte_je_bild=1, cg_max=10000, cg_toleranz=1e-10, selbstkollision=1)


def geraet(s, praezision="fp64", **kw):
    p = dict(P, **kw)
    p["selbstkollision"] = bool(p["selbstkollision"])
    return GeraetSzene(s, ImplizitParameter(**p), praezision=praezision)


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(a)), 1e-300))


def test_naehe_und_federn_wie_orakel():
    s = zerknittert()
    o = O.OrakelSim(s, O.param(**P))
    g = geraet(s)
    erwartet = o.selbst_naehe()
    g.bauen(H)
    o.sys_bauen(H)
    n = g.naehe()
    print("pairs", n["art"].size, "point-triangle", int((n["art"] == 0).sum()))
    assert (erwartet["art"] == 0).sum() > 50 and (erwartet["art"] == 1).sum() > 50
    for k in ("art", "punkte", "gewichte", "normale", "abstand", "ruhe"):
        assert np.array_equal(erwartet[k], n[k]), k
    fa, fb = o.sys_federn(), g.federn()
    for k in ("punkte", "gewichte", "kraft", "block"):
        assert np.array_equal(fa[k], fb[k]), k
    assert np.array_equal(o.v(), g.v())  # the approach damping, in list order
    a, b = o.sys_vektoren(), g.system()
    for k in ("f", "rhs"):
        print(k, rel(a[k], b[k]))
        assert rel(a[k], b[k]) <= 1e-13, k
    assert rel(a["pinv"].reshape(-1, 9), b["pinv"].reshape(-1, 9)) <= 1e-13
    p = np.random.default_rng(3).normal(size=(s.n, 3))
    assert rel(o.sys_mal_a(p), g.mal_a(p)) <= 1e-14
    st = g.statistik()
    assert st["naehe_punkt_dreieck"] == (erwartet["art"] == 0).sum()
    assert st["naehe_kante_kante"] == (erwartet["art"] == 1).sum()
    g.close()


def grob_ueber_fein(lage_oben):
    """A fine 20 x 20 cm square at 10 mm, held, under a coarse one of 100 mm
    triangles: a coarse triangle's box spans about 18 x 18 x 2 proximity cells
    of 6 mm (beyond 512, the hash's block-wide path; the reference accepts up
    to 64 cells per axis, selbstkollision.rs:268-272)."""
    fein, tf, _, _ = szenen.rechteck_netz(200.0, 200.0, 10.0)
    grob, tg, _, _ = szenen.rechteck_netz(200.0, 200.0, 100.0)
    teile = [Teil("fein", fein, tf, fadenlauf=(1.0, 0.0), material=RADICAL_RUBBER_04, reduktion=0.0),
             Teil("grob", grob, tg, fadenlauf=(1.0, 0.0), material=RADICAL_RUBBER_04, reduktion=0.0)]

    def lage(ti, p):
        x, y = p[:, 0] * 1e-3, p[:, 1] * 1e-3
        if ti == 0:
            return np.column_stack([x, y, np.zeros_like(x)])
        return np.column_stack([x + 0.0031, y + 0.0043, lage_oben(x, y)])
    s = bauen(teile, [], lage, [], Simparameter(bild_dt=1.0 / 30.0, daempfung=1.0))
    w = s.w.copy()
    w[s.teil_von_punkt == 0] = 0.0
    s.w = w
    return s


@pytest.mark.parametrize("form", ["flach", "schraeg"])
def test_grosse_quader_wie_orakel(form):
    """Hash entries whose boxes cover more than 512 cells: proximity pairs,
    springs, cuts and their regions equal to the oracle's, element for element."""
    oben = (lambda x, y: 0.0021 + 0.0 * x) if form == "flach" else (lambda x, y: -0.0353 + 0.4 * x)
    s = grob_ueber_fein(oben)
    o = O.OrakelSim(s, O.param(**P))
    g = geraet(s)
    erwartet = o.selbst_naehe()
    g.bauen(H)
    o.sys_bauen(H)
    n = g.naehe()
    print(form, "pairs", n["art"].size, "point-triangle", int((n["art"] == 0).sum()))
    assert n["art"].size > 20
    for k in ("art", "punkte", "gewichte", "normale", "abstand", "ruhe"):
        assert np.array_equal(erwartet[k], n[k]), k
    fa, fb = o.sys_federn(), g.federn()
    for k in ("punkte", "gewichte", "kraft", "block"):
        assert np.array_equal(fa[k], fb[k]), k
    sa, sb = o.selbst_schnitte(), g.schnitte()
    print(form, "cuts", len(sb["t"]))
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), k
    if form == "schraeg":
        assert len(sb["t"]) > 5
    assert g.statistik()["selbst_listen_faktor"] == 1
    g.close()


def test_schnitte_und_gebiete_wie_orakel():
    s = zwei_lagen(lambda x, y: -0.0053 + 0.1 * x, versatz=(0.0031, 0.0043))
    o = O.OrakelSim(s, O.param(**P))
    g = geraet(s)
    erwartet = o.selbst_schnitte()
    g.bauen(H)
    o.sys_bauen(H)
    c = g.schnitte()
    print("cuts", c["t"].size, "regions", np.unique(c["wurzel"]).size)
    assert erwartet["t"].size > 10
    for k in ("kante_dreieck", "t", "w", "richtung", "gewicht", "wurzel"):
        assert np.array_equal(erwartet[k], c[k]), k
    fa, fb = o.sys_federn(), g.federn()
    for k in ("punkte", "gewichte", "kraft", "block"):
        assert np.array_equal(fa[k], fb[k]), k
    a, b = o.sys_vektoren(), g.system()
    assert rel(a["rhs"], b["rhs"]) <= 1e-13
    g.close()


@pytest.mark.parametrize("praezision,tol_x", [("fp64", 1e-9), ("fp32", 1e-4)])
def test_schritte_mit_selbstkollision_wie_orakel(praezision, tol_x):
    """Single steps from the oracle's state (x, v, warm start). The upper layer
    falls at 2 m/s: the first step crosses the lower layer (impacts and their
    response), the next ones carry proximity springs, cuts, regions and push
    springs. The searches run in FP64 in both modes, so their counts agree
    exactly; the impacts depend on the solved positions and are compared in
    FP64 only."""
    s = zwei_lagen(lambda x, y: 0.008 + 0.002 * np.sin(x * 40.0), versatz=(0.0031, 0.0043))
    v = np.zeros_like(s.x)
    v[s.w > 0, 2] = -2.0
    s.v = v
    rtol = 1e-10 if praezision == "fp64" else 1e-6
    o = O.OrakelSim(s, O.param(**dict(P, cg_toleranz=rtol)))
    g = geraet(s, praezision, cg_toleranz=rtol)
    summe = {q: 0 for q in ("naehe_punkt_dreieck", "naehe_kante_kante", "durchdringungen", "selbst_stoesse")}
    worst = 0.0
    for k in range(10):
        x, v = o.lage()
        g.set_state(x, v)
        g.set_dv_vorher(o.dv_vorher())
        vor = o.statistik()
        g.statistik(reset=True)
        o.schritt(H)
        g.schritte(H, 1)
        nach = o.statistik()
        sg = g.statistik()
        for q in summe:
            summe[q] += nach[q] - vor[q]
            if q != "selbst_stoesse" or praezision == "fp64":
                assert nach[q] - vor[q] == sg[q], (k, q, nach[q] - vor[q], sg[q])
        xa, _ = o.lage()
        xb, _ = g.lage()
        d = float(np.max(np.linalg.norm(xa - xb, axis=1)))
        worst = max(worst, d)
        assert d <= tol_x, (k, d)
    print(praezision, summe, "worst", worst)
    assert summe["selbst_stoesse"] > 100 and summe["durchdringungen"] > 100 and summe["naehe_punkt_dreieck"] > 100
    g.close()


def test_drape_mit_selbstkollision_graph_wie_eager():
    """The self-collision kernels replay in a captured graph to the bit."""
    sdf = szenen.koerper_sdf(aufloesung=96, spacing=0.008)
    cfg = szenen.DrapeKonfig(spalten=2, zeilen=2, zelle_mm=160.0, hoehe_m=0.335, jitter_mm=10.0,
                              versatz_mm=(0.0, 0.0))
    s, _ = szenen.drape(cfg, sdf=sdf)
    a = geraet(s, "fp32", cg_toleranz=1e-6)
    b = geraet(s, "fp32", cg_toleranz=1e-6)
    a.schritte(1.0 / 120.0, 20)
    b.schritte_graph(1.0 / 120.0, 20)
    xa, va = a.lage()
    xb, vb = b.lage()
    assert np.array_equal(xa, xb) and np.array_equal(va, vb)
    st = a.statistik()
    print(st)
    a.close(); b.close()


def test_parameter_zwischen_graph_laeufen_wie_eager():
    """Self-collision switched off and on again, its impact iterations and its
    spring stiffness changed between graph replays: the replays follow the new
    parameters as eager steps do, to the bit (a graph captured with the old
    launch sequence is captured again)."""
    from dataclasses import replace
    sdf = szenen.koerper_sdf(aufloesung=96, spacing=0.008)
    cfg = szenen.DrapeKonfig(spalten=2, zeilen=2, zelle_mm=160.0, hoehe_m=0.335, jitter_mm=10.0,
                              versatz_mm=(0.0, 0.0))
    s, _ = szenen.drape(cfg, sdf=sdf)
    a = geraet(s, "fp32", cg_toleranz=1e-6)
    b = geraet(s, "fp32", cg_toleranz=1e-6)
    h = 1.0 / 120.0
    p0 = a.param
    folge = [p0, replace(p0, selbstkollision=False), replace(p0, selbst_stoss_iterationen=3),
             replace(p0, selbst_steifigkeit=3e-3)]
    for p in folge:
        a.set_param(p)
        b.set_param(p)
        a.schritte(h, 8)
        b.schritte_graph(h, 8)
        xa, va = a.lage()
        xb, vb = b.lage()
        assert np.array_equal(xa, xb) and np.array_equal(va, vb), p
    st = a.statistik()
    assert st["naehe_punkt_dreieck"] + st["naehe_kante_kante"] > 0
    a.close(); b.close()


def test_drape_baut_wie_orakel():
    """The drape after 40 oracle steps with self-collision: the FP64 device build
    gives the oracle's proximity list, springs and right-hand side."""
    sdf = szenen.koerper_sdf(aufloesung=96, spacing=0.008)
    cfg = szenen.DrapeKonfig(spalten=2, zeilen=2, zelle_mm=160.0, hoehe_m=0.335, jitter_mm=10.0,
                              versatz_mm=(0.0, 0.0))
    s, _ = szenen.drape(cfg, sdf=sdf)
    lauf = O.OrakelSim(s, O.param(**dict(P, cg_toleranz=1e-6)))
    for _ in range(40):
        lauf.schritt(1.0 / 120.0)
    x, v = lauf.lage()
    # a fresh oracle at that state: the run counter of the cut triangles
    # (selbstkollision.rs:1024-1033) starts at 1 on both sides
    o = O.OrakelSim(s, O.param(**dict(P, cg_toleranz=1e-6)))
    o.setze_lage(x, v)
    g = geraet(s, "fp64", cg_toleranz=1e-6)
    g.set_state(x, v)
    erwartet = o.selbst_naehe()
    g.bauen(1.0 / 120.0)
    o.sys_bauen(1.0 / 120.0)
    n = g.naehe()
    print("pairs", n["art"].size)
    assert n["art"].size > 20
    for k in ("art", "punkte", "gewichte", "normale", "abstand", "ruhe"):
        assert np.array_equal(erwartet[k], n[k]), k
    assert np.array_equal(o.v(), g.v())
    fa, fb = o.sys_federn(), g.federn()
    print("springs", fa["punkte"].shape[0])
    for k in ("punkte", "gewichte", "kraft", "block"):
        assert np.array_equal(fa[k], fb[k]), k
    a, b = o.sys_vektoren(), g.system()
    assert rel(a["rhs"], b["rhs"]) <= 1e-12
    g.close()
