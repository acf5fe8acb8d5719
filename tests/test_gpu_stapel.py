"""Batched substeps (gc_batch_*): several systems in one persistent PCG launch
per solve pass.

With the same launch shape per system (blocks and threads), a system's
reductions group their block partials the same way in a batch as alone, so a
batched substep equals gc_step on each system to the bit; the graph replay of
the batched substep equals the eager one to the bit."""
import numpy as np
import pytest

from garmentcraft_amd import szenen
from garmentcraft_amd.geraet import GeraetSzene, GeraetStapel, ImplizitParameter

pytestmark = pytest.mark.gpu

H = 1.0 / 120.0


@pytest.fixture(scope="module")
def szenen_satz():
    sdf = szenen.koerper_sdf(aufloesung=96, spacing=0.008)
    a, _ = szenen.drape(szenen.DrapeKonfig(spalten=2, zeilen=2, zelle_mm=160.0, hoehe_m=0.335, jitter_mm=10.0,
                                           versatz_mm=(0.0, 0.0)), sdf=sdf)
    b, _ = szenen.drape(szenen.DrapeKonfig(seed=7, spalten=3, zeilen=1, zelle_mm=150.0, hoehe_m=0.33, jitter_mm=10.0,
                                           versatz_mm=(0.0, 0.0)), sdf=sdf)
    rohr = szenen.haengende_roehre()
    return [a, b, rohr]


def param(selbst):
    return ImplizitParameter(schritte_je_bild=1, cg_max=10000, cg_toleranz=1e-6, selbstkollision=selbst)


def systeme(satz, praezision, formen=None, quelle=True):
    """One context per scene; the second drape shares the first one's SDF."""
    out = []
    for k, s in enumerate(satz):
        kw = {}
        if formen is not None:
            kw = dict(bloecke=int(formen[0][k]), threads=formen[1])
        q = out[0] if (quelle and k == 1) else None
        out.append(GeraetSzene(s, param(k == 0), praezision=praezision, koerper_quelle=q, **kw))
    return out


@pytest.mark.parametrize("praezision", ["fp32", "fp64"])
def test_stapel_wie_einzeln(szenen_satz, praezision):
    gs = systeme(szenen_satz, praezision)
    st = GeraetStapel(gs, threads=1024)
    nb, threads = st.bloecke()
    print(praezision, "blocks per system", nb.tolist(), "threads", threads)
    assert (nb >= 1).all()
    einzeln = systeme(szenen_satz, praezision, formen=(nb, threads), quelle=False)
    for k in range(12):
        st.schritte(H, 1)
        for g in einzeln:
            g.schritte(H, 1)
    for a, b in zip(gs, einzeln):
        xa, va = a.lage()
        xb, vb = b.lage()
        assert np.isfinite(xa).all()
        assert np.array_equal(xa, xb) and np.array_equal(va, vb)
        sa, sb = a.statistik(), b.statistik()
        for q in ("cg_iterationen", "loesungen", "kontakte", "naehe_punkt_dreieck", "naehe_kante_kante"):
            assert sa[q] == sb[q], q
    print("iterations", [g.statistik()["cg_iterationen"] for g in gs])
    st.close()
    for g in gs + einzeln:
        g.close()


def test_stapel_graph_wie_eager(szenen_satz):
    a = systeme(szenen_satz, "fp32")
    b = systeme(szenen_satz, "fp32")
    sa, sb = GeraetStapel(a), GeraetStapel(b)
    sa.schritte(H, 10)
    sb.schritte_graph(H, 10)
    for x, y in zip(a, b):
        xa, va = x.lage()
        xb, vb = y.lage()
        assert np.array_equal(xa, xb) and np.array_equal(va, vb)
    sa.close(); sb.close()


def test_autotune_laesst_den_zustand(szenen_satz):
    gs = systeme(szenen_satz, "fp32")
    st = GeraetStapel(gs)
    st.schritte(H, 3)
    vor = [g.lage() for g in gs]
    stat = [g.statistik() for g in gs]
    wahl = st.autotune(H)
    print("autotune", wahl, st.bloecke())
    assert wahl["threads"] in (256, 512, 1024) and wahl["bloecke"] >= len(gs) and wahl["ms_loesung"] > 0
    for g, (x, v), s0 in zip(gs, vor, stat):
        x1, v1 = g.lage()
        assert np.array_equal(x, x1) and np.array_equal(v, v1)
        assert g.statistik()["cg_iterationen"] == s0["cg_iterationen"]
    st.schritte(H, 3)
    for g in gs:
        assert np.isfinite(g.lage()[0]).all()
    st.close()
