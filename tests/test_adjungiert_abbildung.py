"""The per-panel material map of the adjoint rollout (host side, no GPU).

PanelMaterial.zurueck must be the transpose of the Jacobian of
PanelMaterial.werte: for any per-element gradient G, the directional
derivative of the linear functional <G, werte(theta)> along a direction d
equals <zurueck(G), d>. Checked by central differences of werte (a linear map,
so exact up to rounding) with random G and d on the small drape.
"""
import numpy as np
import pytest

from garmentcraft_amd import szenen
from garmentcraft_amd.adjungiert import PanelMaterial, verlust, verlust_gradient


@pytest.fixture(scope="module")
def material():
    cfg = szenen.DrapeKonfig(spalten=2, zeilen=2, zelle_mm=160.0, hoehe_m=0.335, jitter_mm=10.0, versatz_mm=(0.0, 0.0))
    s, _ = szenen.drape(cfg, mit_koerper=False)
    return s, PanelMaterial(s)


def _funktional(mat, th, G):
    md, bk, reib = mat.werte(th)
    return float(np.sum(G["membran"] * md) + np.sum(G["biegung"] * bk) + np.sum(G["reibung"] * reib))


def test_zurueck_ist_die_transponierte_von_werte(material):
    s, mat = material
    rng = np.random.default_rng(3)
    G = {"membran": rng.standard_normal((s.membran_punkte.shape[0], 3)),
         "biegung": rng.standard_normal(s.biegung_punkte.shape[0]),
         "reibung": rng.standard_normal(s.n)}
    th = mat.theta0() * rng.uniform(0.7, 1.3, mat.n_param)
    g = mat.zurueck(G)
    for _ in range(5):
        d = rng.standard_normal(mat.n_param)
        h = 1e-3
        fd = (_funktional(mat, th + h * d, G) - _funktional(mat, th - h * d, G)) / (2 * h)
        assert abs(fd - float(g @ d)) <= 1e-9 * max(1.0, abs(fd))


def test_werte_bei_theta0_sind_die_der_szene(material):
    s, mat = material
    md, bk, reib = mat.werte(mat.theta0())
    assert np.array_equal(md, s.membran_daten[:, 9:12])
    assert np.allclose(bk, 1.0 / s.biegung_nachgiebigkeit, rtol=0, atol=0)
    assert np.array_equal(reib, np.full(s.n, s.reibung[0]))
    assert mat.n_param == 3 * 4 + 1


def test_szene_traegt_die_werte(material):
    s, mat = material
    th = mat.theta0() * 1.1
    s2 = mat.szene(th)
    md, bk, reib = mat.werte(th)
    assert np.array_equal(s2.membran_daten[:, 9:12], md)
    assert np.allclose(1.0 / s2.biegung_nachgiebigkeit, bk, rtol=1e-15)
    assert np.array_equal(s2.reibung, reib)
    # the original scene is untouched
    assert np.array_equal(s.membran_daten[:, 9:12], mat.md0)


def test_verlust_gradient_ist_die_ableitung_des_verlusts():
    rng = np.random.default_rng(4)
    x, z = rng.standard_normal((50, 3)), rng.standard_normal((50, 3))
    g = verlust_gradient(x, z)
    d = rng.standard_normal((50, 3))
    h = 1e-6
    fd = (verlust(x + h * d, z) - verlust(x - h * d, z)) / (2 * h)
    assert abs(fd - float(np.sum(g * d))) <= 1e-8
