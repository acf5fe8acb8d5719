"""gc_solve_ohne_speicher (no allocation; its PCG's grid state and reduction
slots in a static device array) against gc_solve on the same built system
and warm start: equal to the bit, in both precisions, on two calls in a row,
the second on a smaller system."""
import ctypes as C

import numpy as np
import pytest

from garmentcraft_amd import _lib as L
from garmentcraft_amd import szenen
from garmentcraft_amd.geraet import GeraetSzene, ImplizitParameter

pytestmark = pytest.mark.gpu

H = 1.0 / 120.0


def ohne_speicher(g, dv0):
    d = np.ascontiguousarray(dv0, dtype=np.float64).copy()
    lauf = L.CgLauf_c()
    rc = g.lib.gc_solve_ohne_speicher(g.h, d.ctypes.data_as(C.POINTER(C.c_double)), C.byref(lauf))
    assert rc == 0, g.lib.gc_last_error(g.h).decode()
    return d, lauf


@pytest.mark.parametrize("praezision", ["fp64", "fp32"])
def test_ohne_speicher_wie_gc_solve(praezision):
    sdf = szenen.koerper_sdf(aufloesung=96, spacing=0.008)
    gross, _ = szenen.drape(szenen.DrapeKonfig(spalten=2, zeilen=2, zelle_mm=160.0, hoehe_m=0.335, jitter_mm=10.0,
                                               versatz_mm=(0.0, 0.0)), sdf=sdf)
    klein, _ = szenen.drape(szenen.DrapeKonfig(seed=3, spalten=1, zeilen=1, zelle_mm=150.0, hoehe_m=0.33,
                                               jitter_mm=10.0, versatz_mm=(0.0, 0.0)), sdf=sdf)
    assert klein.n < gross.n
    p = ImplizitParameter(schritte_je_bild=1, cg_max=10000, cg_toleranz=1e-6)
    for s in (gross, klein):  # the second call on the smaller system
        g = GeraetSzene(s, p, praezision=praezision)
        g.schritte(H, 14)  # in contact with the body by now
        g.bauen(H)
        dv0 = np.tile(np.array([0.0, 0.0, -9.81 * H]), (s.n, 1))
        a, la = g.loese(dv0)
        b, lb = ohne_speicher(g, dv0)
        print(praezision, s.n, "points:", la, "iterations", lb.iterationen, "held", int((g.fest()["flag"] == 1).sum()))
        assert np.array_equal(a, b)
        assert la.iterationen == lb.iterationen and la.rest == lb.rest
        assert not lb.nicht_konvergiert and lb.iterationen > 0
        g.close()
