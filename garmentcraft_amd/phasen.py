"""Phase split of the persistent PCG's iteration on the 40,758-point drape.

Loads the diagnostic build libgcsim_phasen.so (`make -C garmentcraft_amd/csrc
phasen`, the library with -DGC_PHASEN), whose k_pcg stamps the device wall
clock at the phase boundaries of every iteration in every block:

  A       directions p = S(z + beta p_old) of the tile's points into LDS
  B       element and spring contributions into the tile's LDS slots
  C       row sums: A p, p and A p stored, the block's share of p . A p
  ar_pAp  grid all-reduce of p . A p (the wait for the slowest block included)
  update  dv, r, z and the block's share of r . z
  ar_rz   grid all-reduce of r . z (with the release's L2 write-back)

Block 0's split is summed in GridState; every block's split goes into a
per-block record (stages A-C, update and the two waits). Block 0's time
from its own partial to the summed result is read too.

    python -m garmentcraft_amd.phasen [--fenster 5:25 --fenster 240:260]

prints one JSON line per window of steps.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os

import numpy as np

from . import _lib

PHASEN_SO = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libgcsim_phasen.so")
NAMEN = ["A", "B", "C", "ar_pAp", "update", "ar_rz"]


def laden():
    """Bind the diagnostic build in place of libgcsim.so (before any context)."""
    if _lib._lib is not None and _lib.SO_PFAD != PHASEN_SO:
        raise _lib.GcFehler("libgcsim.so is already loaded in this process")
    if not os.path.exists(PHASEN_SO):
        raise _lib.GcFehler(f"{PHASEN_SO} is not built; run `make -C garmentcraft_amd/csrc phasen`")
    _lib.SO_PFAD = PHASEN_SO
    L = _lib.lib()
    L.gc_phasen_lesen.restype = C.c_int
    L.gc_phasen_lesen.argtypes = [C.c_void_p, C.c_void_p, C.c_int32]
    L.gc_phasen_bloecke.restype = C.c_int
    L.gc_phasen_bloecke.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32]
    L.gc_phasen_kacheln.restype = C.c_int
    L.gc_phasen_kacheln.argtypes = [C.c_void_p, C.c_void_p, C.c_int32]
    return L


def main(argv=None) -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--fenster", action="append", default=None,
                    help="steps a:b (after a steps, time b - a), repeatable; default 5:25 and 240:260")
    ap.add_argument("--praezision", default="fp32")
    a = ap.parse_args(argv)
    fenster = [tuple(int(x) for x in f.split(":")) for f in (a.fenster or ["5:25", "240:260"])]
    L = laden()
    from . import szenen
    from .geraet import GeraetSzene, ImplizitParameter
    s, _ = szenen.drape(sdf=szenen.koerper_sdf())
    g = GeraetSzene(s, ImplizitParameter(schritte_je_bild=1, cg_max=10000, cg_toleranz=1e-6, naht_k=500.0,
                                         selbstkollision=True), praezision=a.praezision)
    if L.gc_phasen_bloecke(g.h, None, 0, 0) != 0:
        raise _lib.GcFehler("gc_phasen_bloecke")
    us = np.zeros(7)
    bl = np.zeros(6 * 1024)
    fertig = 0
    for lo, hi in fenster:
        if lo > fertig:
            g.schritte(1.0 / 120.0, lo - fertig)
        L.gc_phasen_lesen(g.h, us.ctypes.data, 1)
        L.gc_phasen_bloecke(g.h, bl.ctypes.data, 1024, 1)
        g.statistik(reset=True)
        g.schritte(1.0 / 120.0, hi - lo)
        fertig = hi
        n = L.gc_phasen_lesen(g.h, us.ctypes.data, 1)
        L.gc_phasen_bloecke(g.h, bl.ctypes.data, 1024, 1)
        st = g.statistik()
        ms, pcg = g.zeiten()
        b6 = bl.reshape(1024, 6) / max(n, 1)
        aktiv = b6[:, 0] > 0
        b6 = b6[aktiv]
        # columns: A-C, update, waits (as before), then A, B, C alone
        b = np.column_stack([b6[:, 0] + b6[:, 1] + b6[:, 2], b6[:, 3], b6[:, 4], b6[:, 5], b6[:, 0], b6[:, 1], b6[:, 2]])
        pro_block = {}
        for k, name in enumerate(["A-C", "update", "wait_pAp", "wait_rz", "A", "B", "C"]):
            q = np.percentile(b[:, k], [0, 50, 100]) if len(b) else [0.0, 0.0, 0.0]
            pro_block[name] = {"min": round(float(q[0]), 3), "median": round(float(q[1]), 3), "max": round(float(q[2]), 3)}
        # the slowest one-tile blocks against their tiles' work (spring list
        # entries after the window's last rebuild, listed elements, rows)
        kach = np.zeros(4 * 1024, dtype=np.int32)
        nk = L.gc_phasen_kacheln(g.h, kach.ctypes.data, 1024)
        kach = kach.reshape(1024, 4)[:max(nk, 0)]
        arbeit = {}
        if nk > 0 and len(b) == nk:
            ac = b[:, 0]
            reihe = np.argsort(ac)[::-1]
            arbeit = {
                "slowest": [{"block": int(r), "A-C_us": round(float(ac[r]), 2), "A_B_C_us": [round(float(b[r, k]), 2) for k in (4, 5, 6)],
                             "springs": int(kach[r, 2]),
                             "elements": int(kach[r, 3]), "rows": int(kach[r, 1] - kach[r, 0])} for r in reihe[:8]],
                "median_springs": float(np.median(kach[:, 2])), "median_elements": float(np.median(kach[:, 3])),
                "corr_A-C_springs": round(float(np.corrcoef(ac, kach[:, 2])[0, 1]), 3) if kach[:, 2].std() > 0 else None,
                "corr_A-C_elements": round(float(np.corrcoef(ac, kach[:, 3])[0, 1]), 3),
            }
        print(json.dumps({
            "steps": f"{lo + 1}-{hi}", "iterations": n, "cg_iterations": st["cg_iterationen"],
            "tiles": arbeit,
            "block0_us": {k: round(float(us[i]), 3) for i, k in enumerate(NAMEN)},
            "block0_sum_us": round(float(us[:6].sum()), 3),
            "block0_partial_to_result_us": round(float(us[6]), 3),
            "us_per_iteration": round(pcg * 1e3 / max(st["cg_iterationen"], 1), 3),
            "blocks": int(aktiv.sum()), "per_block_us": pro_block,
            "proximity_pairs": st["naehe_punkt_dreieck"] + st["naehe_kante_kante"],
        }), flush=True)
    g.close()


if __name__ == "__main__":
    main()
