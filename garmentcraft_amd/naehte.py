"""Seam pairing by length fraction through the C ABI (gc_naht_planen).

The boundary plan of klon3d-mesh (rand.rs: Randplan, naht_planen, unterteile)
and the stitched point pairs it yields (lib.rs:351-381): pieces are given as
boundary edges (polylines in mm, node to node), seams as two sides, each a
list of runs of whole edges of one piece (klon3d_pattern::naehte::Teilstueck:
piece, edges, backwards). `planen` returns every piece's boundary points (the
first points of its mesh, rand.rs:547-583) and every seam's stations and
sub-seams with their point pairs.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L

ENDE, GEMEINSAM, NUR_A, NUR_B = 0, 1, 2, 3
ARTEN = {ENDE: "Ende", GEMEINSAM: "Gemeinsam", NUR_A: "NurA", NUR_B: "NurB"}


@dataclass
class Lauf:
    """A run of a seam side: edges of one piece in seam order (Teilstueck)."""
    teil: int
    kanten: list
    rueckwaerts: bool = False


@dataclass
class Naht:
    a: list  # [Lauf]
    b: list


@dataclass
class Parameter:
    """Parameter of klon3d-mesh (lib.rs:52-86): the defaults of the reference."""
    kante_mm: float = 10.0
    sehne_mm: float = 0.1
    laenge_mm: float = 0.1
    fang_mm: float = 2.0
    min_rand_mm: float = 0.5


@dataclass
class Stelle:
    art: int
    knoten_a: tuple  # (piece, edge whose start node it is) or None
    knoten_b: tuple
    rho_a: float
    rho_b: float
    versatz_mm: float


@dataclass
class Teilnaht:
    teil_a: int
    teil_b: int
    paare: np.ndarray  # (k, 2) int32: boundary point on A, on B
    rho: np.ndarray    # (k, 2) float64


@dataclass
class Nahtnetz:
    laenge_a_mm: float
    laenge_b_mm: float
    stellen: list = field(default_factory=list)
    teilnaehte: list = field(default_factory=list)


@dataclass
class Randpunkte:
    kante: np.ndarray  # int32 per boundary point
    s: np.ndarray      # arc length from the edge's start (mm)
    xy: np.ndarray     # (n, 2) mm
    knoten: np.ndarray  # per edge: the boundary point of its start node


@dataclass
class Plan:
    teile: list   # [Randpunkte]
    naehte: list  # [Nahtnetz]


def _i32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))


def _f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def planen(teile, naehte, p: Parameter | None = None) -> Plan:
    """teile: per piece a list of boundary edges, each an (m, 2) polyline in mm
    (edge k runs from node k to node k + 1; the last edge closes the outline).
    naehte: [Naht]. Raises GcFehler with the reason when the pattern cannot be
    planned (as the reference's Err)."""
    p = p or Parameter()
    teil_kanten, kante_punkte, punkte = [0], [0], []
    for t in teile:
        for k in t:
            k = _f64(k).reshape(-1, 2)
            punkte.append(k)
            kante_punkte.append(kante_punkte[-1] + k.shape[0])
        teil_kanten.append(teil_kanten[-1] + len(t))
    punkte = _f64(np.concatenate(punkte) if punkte else np.zeros((0, 2)))
    seite_stuecke, st_teil, st_rueck, st_kanten, kanten = [0], [], [], [0], []
    for n in naehte:
        for seite in (n.a, n.b):
            for lauf in seite:
                st_teil.append(lauf.teil)
                st_rueck.append(1 if lauf.rueckwaerts else 0)
                kanten.extend(lauf.kanten)
                st_kanten.append(len(kanten))
            seite_stuecke.append(len(st_teil))
    arr = dict(teil_kanten=_i32(teil_kanten), kante_punkte=_i32(kante_punkte), seite_stuecke=_i32(seite_stuecke),
               stueck_teil=_i32(st_teil + [0]), stueck_rueck=_i32(st_rueck + [0]), stueck_kanten=_i32(st_kanten),
               kanten=_i32(kanten + [0]))
    ein = L.NahtEingabe_c(
        n_teile=len(teile), teil_kanten=arr["teil_kanten"].ctypes.data_as(L.I32),
        kante_punkte=arr["kante_punkte"].ctypes.data_as(L.I32), punkte=punkte.ctypes.data_as(L.D),
        n_naehte=len(naehte), seite_stuecke=arr["seite_stuecke"].ctypes.data_as(L.I32),
        stueck_teil=arr["stueck_teil"].ctypes.data_as(L.I32), stueck_rueck=arr["stueck_rueck"].ctypes.data_as(L.I32),
        stueck_kanten=arr["stueck_kanten"].ctypes.data_as(L.I32), kanten=arr["kanten"].ctypes.data_as(L.I32),
        kante_mm=p.kante_mm, sehne_mm=p.sehne_mm, laenge_mm=p.laenge_mm, fang_mm=p.fang_mm, min_rand_mm=p.min_rand_mm)
    lib = L.lib()
    n_kanten = teil_kanten[-1]
    kap = dict(rand=0, stellen=0, teilnaehte=0, paare=0)
    for _ in range(2):
        b = dict(
            rand_teil=np.zeros(len(teile) + 1, np.int32), rand_kante=np.zeros(max(kap["rand"], 1), np.int32),
            rand_s=np.zeros(max(kap["rand"], 1)), rand_xy=np.zeros(2 * max(kap["rand"], 1)),
            rand_knoten=np.zeros(max(n_kanten, 1), np.int32),
            naht_stellen=np.zeros(len(naehte) + 1, np.int32), stelle_art=np.zeros(max(kap["stellen"], 1), np.int32),
            stelle_knoten=np.zeros(4 * max(kap["stellen"], 1), np.int32),
            stelle_rho=np.zeros(2 * max(kap["stellen"], 1)), stelle_versatz_mm=np.zeros(max(kap["stellen"], 1)),
            naht_teilnaehte=np.zeros(len(naehte) + 1, np.int32),
            teilnaht_teile=np.zeros(2 * max(kap["teilnaehte"], 1), np.int32),
            teilnaht_paare=np.zeros(max(kap["teilnaehte"], 1) + 1, np.int32),
            paar_punkte=np.zeros(2 * max(kap["paare"], 1), np.int32), paar_rho=np.zeros(2 * max(kap["paare"], 1)),
            naht_laengen_mm=np.zeros(2 * max(len(naehte), 1)))
        aus = L.NahtAusgabe_c(kap_rand=kap["rand"], kap_stellen=kap["stellen"], kap_teilnaehte=kap["teilnaehte"],
                              kap_paare=kap["paare"])
        for name, a in b.items():
            setattr(aus, name, a.ctypes.data_as(L.I32 if a.dtype == np.int32 else L.D))
        rc = lib.gc_naht_planen(C.byref(ein), C.byref(aus))
        if rc == 1:
            raise L.GcFehler("gc_naht_planen: " + aus.fehler.decode())
        if rc not in (0, 2):
            raise L.GcFehler(f"gc_naht_planen: bad arguments ({rc})")
        if rc == 0:
            break
        kap = dict(rand=aus.n_rand, stellen=aus.n_stellen, teilnaehte=aus.n_teilnaehte, paare=aus.n_paare)
    else:
        raise L.GcFehler("gc_naht_planen: capacities did not settle")
    plan = Plan(teile=[], naehte=[])
    k0 = 0
    for t in range(len(teile)):
        a, e = b["rand_teil"][t], b["rand_teil"][t + 1]
        nk = teil_kanten[t + 1] - teil_kanten[t]
        plan.teile.append(Randpunkte(kante=b["rand_kante"][a:e].copy(), s=b["rand_s"][a:e].copy(),
                                     xy=b["rand_xy"][2 * a:2 * e].reshape(-1, 2).copy(),
                                     knoten=b["rand_knoten"][k0:k0 + nk].copy()))
        k0 += nk
    for n in range(len(naehte)):
        nn = Nahtnetz(laenge_a_mm=float(b["naht_laengen_mm"][2 * n]), laenge_b_mm=float(b["naht_laengen_mm"][2 * n + 1]))
        for i in range(b["naht_stellen"][n], b["naht_stellen"][n + 1]):
            kn = b["stelle_knoten"][4 * i:4 * i + 4]
            nn.stellen.append(Stelle(art=int(b["stelle_art"][i]),
                                     knoten_a=None if kn[0] < 0 else (int(kn[0]), int(kn[1])),
                                     knoten_b=None if kn[2] < 0 else (int(kn[2]), int(kn[3])),
                                     rho_a=float(b["stelle_rho"][2 * i]), rho_b=float(b["stelle_rho"][2 * i + 1]),
                                     versatz_mm=float(b["stelle_versatz_mm"][i])))
        for q in range(b["naht_teilnaehte"][n], b["naht_teilnaehte"][n + 1]):
            a, e = b["teilnaht_paare"][q], b["teilnaht_paare"][q + 1]
            nn.teilnaehte.append(Teilnaht(teil_a=int(b["teilnaht_teile"][2 * q]), teil_b=int(b["teilnaht_teile"][2 * q + 1]),
                                          paare=b["paar_punkte"][2 * a:2 * e].reshape(-1, 2).copy(),
                                          rho=b["paar_rho"][2 * a:2 * e].reshape(-1, 2).copy()))
        plan.naehte.append(nn)
    return plan
