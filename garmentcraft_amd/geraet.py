"""A garment system resident on the MI355X: the Python face of one gc_ctx.

:class:`GeraetSzene` uploads a :class:`~garmentcraft_amd.szene.Szene` once
(topology, stiffnesses, rest lengths, bodies, state) and then steps it on the
device. The Loeser-level probes (build, solve, operator, residual, load) are
exposed for the parity checks against the reference's solver contract
(implizit.rs:304-351).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

from . import _lib as L
from .szene import GitterSdf, Netz, Szene, Zylinder


@dataclass
class ImplizitParameter:
    """implizit.rs:456-531 (the fields this path reads)."""
    schritte_je_bild: int = 5
    cg_max: int = 50
    cg_toleranz: float = 1e-5
    naht_k: float = 500.0
    kontakt_rand: float = 1e-4
    normalkraft_aus_last: bool = False
    selbstkollision: bool = False
    selbst_zusatzdicke: float = 2.5e-3
    selbst_steifigkeit: float = 1e-3
    selbst_stoss_iterationen: int = 1
    selbst_schnitte_an_naehten: bool = False
    durchstoss_pruefen: bool = True
    koerperpunkt_durchgaenge: int = 0
    koerperpunkt_tiefe: float = 0.01


@dataclass
class CgLauf:
    """implizit.rs:293-299."""
    iterationen: int
    nicht_konvergiert: bool
    rest: float


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def _p(a, t=L.D):
    return a.ctypes.data_as(t)


class GeraetSzene:
    """One scene on one device. ``praezision``: "fp64" or "fp32" (FP32 bulk,
    FP64 reductions)."""

    def __init__(self, s: Szene, param: Optional[ImplizitParameter] = None, praezision: str = "fp64",
                 bloecke: int = 0, threads: int = 0, geraet: int = 0, lds_kb: int = 0,
                 koerper_quelle: Optional["GeraetSzene"] = None):
        self.lib = L.lib()
        opt = L.Optionen(L.GC_FP32 if praezision == "fp32" else L.GC_FP64, geraet, bloecke, threads, lds_kb)
        err = C.create_string_buffer(512)
        h = self.lib.gc_create(C.byref(opt), err, 512)
        if not h:
            raise L.GcFehler("gc_create: " + err.value.decode())
        self.h = C.c_void_p(h)
        self.n = s.n
        self.praezision = praezision
        self.param = param or ImplizitParameter()
        self.bild_dt = s.param.bild_dt
        self._szene_param = s.param
        self.set_param(self.param)
        self._topologie(s)
        # grid bodies of koerper_quelle (same index) are shared, not uploaded again
        self._quelle = koerper_quelle
        for k, kb in enumerate(s.koerper):
            if koerper_quelle is not None and isinstance(kb, GitterSdf):
                i = self.lib.gc_add_sdf_von(self.h, koerper_quelle.h, k, _p(_f64(kb.versatz)), _p(_f64(kb.geschwindigkeit)))
                if i < 0:
                    raise L.GcFehler("gc_add_sdf_von: " + self.lib.gc_last_error(self.h).decode())
            else:
                self.add_koerper(kb)
        self.set_state(s.x, s.v)

    # -- plumbing -----------------------------------------------------------
    def _ck(self, rc: int, was: str) -> None:
        if rc != 0:
            raise L.GcFehler(f"{was}: {self.lib.gc_last_error(self.h).decode()}")

    def close(self) -> None:
        if getattr(self, "h", None) is not None:
            self.lib.gc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def name(self) -> str:
        return self.lib.gc_name(self.h).decode()

    # -- set-up -------------------------------------------------------------
    def _topologie(self, s: Szene) -> None:
        keep = []

        def arr(a, dt):
            a = np.ascontiguousarray(a, dtype=dt)
            keep.append(a)
            return a

        k = np.where(s.membran_kante, 0.0, 1.0 / s.abstand_nachgiebigkeit)  # implizit.rs:732-736
        bk = 1.0 / s.biegung_nachgiebigkeit                                   # implizit.rs:777
        t = L.Topologie()
        t.n = s.n
        t.masse = _p(arr(s.masse, np.float64)); t.w = _p(arr(s.w, np.float64))
        t.kollisionsabstand = _p(arr(s.kollisionsabstand, np.float64)); t.reibung = _p(arr(s.reibung, np.float64))
        t.n_kanten = s.abstand_punkte.shape[0]
        t.kanten_punkte = _p(arr(s.abstand_punkte, np.uint32), L.U32)
        t.kanten_k = _p(arr(k, np.float64)); t.kanten_ruhe = _p(arr(s.abstand_ruhe, np.float64))
        t.n_naehte = s.naht_punkte.shape[0]
        t.naht_punkte = _p(arr(s.naht_punkte, np.uint32), L.U32); t.naht_ruhe = _p(arr(s.naht_ruhe, np.float64))
        t.n_biegungen = s.biegung_punkte.shape[0]
        t.biegung_punkte = _p(arr(s.biegung_punkte, np.uint32), L.U32)
        t.biegung_k = _p(arr(bk, np.float64)); t.biegung_ruhe = _p(arr(s.biegung_ruhe, np.float64))
        t.n_membran = s.membran_punkte.shape[0]
        t.membran_punkte = _p(arr(s.membran_punkte, np.uint32), L.U32)
        t.membran_daten = _p(arr(s.membran_daten, np.float64))
        t.n_dreiecke = s.dreiecke.shape[0]
        t.dreiecke = _p(arr(s.dreiecke, np.uint32), L.U32)
        for name in ("selbst_e_punkt", "selbst_e_dreieck", "selbst_a_dreieck"):
            a = getattr(s, name, None)
            if a is not None:
                setattr(t, name, _p(arr(a, np.float64)))
        self._ck(self.lib.gc_set_topology(self.h, C.byref(t)), "gc_set_topology")

    def set_param(self, p: ImplizitParameter) -> None:
        self.param = p
        c = L.Param()
        for q in range(3):
            c.schwerkraft[q] = self._szene_param.schwerkraft[q]
        c.daempfung = self._szene_param.daempfung
        c.cg_max = p.cg_max
        c.cg_toleranz = p.cg_toleranz
        c.naht_k = p.naht_k
        c.kontakt_rand = p.kontakt_rand
        c.normalkraft_aus_last = 1 if p.normalkraft_aus_last else 0
        c.selbstkollision = 1 if p.selbstkollision else 0
        c.selbst_zusatzdicke = p.selbst_zusatzdicke
        c.selbst_steifigkeit = p.selbst_steifigkeit
        c.selbst_stoss_iterationen = p.selbst_stoss_iterationen
        c.selbst_schnitte_an_naehten = 1 if p.selbst_schnitte_an_naehten else 0
        c.durchstoss_pruefen = 1 if p.durchstoss_pruefen else 0
        c.koerperpunkt_durchgaenge = p.koerperpunkt_durchgaenge
        c.koerperpunkt_tiefe = p.koerperpunkt_tiefe
        self._ck(self.lib.gc_set_param(self.h, C.byref(c)), "gc_set_param")

    def add_koerper(self, kb) -> int:
        if isinstance(kb, Zylinder):
            i = self.lib.gc_add_zylinder(self.h, _p(_f64(kb.punkt)), _p(_f64(kb.achse)), kb.radius, _p(_f64(kb.geschwindigkeit)))
        elif isinstance(kb, GitterSdf):
            w = np.ascontiguousarray(kb.werte, dtype=np.float32)
            dims = np.array(kb.dims, dtype=np.int32)
            i = self.lib.gc_add_sdf(self.h, _p(w, L.F32), _p(dims, L.I32), _p(_f64(kb.origin)), kb.spacing,
                                    _p(_f64(kb.versatz)), _p(_f64(kb.geschwindigkeit)))
        elif isinstance(kb, Netz):
            pk = _f64(kb.punkte).reshape(-1, 3)
            dk = np.ascontiguousarray(kb.dreiecke, dtype=np.uint32).reshape(-1, 3)
            i = self.lib.gc_add_netz(self.h, _p(pk), pk.shape[0], _p(dk, L.U32), dk.shape[0], kb.suchtiefe,
                                     _p(_f64(kb.versatz)), _p(_f64(kb.geschwindigkeit)))
            if i >= 0 and kb.punktgeschwindigkeit is not None:
                self.netz_punktgeschwindigkeit(i, kb.punktgeschwindigkeit)
        else:
            raise TypeError(f"unknown body {type(kb)}")
        if i < 0:
            raise L.GcFehler("add body: " + self.lib.gc_last_error(self.h).decode())
        return i

    # -- mesh bodies (netzkoerper.rs) ---------------------------------------
    def netz_nachfuehren(self, k: int, punkte) -> None:
        self._ck(self.lib.gc_netz_nachfuehren(self.h, k, _p(_f64(punkte))), "gc_netz_nachfuehren")

    def netz_punktgeschwindigkeit(self, k: int, v) -> None:
        if v is None:
            self._ck(self.lib.gc_netz_punktgeschwindigkeit(self.h, k, None, 0), "gc_netz_punktgeschwindigkeit")
        else:
            a = _f64(v).reshape(-1, 3)
            self._ck(self.lib.gc_netz_punktgeschwindigkeit(self.h, k, _p(a), a.shape[0]), "gc_netz_punktgeschwindigkeit")

    def netz_info(self, k: int) -> dict:
        a = L.NetzAngaben_c()
        self._ck(self.lib.gc_netz_info(self.h, k, C.byref(a)), "gc_netz_info")
        return {"n_punkte": a.n_punkte, "n_dreiecke": a.n_dreiecke, "randkanten": a.randkanten, "volumen": a.volumen,
                "ursprung": np.array(a.ursprung[:]), "zelle": a.zelle, "zahl": np.array(a.zahl[:], dtype=np.int64),
                "n_eintraege": a.n_eintraege}

    def netz_lesen(self, k: int) -> dict:
        i = self.netz_info(k)
        n, nt = i["n_punkte"], i["n_dreiecke"]
        nz = int(np.prod(i["zahl"]))
        fn, pn = np.zeros((nt, 3)), np.zeros((n, 3))
        nb = np.zeros((nt, 3), dtype=np.uint32)
        kg = np.zeros((nt, 4))
        an = np.zeros(nz + 1, dtype=np.uint32)
        ei = np.zeros(max(1, i["n_eintraege"]), dtype=np.uint32)
        self._ck(self.lib.gc_netz_lesen(self.h, k, _p(fn), _p(pn), _p(nb, L.U32), _p(kg), _p(an, L.U32), _p(ei, L.U32)),
                 "gc_netz_lesen")
        return {"flaechennormale": fn, "punktnormale": pn, "nachbar": nb, "kugel": kg, "anfang": an,
                "eintraege": ei[:i["n_eintraege"]]}

    def netz_naechster(self, k: int, x, radius: float) -> dict:
        x = _f64(x).reshape(-1, 3)
        m = x.shape[0]
        da = np.zeros(m, dtype=np.int32)
        ab, pu, no = np.zeros(m), np.zeros((m, 3)), np.zeros((m, 3))
        dr, rg = np.zeros(m, dtype=np.int32), np.zeros(m, dtype=np.int32)
        self._ck(self.lib.gc_netz_naechster(self.h, k, m, _p(x), radius, _p(da, L.I32), _p(ab), _p(pu), _p(no),
                                            _p(dr, L.I32), _p(rg, L.I32)), "gc_netz_naechster")
        return {"da": da.astype(bool), "abstand": ab, "punkt": pu, "normale": no, "dreieck": dr, "region": rg}

    def netz_durchstoss(self, k: int, a, b, starr=(0.0, 0.0, 0.0)) -> dict:
        a = _f64(a).reshape(-1, 3)
        b = _f64(b).reshape(-1, 3)
        m = a.shape[0]
        da = np.zeros(m, dtype=np.int32)
        p, no, vk = np.zeros((m, 3)), np.zeros((m, 3)), np.zeros((m, 3))
        self._ck(self.lib.gc_netz_durchstoss(self.h, k, m, _p(a), _p(b), _p(_f64(starr)), _p(da, L.I32), _p(p), _p(no),
                                             _p(vk)), "gc_netz_durchstoss")
        return {"da": da.astype(bool), "p": p, "normale": no, "vk": vk}

    def netz_sdf_backen(self, k: int, dims, origin, spacing: float, band: float) -> GitterSdf:
        """The mesh body k's signed distance on a node grid of dims (nx, ny,
        nz), exact within ``band`` of the surface, as a grid-SDF body."""
        nx, ny, nz = (int(a) for a in dims)
        w = np.zeros((nz, ny, nx), dtype=np.float32)
        d = np.array([nx, ny, nz], dtype=np.int32)
        self._ck(self.lib.gc_netz_sdf_backen(self.h, k, _p(d, L.I32), _p(_f64(origin)), spacing, band, _p(w, L.F32)),
                 "gc_netz_sdf_backen")
        return GitterSdf(werte=w, origin=tuple(float(a) for a in origin), spacing=spacing)

    # -- the XPBD path (cpu.rs:76-246) -------------------------------------
    def xpbd_topologie(self, s: Szene) -> None:
        """Hand the scene's XPBD constraint lists (colour-ordered, with their
        compliances) to the device."""
        keep = []

        def arr(a, dt):
            a = np.ascontiguousarray(a, dtype=dt)
            keep.append(a)
            return a

        t = L.XpbdTopologie()
        t.n_abstand = s.abstand_punkte.shape[0]
        t.abstand_punkte = _p(arr(s.abstand_punkte, np.uint32), L.U32)
        t.abstand_ruhe = _p(arr(s.abstand_ruhe, np.float64))
        t.abstand_nachgiebigkeit = _p(arr(s.abstand_nachgiebigkeit, np.float64))
        t.membran_kante = _p(arr(s.membran_kante, np.uint8), C.POINTER(C.c_uint8))
        t.n_naehte = s.naht_punkte.shape[0]
        t.naht_punkte = _p(arr(s.naht_punkte, np.uint32), L.U32)
        t.naht_ruhe = _p(arr(s.naht_ruhe, np.float64))
        t.naht_nachgiebigkeit = s.param.naht_nachgiebigkeit
        t.n_biegungen = s.biegung_punkte.shape[0]
        t.biegung_punkte = _p(arr(s.biegung_punkte, np.uint32), L.U32)
        t.biegung_ruhe = _p(arr(s.biegung_ruhe, np.float64))
        t.biegung_nachgiebigkeit = _p(arr(s.biegung_nachgiebigkeit, np.float64))
        self._ck(self.lib.gc_xpbd_topology(self.h, C.byref(t)), "gc_xpbd_topology")

    def xpbd_bild(self, unterschritte: Optional[int] = None, bild_dt: Optional[float] = None) -> dict:
        """One XPBD frame (CpuReferenz::bild); returns its contact statistics."""
        st = L.XpbdStatistik_c()
        n = self._szene_param.unterschritte if unterschritte is None else unterschritte
        dt = self._szene_param.bild_dt if bild_dt is None else bild_dt
        self._ck(self.lib.gc_xpbd_bild(self.h, dt, int(n), C.byref(st)), "gc_xpbd_bild")
        return {"kontakte": st.kontakte, "haftend": st.haftend, "tiefe": st.tiefe, "tangential": st.tangential}

    def xpbd_laeufe(self) -> dict:
        a, n, b, m = (C.c_int32() for _ in range(4))
        self._ck(self.lib.gc_xpbd_laeufe(self.h, C.byref(a), C.byref(n), C.byref(b), C.byref(m)), "gc_xpbd_laeufe")
        return {"abstand": a.value, "naehte": n.value, "biegungen": b.value, "membran_stufen": m.value}

    def set_state(self, x=None, v=None) -> None:
        self._ck(self.lib.gc_set_state(self.h, _p(_f64(x)) if x is not None else None,
                                       _p(_f64(v)) if v is not None else None), "gc_set_state")

    def lage(self):
        x = np.zeros((self.n, 3)); v = np.zeros((self.n, 3))
        self._ck(self.lib.gc_get_state(self.h, _p(x), _p(v)), "gc_get_state")
        return x, v

    def set_naht_ruhe(self, r) -> None:
        self._ck(self.lib.gc_set_naht_ruhe(self.h, _p(_f64(r))), "gc_set_naht_ruhe")

    def set_w(self, w) -> None:
        self._ck(self.lib.gc_set_w(self.h, _p(_f64(w))), "gc_set_w")

    def set_dv_vorher(self, dv) -> None:
        if dv is None:
            self._ck(self.lib.gc_set_dv_vorher(self.h, None, 0), "gc_set_dv_vorher")
        else:
            d = _f64(dv)
            self._ck(self.lib.gc_set_dv_vorher(self.h, _p(d), d.shape[0]), "gc_set_dv_vorher")

    def set_koerper_lage(self, k: int, lage) -> None:
        self._ck(self.lib.gc_set_koerper_lage(self.h, k, _p(_f64(lage))), "gc_set_koerper_lage")

    # -- stepping -----------------------------------------------------------
    def schritte(self, h: float, n: int = 1) -> None:
        self._ck(self.lib.gc_step(self.h, h, n), "gc_step")

    def schritte_graph(self, h: float, n: int) -> None:
        self._ck(self.lib.gc_step_graph(self.h, h, n), "gc_step_graph")

    def bild(self, bild_dt: Optional[float] = None, schritte_je_bild: Optional[int] = None) -> None:
        dt = self.bild_dt if bild_dt is None else bild_dt
        n = self.param.schritte_je_bild if schritte_je_bild is None else schritte_je_bild
        self._ck(self.lib.gc_bild(self.h, dt, n), "gc_bild")

    def statistik(self, reset: bool = False) -> dict:
        st = L.Statistik_c()
        self._ck(self.lib.gc_read_statistik(self.h, C.byref(st), 1 if reset else 0), "gc_read_statistik")
        return {f: getattr(st, f) for f, _ in L.Statistik_c._fields_}

    def zeiten(self):
        a = C.c_double(); b = C.c_double()
        self._ck(self.lib.gc_zeiten(self.h, C.byref(a), C.byref(b)), "gc_zeiten")
        return a.value, b.value

    # -- Loeser-level probes ------------------------------------------------
    def bauen(self, h: float) -> None:
        self._ck(self.lib.gc_build(self.h, h), "gc_build")

    def fest(self):
        n = self.n
        fl = np.zeros(n, np.int32); nr = np.zeros((n, 3)); vn = np.zeros(n); vk = np.zeros((n, 3))
        self._ck(self.lib.gc_get_fest(self.h, _p(fl, L.I32), _p(nr), _p(vn), _p(vk)), "gc_get_fest")
        return {"flag": fl, "normale": nr, "vn_soll": vn, "v_koerper": vk}

    def setze_fest(self, flag, normale, vn_soll, v_koerper) -> None:
        self._ck(self.lib.gc_set_fest(self.h, _p(np.ascontiguousarray(flag, np.int32), L.I32), _p(_f64(normale)),
                                      _p(_f64(vn_soll)), _p(_f64(v_koerper))), "gc_set_fest")

    def loese(self, dv):
        d = _f64(dv).copy()
        lauf = L.CgLauf_c()
        self._ck(self.lib.gc_solve(self.h, _p(d), C.byref(lauf)), "gc_solve")
        return d, CgLauf(lauf.iterationen, bool(lauf.nicht_konvergiert), lauf.rest)

    def mal_a(self, p):
        o = np.zeros((self.n, 3))
        self._ck(self.lib.gc_apply_A(self.h, _p(_f64(p)), _p(o)), "gc_apply_A")
        return o

    def mal_h(self, p):
        o = np.zeros((self.n, 3))
        self._ck(self.lib.gc_apply_H(self.h, _p(_f64(p)), _p(o)), "gc_apply_H")
        return o

    def rest(self, dv):
        o = np.zeros((self.n, 3))
        self._ck(self.lib.gc_residual(self.h, _p(_f64(dv)), _p(o)), "gc_residual")
        return o

    def last(self):
        o = np.zeros((self.n, 3))
        self._ck(self.lib.gc_read_load(self.h, _p(o)), "gc_read_load")
        return o

    def system(self):
        n = self.n
        rhs, f, last, pinv = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3, 3))
        self._ck(self.lib.gc_read_system(self.h, _p(rhs), _p(f), _p(last), _p(pinv)), "gc_read_system")
        return {"rhs": rhs, "f": f, "last": last, "pinv": pinv}

    def records(self, n_kanten_geraet: int, n_biegungen: int):
        r = np.zeros((n_kanten_geraet, 3)); q = np.zeros(n_kanten_geraet); g = np.zeros((n_biegungen, 4, 3))
        self._ck(self.lib.gc_read_records(self.h, _p(r), _p(q), _p(g)), "gc_read_records")
        return {"richtung": r, "quer": q, "grad": g}

    # -- self-collision lists of the last build / step ---------------------
    def naehe(self) -> dict:
        m = self.lib.gc_read_naehe(self.h, 0, None, None, None, None, None, None)
        self._ck(0 if m >= 0 else 1, "gc_read_naehe")
        art = np.zeros(m, np.int32); p = np.zeros((m, 4), np.uint32); g = np.zeros((m, 4)); nr = np.zeros((m, 3))
        ab = np.zeros(m); ru = np.zeros(m)
        self.lib.gc_read_naehe(self.h, m, _p(art, L.I32), _p(p, L.U32), _p(g), _p(nr), _p(ab), _p(ru))
        return {"art": art, "punkte": p, "gewichte": g, "normale": nr, "abstand": ab, "ruhe": ru}

    def federn(self) -> dict:
        m = self.lib.gc_read_federn(self.h, 0, None, None, None, None)
        self._ck(0 if m >= 0 else 1, "gc_read_federn")
        p = np.zeros((m, 5), np.uint32); g = np.zeros((m, 5)); k = np.zeros((m, 3)); b = np.zeros((m, 3, 3))
        self.lib.gc_read_federn(self.h, m, _p(p, L.U32), _p(g), _p(k), _p(b))
        return {"punkte": p, "gewichte": g, "kraft": k, "block": b}

    def schnitte(self) -> dict:
        m = self.lib.gc_read_schnitte(self.h, 0, None, None, None, None, None, None)
        self._ck(0 if m >= 0 else 1, "gc_read_schnitte")
        kd = np.zeros((m, 2), np.uint32); t = np.zeros(m); w = np.zeros((m, 3)); r = np.zeros((m, 3)); g = np.zeros(m)
        wz = np.zeros(m, np.uint32)
        self.lib.gc_read_schnitte(self.h, m, _p(kd, L.U32), _p(t), _p(w), _p(r), _p(g), _p(wz, L.U32))
        return {"kante_dreieck": kd, "t": t, "w": w, "richtung": r, "gewicht": g, "wurzel": wz}

    def stoesse(self) -> dict:
        m = self.lib.gc_read_stoesse(self.h, 0, None, None, None, None, None, None, None)
        self._ck(0 if m >= 0 else 1, "gc_read_stoesse")
        art = np.zeros(m, np.int32); p = np.zeros((m, 4), np.uint32); lage = np.zeros((m, 2)); zeit = np.zeros(m)
        seite = np.zeros(m); nr = np.zeros((m, 3)); gilt = np.zeros(m, np.int32)
        self.lib.gc_read_stoesse(self.h, m, _p(art, L.I32), _p(p, L.U32), _p(lage), _p(zeit), _p(seite), _p(nr),
                                 _p(gilt, L.I32))
        return {"art": art, "punkte": p, "lage": lage, "zeit": zeit, "seite": seite, "normale": nr, "gilt": gilt}

    def v(self):
        o = np.zeros((self.n, 3))
        self._ck(self.lib.gc_get_v(self.h, _p(o)), "gc_get_v")
        return o

    # -- material values, saved states, adjoint ----------------------------
    def set_steifigkeit(self, kanten_k=None, biegung_k=None, membran_ek_es_g=None, reibung=None) -> None:
        """New stiffnesses / friction on the uploaded topology (None keeps a group)."""
        keep = [None if a is None else _f64(a) for a in (kanten_k, biegung_k, membran_ek_es_g, reibung)]
        ptr = [None if a is None else _p(a) for a in keep]
        self._ck(self.lib.gc_set_steifigkeit(self.h, *ptr), "gc_set_steifigkeit")

    def sichern(self, slot: int) -> None:
        self._ck(self.lib.gc_zustand_sichern(self.h, slot), "gc_zustand_sichern")

    def entscheidungen(self, modus: int, slot: int = 0) -> None:
        """Record (1), replay (2) or stop holding (0) the steps' discrete decisions."""
        self._ck(self.lib.gc_entscheidungen(self.h, int(modus), int(slot)), "gc_entscheidungen")

    def entscheidungen_abweichungen(self, reset: bool = False) -> int:
        n = C.c_int64()
        self._ck(self.lib.gc_entscheidungen_abweichungen(self.h, C.byref(n), 1 if reset else 0),
                 "gc_entscheidungen_abweichungen")
        return int(n.value)

    def laden(self, slot: int) -> None:
        self._ck(self.lib.gc_zustand_laden(self.h, slot), "gc_zustand_laden")

    def adjoint_setzen(self, gx=None, gv=None) -> None:
        a = None if gx is None else _f64(gx)
        b = None if gv is None else _f64(gv)
        self._ck(self.lib.gc_adjoint_set(self.h, None if a is None else _p(a), None if b is None else _p(b)),
                 "gc_adjoint_set")

    def adjoint_lesen(self):
        gx = np.zeros((self.n, 3)); gv = np.zeros((self.n, 3))
        self._ck(self.lib.gc_adjoint_get(self.h, _p(gx), _p(gv)), "gc_adjoint_get")
        return gx, gv

    def adjoint_schritt(self, h: float) -> None:
        self._ck(self.lib.gc_adjoint_step(self.h, h), "gc_adjoint_step")

    def schritt_aufnehmen(self, h: float, slot: int) -> None:
        """One step after saving the state into slot, keeping its solutions (gc_step_aufnehmen)."""
        self._ck(self.lib.gc_step_aufnehmen(self.h, h, int(slot)), "gc_step_aufnehmen")

    def adjoint_schritt_aus(self, h: float, slot: int) -> None:
        """The adjoint step from slot's state with its recorded solutions (gc_adjoint_step_aus)."""
        self._ck(self.lib.gc_adjoint_step_aus(self.h, h, int(slot)), "gc_adjoint_step_aus")

    def adjoint_gradient(self, s: Szene, reset: bool = False) -> dict:
        """Accumulated parameter gradients: per input edge, seam, bending
        constraint, membrane (E_k, E_s, G) and point (friction)."""
        dk = np.zeros(s.abstand_punkte.shape[0]); dn = np.zeros(s.naht_punkte.shape[0])
        db = np.zeros(s.biegung_punkte.shape[0]); dm = np.zeros((s.membran_punkte.shape[0], 3)); dr = np.zeros(self.n)
        self._ck(self.lib.gc_adjoint_gradient(self.h, _p(dk), _p(dn), _p(db), _p(dm), _p(dr), 1 if reset else 0),
                 "gc_adjoint_gradient")
        return {"kanten": dk, "naehte": dn, "biegung": db, "membran": dm, "reibung": dr}

    def adjoint_statistik(self, reset: bool = False) -> dict:
        st = L.AdjointStat_c()
        self._ck(self.lib.gc_adjoint_statistik(self.h, C.byref(st), 1 if reset else 0), "gc_adjoint_statistik")
        return {f: getattr(st, f) for f, _ in L.AdjointStat_c._fields_}


class GeraetStapel:
    """Many device systems stepped as one (gc_batch_*): every substep builds all
    of them, solves them in one persistent PCG launch per solve pass (each
    system with its own group of blocks, reductions and stopping test) and
    updates all of them. The systems stay usable for reading state and
    statistics; they must not be stepped on their own while the batch runs."""

    def __init__(self, systeme: Sequence[GeraetSzene], threads: int = 0, bloecke: int = 0):
        self.lib = L.lib()
        self.systeme = list(systeme)
        arr = (C.c_void_p * len(self.systeme))(*[g.h.value for g in self.systeme])
        err = C.create_string_buffer(512)
        h = self.lib.gc_batch_create(arr, len(self.systeme), threads, bloecke, err, 512)
        if not h:
            raise L.GcFehler("gc_batch_create: " + err.value.decode())
        self.h = C.c_void_p(h)

    def _ck(self, rc: int, was: str) -> None:
        if rc != 0:
            raise L.GcFehler(f"{was}: {self.lib.gc_batch_last_error(self.h).decode()}")

    def close(self) -> None:
        if getattr(self, "h", None) is not None:
            self.lib.gc_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def schritte(self, h: float, n: int = 1) -> None:
        self._ck(self.lib.gc_batch_step(self.h, h, n), "gc_batch_step")

    def schritte_graph(self, h: float, n: int) -> None:
        self._ck(self.lib.gc_batch_step_graph(self.h, h, n), "gc_batch_step_graph")

    def autotune(self, h: float) -> dict:
        t = C.c_int32(); b = C.c_int32(); ms = C.c_double()
        self._ck(self.lib.gc_batch_autotune(self.h, h, C.byref(t), C.byref(b), C.byref(ms)), "gc_batch_autotune")
        return {"threads": t.value, "bloecke": b.value, "ms_loesung": ms.value}

    def bloecke(self):
        nb = np.zeros(len(self.systeme), dtype=np.int32)
        t = C.c_int32()
        self._ck(self.lib.gc_batch_bloecke(self.h, _p(nb, L.I32), C.byref(t)), "gc_batch_bloecke")
        return nb, t.value

    def zeiten(self):
        a = C.c_double(); b = C.c_double()
        self._ck(self.lib.gc_batch_zeiten(self.h, C.byref(a), C.byref(b)), "gc_batch_zeiten")
        return a.value, b.value
