"""ctypes binding of libgcsim.so (the C ABI declared in include/gcsim.h).

Loading fails loudly: there is no CPU fallback behind this module.
"""
from __future__ import annotations

import ctypes as C
import os

_HIER = os.path.dirname(os.path.abspath(__file__))
SO_PFAD = os.path.join(_HIER, "libgcsim.so")

GC_FP64 = 0
GC_FP32 = 1

D = C.POINTER(C.c_double)
U32 = C.POINTER(C.c_uint32)
I32 = C.POINTER(C.c_int32)
F32 = C.POINTER(C.c_float)


class GcFehler(RuntimeError):
    """An entry point returned a nonzero status."""


class Optionen(C.Structure):
    _fields_ = [("praezision", C.c_int32), ("geraet", C.c_int32), ("bloecke", C.c_int32), ("threads", C.c_int32),
                ("lds_kb", C.c_int32)]


class Topologie(C.Structure):
    _fields_ = [
        ("n", C.c_int32), ("masse", D), ("w", D), ("kollisionsabstand", D), ("reibung", D),
        ("n_kanten", C.c_int32), ("kanten_punkte", U32), ("kanten_k", D), ("kanten_ruhe", D),
        ("n_naehte", C.c_int32), ("naht_punkte", U32), ("naht_ruhe", D),
        ("n_biegungen", C.c_int32), ("biegung_punkte", U32), ("biegung_k", D), ("biegung_ruhe", D),
        ("n_membran", C.c_int32), ("membran_punkte", U32), ("membran_daten", D),
        ("n_dreiecke", C.c_int32), ("dreiecke", U32), ("selbst_e_punkt", D), ("selbst_e_dreieck", D),
        ("selbst_a_dreieck", D),
    ]


class Param(C.Structure):
    _fields_ = [
        ("schwerkraft", C.c_double * 3), ("daempfung", C.c_double), ("cg_max", C.c_int32),
        ("cg_toleranz", C.c_double), ("naht_k", C.c_double), ("kontakt_rand", C.c_double),
        ("normalkraft_aus_last", C.c_int32),
        ("selbstkollision", C.c_int32), ("selbst_zusatzdicke", C.c_double), ("selbst_steifigkeit", C.c_double),
        ("selbst_stoss_iterationen", C.c_int32), ("selbst_schnitte_an_naehten", C.c_int32),
        ("durchstoss_pruefen", C.c_int32), ("koerperpunkt_durchgaenge", C.c_int32), ("koerperpunkt_tiefe", C.c_double),
    ]


class NahtEingabe_c(C.Structure):
    _fields_ = [
        ("n_teile", C.c_int32), ("teil_kanten", I32), ("kante_punkte", I32), ("punkte", D),
        ("n_naehte", C.c_int32), ("seite_stuecke", I32), ("stueck_teil", I32), ("stueck_rueck", I32),
        ("stueck_kanten", I32), ("kanten", I32),
        ("kante_mm", C.c_double), ("sehne_mm", C.c_double), ("laenge_mm", C.c_double), ("fang_mm", C.c_double),
        ("min_rand_mm", C.c_double),
    ]


class NahtAusgabe_c(C.Structure):
    _fields_ = [
        ("kap_rand", C.c_int64), ("n_rand", C.c_int64), ("rand_teil", I32), ("rand_kante", I32), ("rand_s", D),
        ("rand_xy", D), ("rand_knoten", I32),
        ("kap_stellen", C.c_int64), ("n_stellen", C.c_int64), ("naht_stellen", I32), ("stelle_art", I32),
        ("stelle_knoten", I32), ("stelle_rho", D), ("stelle_versatz_mm", D),
        ("kap_teilnaehte", C.c_int64), ("n_teilnaehte", C.c_int64), ("naht_teilnaehte", I32),
        ("teilnaht_teile", I32), ("teilnaht_paare", I32),
        ("kap_paare", C.c_int64), ("n_paare", C.c_int64), ("paar_punkte", I32), ("paar_rho", D),
        ("naht_laengen_mm", D), ("fehler", C.c_char * 256),
    ]


class CgLauf_c(C.Structure):
    _fields_ = [("iterationen", C.c_int32), ("nicht_konvergiert", C.c_int32), ("rest", C.c_double)]


class Statistik_c(C.Structure):
    _fields_ = [
        ("cg_iterationen", C.c_int64), ("loesungen", C.c_int64), ("nicht_konvergiert", C.c_int64),
        ("kontakte", C.c_int64), ("haftend", C.c_int64), ("freigegeben", C.c_int64), ("rest_max", C.c_double),
        ("barriere_spins_max", C.c_int64),
        ("naehe_punkt_dreieck", C.c_int64), ("naehe_kante_kante", C.c_int64), ("selbst_stoesse", C.c_int64),
        ("durchdringungen", C.c_int64), ("durchdringungsgebiete", C.c_int64), ("durchstoesse", C.c_int64),
        ("cg_rundungsgrenze", C.c_int64), ("cg_neustarts", C.c_int64), ("selbst_listen_faktor", C.c_int64),
        ("koerperpunkte", C.c_int64),
    ]


class AdjointStat_c(C.Structure):
    _fields_ = [("schritte", C.c_int64), ("cg_iterationen", C.c_int64), ("nicht_konvergiert", C.c_int64),
                ("abweichungen", C.c_int64), ("eintritte", C.c_int64), ("rest_max", C.c_double)]


class NetzAngaben_c(C.Structure):
    _fields_ = [("n_punkte", C.c_int32), ("n_dreiecke", C.c_int32), ("randkanten", C.c_int64), ("volumen", C.c_double),
                ("ursprung", C.c_double * 3), ("zelle", C.c_double), ("zahl", C.c_int64 * 3), ("n_eintraege", C.c_int64)]


class XpbdTopologie(C.Structure):
    _fields_ = [
        ("n_abstand", C.c_int32), ("abstand_punkte", U32), ("abstand_ruhe", D), ("abstand_nachgiebigkeit", D),
        ("membran_kante", C.POINTER(C.c_uint8)),
        ("n_naehte", C.c_int32), ("naht_punkte", U32), ("naht_ruhe", D), ("naht_nachgiebigkeit", C.c_double),
        ("n_biegungen", C.c_int32), ("biegung_punkte", U32), ("biegung_ruhe", D), ("biegung_nachgiebigkeit", D),
    ]


class XpbdStatistik_c(C.Structure):
    _fields_ = [("kontakte", C.c_int64), ("haftend", C.c_int64), ("tiefe", C.c_double), ("tangential", C.c_double)]


# (name, restype, argtypes) for every entry point of include/gcsim.h
SIGNATUREN = [
    ("gc_create", C.c_void_p, [C.POINTER(Optionen), C.c_char_p, C.c_int32]),
    ("gc_destroy", None, [C.c_void_p]),
    ("gc_last_error", C.c_char_p, [C.c_void_p]),
    ("gc_name", C.c_char_p, [C.c_void_p]),
    ("gc_set_topology", C.c_int, [C.c_void_p, C.POINTER(Topologie)]),
    ("gc_set_param", C.c_int, [C.c_void_p, C.POINTER(Param)]),
    ("gc_set_naht_ruhe", C.c_int, [C.c_void_p, D]),
    ("gc_set_w", C.c_int, [C.c_void_p, D]),
    ("gc_add_zylinder", C.c_int, [C.c_void_p, D, D, C.c_double, D]),
    ("gc_add_sdf", C.c_int, [C.c_void_p, F32, I32, D, C.c_double, D, D]),
    ("gc_set_koerper_lage", C.c_int, [C.c_void_p, C.c_int32, D]),
    ("gc_set_state", C.c_int, [C.c_void_p, D, D]),
    ("gc_get_state", C.c_int, [C.c_void_p, D, D]),
    ("gc_set_dv_vorher", C.c_int, [C.c_void_p, D, C.c_int32]),
    ("gc_build", C.c_int, [C.c_void_p, C.c_double]),
    ("gc_set_fest", C.c_int, [C.c_void_p, I32, D, D, D]),
    ("gc_get_fest", C.c_int, [C.c_void_p, I32, D, D, D]),
    ("gc_solve", C.c_int, [C.c_void_p, D, C.POINTER(CgLauf_c)]),
    ("gc_solve_ohne_speicher", C.c_int, [C.c_void_p, D, C.POINTER(CgLauf_c)]),
    ("gc_apply_A", C.c_int, [C.c_void_p, D, D]),
    ("gc_apply_H", C.c_int, [C.c_void_p, D, D]),
    ("gc_residual", C.c_int, [C.c_void_p, D, D]),
    ("gc_read_load", C.c_int, [C.c_void_p, D]),
    ("gc_read_system", C.c_int, [C.c_void_p, D, D, D, D]),
    ("gc_read_records", C.c_int, [C.c_void_p, D, D, D]),
    ("gc_step", C.c_int, [C.c_void_p, C.c_double, C.c_int32]),
    ("gc_bild", C.c_int, [C.c_void_p, C.c_double, C.c_int32]),
    ("gc_step_graph", C.c_int, [C.c_void_p, C.c_double, C.c_int32]),
    ("gc_read_statistik", C.c_int, [C.c_void_p, C.POINTER(Statistik_c), C.c_int32]),
    ("gc_zeiten", C.c_int, [C.c_void_p, D, D]),
    ("gc_read_naehe", C.c_int, [C.c_void_p, C.c_int32, I32, U32, D, D, D, D]),
    ("gc_read_federn", C.c_int, [C.c_void_p, C.c_int32, U32, D, D, D]),
    ("gc_read_schnitte", C.c_int, [C.c_void_p, C.c_int32, U32, D, D, D, D, U32]),
    ("gc_read_stoesse", C.c_int, [C.c_void_p, C.c_int32, I32, U32, D, D, D, D, I32]),
    ("gc_get_v", C.c_int, [C.c_void_p, D]),
    ("gc_add_sdf_von", C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, D, D]),
    ("gc_batch_create", C.c_void_p, [C.POINTER(C.c_void_p), C.c_int32, C.c_int32, C.c_int32, C.c_char_p, C.c_int32]),
    ("gc_batch_destroy", None, [C.c_void_p]),
    ("gc_batch_last_error", C.c_char_p, [C.c_void_p]),
    ("gc_batch_step", C.c_int, [C.c_void_p, C.c_double, C.c_int32]),
    ("gc_batch_step_graph", C.c_int, [C.c_void_p, C.c_double, C.c_int32]),
    ("gc_batch_autotune", C.c_int, [C.c_void_p, C.c_double, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                    C.POINTER(C.c_double)]),
    ("gc_batch_bloecke", C.c_int, [C.c_void_p, I32, C.POINTER(C.c_int32)]),
    ("gc_batch_zeiten", C.c_int, [C.c_void_p, D, D]),
    ("gc_adjoint_step", C.c_int, [C.c_void_p, C.c_double]),
    ("gc_step_aufnehmen", C.c_int, [C.c_void_p, C.c_double, C.c_int32]),
    ("gc_adjoint_step_aus", C.c_int, [C.c_void_p, C.c_double, C.c_int32]),
    ("gc_adjoint_set", C.c_int, [C.c_void_p, D, D]),
    ("gc_adjoint_get", C.c_int, [C.c_void_p, D, D]),
    ("gc_adjoint_gradient", C.c_int, [C.c_void_p, D, D, D, D, D, C.c_int32]),
    ("gc_adjoint_statistik", C.c_int, [C.c_void_p, C.POINTER(AdjointStat_c), C.c_int32]),
    ("gc_set_steifigkeit", C.c_int, [C.c_void_p, D, D, D, D]),
    ("gc_zustand_sichern", C.c_int, [C.c_void_p, C.c_int32]),
    ("gc_zustand_laden", C.c_int, [C.c_void_p, C.c_int32]),
    ("gc_entscheidungen", C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
    ("gc_entscheidungen_abweichungen", C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.c_int32]),
    ("gc_naht_planen", C.c_int, [C.POINTER(NahtEingabe_c), C.POINTER(NahtAusgabe_c)]),
    ("gc_add_netz", C.c_int, [C.c_void_p, D, C.c_int32, U32, C.c_int32, C.c_double, D, D]),
    ("gc_netz_nachfuehren", C.c_int, [C.c_void_p, C.c_int32, D]),
    ("gc_netz_punktgeschwindigkeit", C.c_int, [C.c_void_p, C.c_int32, D, C.c_int32]),
    ("gc_netz_info", C.c_int, [C.c_void_p, C.c_int32, C.POINTER(NetzAngaben_c)]),
    ("gc_netz_lesen", C.c_int, [C.c_void_p, C.c_int32, D, D, U32, D, U32, U32]),
    ("gc_netz_naechster", C.c_int, [C.c_void_p, C.c_int32, C.c_int32, D, C.c_double, I32, D, D, D, I32, I32]),
    ("gc_netz_durchstoss", C.c_int, [C.c_void_p, C.c_int32, C.c_int32, D, D, D, I32, D, D, D]),
    ("gc_netz_sdf_backen", C.c_int, [C.c_void_p, C.c_int32, I32, D, C.c_double, C.c_double, F32]),
    ("gc_xpbd_topology", C.c_int, [C.c_void_p, C.POINTER(XpbdTopologie)]),
    ("gc_xpbd_bild", C.c_int, [C.c_void_p, C.c_double, C.c_int32, C.POINTER(XpbdStatistik_c)]),
    ("gc_xpbd_laeufe", C.c_int32, [C.c_void_p, I32, I32, I32, I32]),
]

_lib = None


def lib():
    """Load libgcsim.so (built by __graft_entry__.build() or `make -C garmentcraft_amd/csrc`)."""
    global _lib
    if _lib is None:
        if not os.path.exists(SO_PFAD):
            raise GcFehler(f"{SO_PFAD} is not built; run `make -C garmentcraft_amd/csrc`")
        L = C.CDLL(SO_PFAD)
        for name, res, args in SIGNATUREN:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib
