"""The C ABI: libgcsim.so loads and exports every entry point that
include/gcsim.h declares (no device calls: this runs without a GPU)."""
import ctypes
import os
import re

from garmentcraft_amd import _lib

WURZEL = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def deklariert():
    text = open(os.path.join(WURZEL, "include", "gcsim.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gc_[A-Za-z_]+)\s*\(", text)))


def test_alle_deklarierten_symbole_exportiert():
    namen = deklariert()
    assert len(namen) >= 25
    lib = ctypes.CDLL(_lib.SO_PFAD)
    fehlt = [n for n in namen if not hasattr(lib, n)]
    assert not fehlt, fehlt


def test_binding_deckt_den_header():
    assert sorted(n for n, _, _ in _lib.SIGNATUREN) == deklariert()


def test_keine_gpu_meldet_sich_laut():
    """gc_create without a device returns NULL with a reason (the reference's
    GpuLoeser::neu() -> Err, gpu.rs:1136), it never falls back to the CPU."""
    import torch  # noqa: F401  (only to ask whether a device exists)
    if torch.cuda.is_available():
        return
    L = _lib.lib()
    err = ctypes.create_string_buffer(256)
    opt = _lib.Optionen(0, 0, 0, 0, 0)
    h = L.gc_create(ctypes.byref(opt), err, 256)
    assert not h
    assert err.value


# header struct -> its ctypes mirror in garmentcraft_amd/_lib.py
STRUKTUREN = {"gc_optionen": "Optionen", "gc_topologie": "Topologie", "gc_param": "Param", "gc_cg_lauf": "CgLauf_c",
              "gc_statistik": "Statistik_c", "gc_adjoint_stat": "AdjointStat_c", "gc_naht_eingabe": "NahtEingabe_c",
              "gc_naht_ausgabe": "NahtAusgabe_c", "gc_netz_angaben": "NetzAngaben_c",
              "gc_xpbd_topologie": "XpbdTopologie", "gc_xpbd_statistik": "XpbdStatistik_c"}
C_TYPEN = {"uint8_t": ctypes.c_uint8, "int32_t": ctypes.c_int32, "uint32_t": ctypes.c_uint32, "double": ctypes.c_double, "int64_t": ctypes.c_int64,
           "uint64_t": ctypes.c_uint64, "char": ctypes.c_char}


def strukturen():
    """Field names, element types and array lengths of every typedef struct in the header, in order."""
    text = open(os.path.join(WURZEL, "include", "gcsim.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    aus = {}
    for name, koerper in re.findall(r"typedef struct (gc_\w+) \{(.*?)\} \1;", text, flags=re.S):
        felder = []
        for decl in koerper.split(";"):
            decl = " ".join(decl.split())
            if not decl:
                continue
            m = re.match(r"(?:const )?(\w+)\s*(\*?)\s*(\w+(?:\s*,\s*\w+)*)(?:\[(\d+)\])?$", decl)
            assert m, (name, decl)
            typ, zeiger, feld, laenge = m.groups()
            for f in feld.split(","):
                felder.append((f.strip(), typ, bool(zeiger), int(laenge) if laenge else 0))
        aus[name] = felder
    return aus


def test_strukturen_wie_im_header():
    """Every struct the header hands across the boundary has a ctypes mirror
    with the same fields in the same order, the same scalar types, pointers
    where the header has pointers and the same array lengths, so the two
    layouts agree (an added or moved field fails here, without a GPU)."""
    kopf = strukturen()
    assert set(STRUKTUREN) <= set(kopf), sorted(set(STRUKTUREN) - set(kopf))
    for c_name, py_name in STRUKTUREN.items():
        py = getattr(_lib, py_name)
        felder = kopf[c_name]
        assert [f for f, *_ in felder] == [f for f, _ in py._fields_], c_name
        for (f, typ, zeiger, laenge), (_, ct) in zip(felder, py._fields_):
            if zeiger:
                assert issubclass(ct, ctypes._Pointer) or ct in (ctypes.c_void_p,), (c_name, f, ct)
                continue
            erwartet = C_TYPEN[typ]
            if laenge:
                assert issubclass(ct, ctypes.Array) and ct._length_ == laenge and ct._type_ == erwartet, (c_name, f, ct)
            else:
                assert ct == erwartet, (c_name, f, ct)
