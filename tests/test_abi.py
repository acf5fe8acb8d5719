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
