"""The C-ABI library loads and exports every symbol include/verona_hip.h
declares (no compute calls: runs without a GPU)."""
import ctypes
import os
import re

from verona_amd import _lib

HDR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")


def declared_functions():
    names = []
    for fn in sorted(os.listdir(HDR)):
        if fn.endswith(".h"):
            txt = open(os.path.join(HDR, fn)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names += re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(vrn_[a-z_0-9]+)\s*\(", txt, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_api():
    names = declared_functions()
    for must in ("vrn_create", "vrn_fluxes", "vrn_update", "vrn_cons_to_prim", "vrn_prim_to_cons",
                 "vrn_limit", "vrn_fill_ghosts", "vrn_step"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    assert set(declared_functions()) <= set(_lib.SIGNATURES)


def test_last_error_is_callable_without_gpu():
    lib = _lib.load()
    assert isinstance(lib.vrn_last_error(), bytes)
    # argument validation happens before any device call
    h = ctypes.c_void_p()
    rc = lib.vrn_create(4, 8, 8, 8, 1.4, ctypes.byref(h))
    assert rc == -1
    assert b"ndim" in lib.vrn_last_error()
