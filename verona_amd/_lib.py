"""ctypes binding of libverona_hip.so (the C ABI declared in include/verona_hip.h).

The shared library is the product's compute path; there is no CPU fallback.
If it is missing, or no GPU is present when compute is requested, the calls
raise instead of silently substituting another implementation.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("VERONA_HIP_LIB", os.path.join(_HERE, "libverona_hip.so"))
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "verona_hip.h")

VRN_OK = 0
REC = {"minmod": 0, "ppm": 1, "wenoz": 2}
RIEMANN = {"hll": 0, "hllc": 1}
INTEGRATOR = {"rk2": 0, "rk3": 1}
PRECISION = {"fp64": 0, "fp32": 1, "fp64fast": 2}
BC = {"none": 0, "periodic": 1, "outflow": 2, "nozzle": 3}
SLOT = {"P": 0, "U": 1, "Uhalf": 2, "Fx": 3, "Fy": 4, "Fz": 5}


class StepParams(C.Structure):
    _fields_ = [
        ("ref_mode", C.c_int), ("rec", C.c_int), ("riemann", C.c_int), ("integrator", C.c_int),
        ("precision", C.c_int), ("n_iter", C.c_int), ("tol", C.c_double), ("floor_", C.c_double),
        ("dt", C.c_double), ("cfl", C.c_double), ("dx", C.c_double), ("dy", C.c_double),
        ("dz", C.c_double), ("bc", C.c_int * 6), ("nozzle_phi", C.c_double),
    ]


class VeronaHIPError(RuntimeError):
    pass


_lib = None

_P = C.c_void_p
_D = C.c_double
_I = C.c_int
_I64P = C.POINTER(C.c_int64)
_DP = C.POINTER(C.c_double)

# name -> (restype, argtypes); every symbol declared in include/verona_hip.h
SIGNATURES = {
    "vrn_create": (_I, [_I, _I, _I, _I, _D, C.POINTER(_P)]),
    "vrn_destroy": (_I, [_P]),
    "vrn_sync": (_I, [_P]),
    "vrn_last_error": (C.c_char_p, []),
    "vrn_upload": (_I, [_P, _I, _DP]),
    "vrn_download": (_I, [_P, _I, _DP]),
    "vrn_prim_to_cons": (_I, [_P]),
    "vrn_copy": (_I, [_P, _I, _I]),
    "vrn_slot_bytes": (_I, [_P, _I64P]),
    "vrn_fluxes": (_I, [_P, _I, _D, _I, _I]),
    "vrn_update": (_I, [_P, _I, _D, _D, _D, _D]),
    "vrn_cons_to_prim": (_I, [_P, _I, _I, _D]),
    "vrn_limit": (_I, [_P, _D]),
    "vrn_fill_ghosts": (_I, [_P, C.POINTER(_I)]),
    "vrn_set_nozzle": (_I, [_P, _D, _D, _D, _D, _D, _D, _D]),
    "vrn_get_counters": (_I, [_P, _I64P, _I64P]),
    "vrn_reset_counters": (_I, [_P]),
    "vrn_step": (_I, [_P, C.POINTER(StepParams), _DP]),
    "vrn_timing_enable": (_I, [_P, _I]),
    "vrn_timing_report": (_I, [_P, C.c_char_p, _I, _I64P, _DP, _I, C.POINTER(_I)]),
    "vrn_set_partition": (_I, [_P, _I, _I, _I]),
    "vrn_get_partition": (_I, [_P, C.POINTER(_I), C.POINTER(_I), C.POINTER(_I)]),
    "vrn_set_xy_fusion": (_I, [_P, _I]),
}


def load():
    """Load the library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise VeronaHIPError(
                f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                "(hipcc --offload-arch=gfx950); there is no CPU fallback")
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


def check(rc, what=""):
    if rc != VRN_OK:
        msg = load().vrn_last_error().decode(errors="replace")
        raise VeronaHIPError(f"{what}: error {rc}: {msg}")
    return rc


def dptr(a):
    """double* of a contiguous float64 numpy array (C or Fortran order)."""
    import numpy as np
    assert a.dtype == np.float64 and (a.flags["C_CONTIGUOUS"] or a.flags["F_CONTIGUOUS"])
    return a.ctypes.data_as(_DP)
