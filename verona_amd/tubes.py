"""Batched 1D shock tubes on the GPU: many Verona1D.HARM_HLL problems at once.

    tb = Tubes(lengths, gamma=5/3)
    tb.upload(rho, u, ux, dt, dx)           # concatenated SoA, per-tube dt/dx
    tb.set_method("ppm", "hllc", n_iter=50, tol=1e-10)
    shape = tb.autotune()                   # picks (block, grid) on the real data
    ms = tb.run(500, *shape)                # one persistent hipGraph launch
    rho, u, ux = tb.download()

The launch shape (threads per workgroup, persistent workgroups) is chosen by
`autotune`, which times each candidate on a checkpointed copy of the actual
state and caches the winner per (device, problem signature).
"""
import ctypes as C
import hashlib
import json
import os

import numpy as np

from . import _lib
from ._lib import REC, RIEMANN, VeronaHIPError, dptr

_D = C.c_double
_I = C.c_int
_P = C.c_void_p

TUBE_SIGNATURES = {
    "vrn_tubes_create": (_I, [_I, C.POINTER(C.c_int64), _D, C.POINTER(_P)]),
    "vrn_tubes_destroy": (_I, [_P]),
    "vrn_tubes_last_error": (C.c_char_p, []),
    "vrn_tubes_upload": (_I, [_P, _lib._DP, _lib._DP, _lib._DP, _lib._DP, _lib._DP]),
    "vrn_tubes_download": (_I, [_P, _lib._DP, _lib._DP, _lib._DP]),
    "vrn_tubes_set_method": (_I, [_P, _I, _I, _I, _D]),
    "vrn_tubes_set_precision": (_I, [_P, _I]),
    "vrn_tubes_run": (_I, [_P, _I, _I, _I, _I, _lib._DP]),
    "vrn_tubes_checkpoint": (_I, [_P, _I]),
    "vrn_tubes_info": (_I, [_P, C.POINTER(C.c_int64), C.POINTER(_I)]),
}
_lib.SIGNATURES.update(TUBE_SIGNATURES)

BLOCKS = (64, 128, 256, 512, 1024)
WAVES_PER_CU = (1, 2, 4, 8, 12, 16, 20)
_CACHE = {}


def _check(rc, what):
    if rc != 0:
        msg = _lib.load().vrn_tubes_last_error().decode(errors="replace")
        raise VeronaHIPError(f"{what}: error {rc}: {msg}")


class Tubes:
    def __init__(self, lengths, gamma=5.0 / 3.0):
        self.lib = _lib.load()
        for name, (res, args) in TUBE_SIGNATURES.items():
            f = getattr(self.lib, name)
            f.restype, f.argtypes = res, args
        self.lengths = np.ascontiguousarray(lengths, dtype=np.int64)
        self.offsets = np.concatenate([[0], np.cumsum(self.lengths)])
        self.ncell = int(self.offsets[-1])
        h = _P()
        _check(self.lib.vrn_tubes_create(len(self.lengths), self.lengths.ctypes.data_as(C.POINTER(C.c_int64)),
                                         float(gamma), C.byref(h)), "vrn_tubes_create")
        self.h = h
        self.gamma = float(gamma)
        self.method = None
        self.precision = "fp64"
        n = C.c_int64()
        cu = _I()
        _check(self.lib.vrn_tubes_info(self.h, C.byref(n), C.byref(cu)), "vrn_tubes_info")
        self.cu_count = cu.value

    def close(self):
        if getattr(self, "h", None):
            self.lib.vrn_tubes_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *e):
        self.close()

    def upload(self, rho, u, ux, dt, dx):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (rho, u, ux)]
        for a in arrs:
            if a.shape != (self.ncell,):
                raise ValueError(f"expected {self.ncell} cells, got {a.shape}")
        dt = np.ascontiguousarray(np.broadcast_to(dt, self.lengths.shape), dtype=np.float64)
        dx = np.ascontiguousarray(np.broadcast_to(dx, self.lengths.shape), dtype=np.float64)
        _check(self.lib.vrn_tubes_upload(self.h, dptr(arrs[0]), dptr(arrs[1]), dptr(arrs[2]), dptr(dt),
                                         dptr(dx)), "vrn_tubes_upload")

    def download(self):
        out = [np.empty(self.ncell) for _ in range(3)]
        _check(self.lib.vrn_tubes_download(self.h, dptr(out[0]), dptr(out[1]), dptr(out[2])),
               "vrn_tubes_download")
        return out

    def set_method(self, rec="ppm", riemann="hllc", n_iter=50, tol=1e-10):
        _check(self.lib.vrn_tubes_set_method(self.h, REC[rec], RIEMANN[riemann], int(n_iter), float(tol)),
               "vrn_tubes_set_method")
        self.method = (rec, riemann, int(n_iter), float(tol))

    def set_precision(self, precision="fp64"):
        """"fp64" (default): Verona1D's FP64 operation order, bit-reproducible;
        "fp64fast": FMA contraction and reciprocal-based division."""
        codes = {"fp64": 0, "fp64fast": 2}
        if precision not in codes:
            raise ValueError(f"precision must be one of {sorted(codes)}, got {precision!r}")
        _check(self.lib.vrn_tubes_set_precision(self.h, codes[precision]), "vrn_tubes_set_precision")
        self.precision = precision

    def run(self, nsteps, block=128, grid=None, use_graph=True):
        if grid is None:
            grid = 2 * self.cu_count
        ms = _D(0.0)
        _check(self.lib.vrn_tubes_run(self.h, int(nsteps), int(block), int(grid), int(bool(use_graph)),
                                      C.byref(ms)), "vrn_tubes_run")
        return ms.value

    # ------------------------------------------------------------------ autotuner
    def signature(self):
        m = hashlib.sha1()
        m.update(self.lengths.tobytes())
        m.update(repr((self.gamma, self.method, self.precision, self.cu_count)).encode())
        return m.hexdigest()

    def autotune(self, probe_steps=2, cache_file=None, verbose=False):
        """Time every (block, workgroups-per-CU) candidate for probe_steps on a
        checkpoint of the current state, restore the state, and return the
        fastest (block, grid).  Results are cached per problem signature."""
        if self.method is None:
            raise VeronaHIPError("set_method before autotune")
        key = self.signature()
        if key in _CACHE:
            return _CACHE[key]
        if cache_file and os.path.exists(cache_file):
            try:
                disk = json.load(open(cache_file))
                if key in disk:
                    _CACHE[key] = tuple(disk[key]["shape"])
                    return _CACHE[key]
            except (OSError, ValueError):
                pass
        _check(self.lib.vrn_tubes_checkpoint(self.h, 0), "checkpoint")
        results = []
        for block in BLOCKS:
            for wpc in WAVES_PER_CU:
                grid = max(1, (wpc * 64 * self.cu_count) // block)
                self.run(probe_steps, block, grid, use_graph=False)  # warm
                _check(self.lib.vrn_tubes_checkpoint(self.h, 1), "restore")
                ms = self.run(probe_steps, block, grid, use_graph=False)
                _check(self.lib.vrn_tubes_checkpoint(self.h, 1), "restore")
                results.append((ms, block, grid))
                if verbose:
                    print(f"  autotune block={block} grid={grid}: {ms:.3f} ms", flush=True)
        best = min(results)
        shape = (best[1], best[2])
        _CACHE[key] = shape
        if cache_file:
            try:
                disk = json.load(open(cache_file)) if os.path.exists(cache_file) else {}
                disk[key] = {"shape": list(shape), "trials": [[r[0], r[1], r[2]] for r in results]}
                os.makedirs(os.path.dirname(os.path.abspath(cache_file)), exist_ok=True)
                json.dump(disk, open(cache_file, "w"))
            except (OSError, ValueError):
                pass
        self.autotune_trials = results
        return shape


def shock_tube_batch(ntubes, lmin=256, lmax=8192, seed=42, gamma=5.0 / 3.0):
    """Seeded batch of uneven tubes: lengths uniform in [lmin, lmax]; per side
    rho log-uniform in [0.1, 10], p log-uniform in [1e-4, 10], v uniform in
    [-0.9, 0.9]; diaphragm at the middle; dx = 1/L, dt = 0.4 dx.  Draws come
    from one SplitMix64 stream (top 53 bits)."""
    from .bc import uniform_stream
    it = uniform_stream(seed)
    lengths = np.array([lmin + int(next(it) * (lmax - lmin + 1)) for _ in range(ntubes)], dtype=np.int64)
    total = int(lengths.sum())
    rho = np.empty(total)
    u = np.empty(total)
    ux = np.empty(total)
    off = 0
    for L in lengths:
        sides = []
        for _ in range(2):
            r = 10 ** (-1 + 2 * next(it))
            p = 10 ** (-4 + 5 * next(it))
            v = -0.9 + 1.8 * next(it)
            sides.append((r, p / (gamma - 1), v / np.sqrt(1 - v * v)))
        h = L // 2
        rho[off:off + h], u[off:off + h], ux[off:off + h] = sides[0]
        rho[off + h:off + L], u[off + h:off + L], ux[off + h:off + L] = sides[1]
        off += L
    dx = 1.0 / lengths
    dt = 0.4 * dx
    return lengths, rho, u, ux, dt, dx
