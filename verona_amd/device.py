"""DeviceGrid: one libverona_hip context (device buffers + HIP stream) for a 2D or 3D grid.

Host arrays use Verona's byte order.  In numpy they are addressed like the
Julia arrays, ``arr[v, i, j, k]`` with shape ``(5, Nx+6, Ny+6, Nz+6)`` in
Fortran order (2D: ``(4, Nx+6, Ny+6)``), 0-based.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import BC, INTEGRATOR, PRECISION, REC, RIEMANN, SLOT, StepParams, check, dptr


def host_array(ndim, nx, ny, nz=1):
    """Zeroed host array in Verona layout (ghosts included)."""
    if ndim == 3:
        return np.zeros((5, nx + 6, ny + 6, nz + 6), dtype=np.float64, order="F")
    return np.zeros((4, nx + 6, ny + 6), dtype=np.float64, order="F")


class DeviceGrid:
    def __init__(self, ndim, nx, ny, nz=1, gamma=4.0 / 3.0):
        self.lib = _lib.load()
        self.ndim, self.nx, self.ny, self.nz = ndim, nx, ny, (nz if ndim == 3 else 1)
        self.gamma = float(gamma)
        self.nvar = 5 if ndim == 3 else 4
        h = C.c_void_p()
        check(self.lib.vrn_create(ndim, nx, ny, self.nz, self.gamma, C.byref(h)), "vrn_create")
        self.h = h
        self._params = StepParams()

    # ------------------------------------------------------------------ lifecycle
    def close(self):
        if getattr(self, "h", None):
            self.lib.vrn_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def shape(self):
        if self.ndim == 3:
            return (5, self.nx + 6, self.ny + 6, self.nz + 6)
        return (4, self.nx + 6, self.ny + 6)

    # ------------------------------------------------------------------ data
    def upload(self, slot, arr):
        a = np.asarray(arr, dtype=np.float64)
        if a.shape != self.shape:
            raise ValueError(f"expected shape {self.shape}, got {a.shape}")
        a = np.asfortranarray(a)
        check(self.lib.vrn_upload(self.h, SLOT[slot], dptr(a)), f"vrn_upload({slot})")

    def download(self, slot, out=None):
        if out is None:
            out = np.empty(self.shape, dtype=np.float64, order="F")
        check(self.lib.vrn_download(self.h, SLOT[slot], dptr(out)), f"vrn_download({slot})")
        return out

    def sync(self):
        check(self.lib.vrn_sync(self.h), "vrn_sync")

    def copy(self, dst, src):
        check(self.lib.vrn_copy(self.h, SLOT[dst], SLOT[src]), "vrn_copy")

    def slot_bytes(self):
        b = C.c_int64()
        check(self.lib.vrn_slot_bytes(self.h, C.byref(b)), "vrn_slot_bytes")
        return b.value

    # ------------------------------------------------------------------ reference kernel objects
    def prim_to_cons(self):
        check(self.lib.vrn_prim_to_cons(self.h), "vrn_prim_to_cons")

    def fluxes(self, dim, floor, rec="wenoz", riemann="hll"):
        check(self.lib.vrn_fluxes(self.h, dim, floor, REC[rec], RIEMANN[riemann]), "vrn_fluxes")

    def update(self, dst, dt, dx, dy, dz=1.0):
        check(self.lib.vrn_update(self.h, SLOT[dst], dt, dx, dy, dz), "vrn_update")

    def cons_to_prim(self, src, n_iter, tol):
        check(self.lib.vrn_cons_to_prim(self.h, SLOT[src], int(n_iter), tol), "vrn_cons_to_prim")

    def limit(self, floor):
        check(self.lib.vrn_limit(self.h, floor), "vrn_limit")

    def fill_ghosts(self, bc):
        arr = (C.c_int * 6)(*[BC[b] if isinstance(b, str) else int(b) for b in bc])
        check(self.lib.vrn_fill_ghosts(self.h, arr), "vrn_fill_ghosts")

    def set_nozzle(self, radius, dx, dy, rho, u, umag, eps):
        check(self.lib.vrn_set_nozzle(self.h, radius, dx, dy, rho, u, umag, eps), "vrn_set_nozzle")

    def counters(self):
        a, b = C.c_int64(), C.c_int64()
        check(self.lib.vrn_get_counters(self.h, C.byref(a), C.byref(b)), "vrn_get_counters")
        return a.value, b.value

    def reset_counters(self):
        check(self.lib.vrn_reset_counters(self.h), "vrn_reset_counters")

    # ------------------------------------------------------------------ fused step
    def set_step(self, *, ref_mode=True, rec="wenoz", riemann="hll", integrator="rk2",
                 precision="fp64", n_iter=50, tol=1e-10, floor=1e-7, dt=0.0, cfl=0.0,
                 dx=1.0, dy=1.0, dz=1.0, bc=("none",) * 6):
        p = self._params
        p.ref_mode = int(bool(ref_mode))
        p.rec = REC[rec]
        p.riemann = RIEMANN[riemann]
        p.integrator = INTEGRATOR[integrator]
        p.precision = PRECISION[precision]
        p.n_iter = int(n_iter)
        p.tol = float(tol)
        p.floor_ = float(floor)
        p.dt = float(dt)
        p.cfl = float(cfl)
        p.dx, p.dy, p.dz = float(dx), float(dy), float(dz)
        for f, b in enumerate(bc):
            p.bc[f] = BC[b] if isinstance(b, str) else int(b)
        p.nozzle_phi = 0.0
        return p

    def step(self, phi=0.0, want_dt=False):
        self._params.nozzle_phi = float(phi)
        dt = C.c_double(0.0)
        check(self.lib.vrn_step(self.h, C.byref(self._params), C.byref(dt) if want_dt else None),
              "vrn_step")
        return dt.value if want_dt else None

    # ------------------------------------------------------------------ work partition
    def set_partition(self, rows_per_wave, xseg, seg_len):
        check(self.lib.vrn_set_partition(self.h, int(rows_per_wave), int(xseg), int(seg_len)),
              "vrn_set_partition")

    def set_xy_fusion(self, on=True):
        """3D fast FP64 form: x direction inside the y march (default) or as
        its own sweep; the same arithmetic either way."""
        check(self.lib.vrn_set_xy_fusion(self.h, int(bool(on))), "vrn_set_xy_fusion")

    def partition(self):
        r, x, s = C.c_int(), C.c_int(), C.c_int()
        check(self.lib.vrn_get_partition(self.h, C.byref(r), C.byref(x), C.byref(s)), "vrn_get_partition")
        return r.value, x.value, s.value

    # ------------------------------------------------------------------ timing
    def timing(self, enable=True):
        # True / 1: every kernel; 2: the directional sweeps only
        check(self.lib.vrn_timing_enable(self.h, int(enable)), "vrn_timing_enable")

    def timing_report(self):
        cap, nl = 16, 32
        names = C.create_string_buffer(cap * nl)
        counts = (C.c_int64 * cap)()
        means = (C.c_double * cap)()
        n = C.c_int(0)
        check(self.lib.vrn_timing_report(self.h, names, nl, counts, means, cap, C.byref(n)),
              "vrn_timing_report")
        out = {}
        for k in range(n.value):
            nm = names.raw[k * nl:(k + 1) * nl].split(b"\0", 1)[0].decode()
            out[nm] = (int(counts[k]), float(means[k]))
        return out
