"""vrn_c2p_solve_noalloc against vrn_c2p_solve (include/verona_c2p.h).

The no-allocation entry point runs the same kernels on caller-owned device
arrays with its working memory in the library's fixed device array.  Its P
and statistics must equal the handle-based solve's bit for bit, on two
calls in a row with the second smaller (the array keeps the first call's
counters and rescue list, which must not leak into the second), for the
three guesses it offers, and with an iteration limit low enough that the
rescue list (the cascade after 1DW) is used.  Sizes beyond the array's
capacity and the learned guess are rejected before any device work.

The caller-owned device arrays come from the HIP runtime the library itself
is linked against (hipMalloc / hipMemcpy through ctypes): test plumbing only.
"""
import ctypes as C

import numpy as np
import pytest

from helpers import progress
from verona_amd._lib import VeronaHIPError
from verona_amd.c2p import Con2PrimBatch, c4_states, solve_noalloc, warm_guess

pytestmark = pytest.mark.gpu
G = 4.0 / 3.0
MAX_CELLS = 1 << 20
H2D, D2H = 1, 2


def _hip():
    from verona_amd import _lib
    _lib.load()
    # the HIP runtime the library loaded (same handle: already mapped)
    with open("/proc/self/maps") as f:
        paths = {ln.split()[-1] for ln in f if "libamdhip64.so" in ln}
    assert paths, "libamdhip64 is not loaded"
    hip = C.CDLL(sorted(paths)[0])
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipDeviceSynchronize.argtypes = []
    return hip


class DevArray:
    """A device copy of a (5, n) float64 array, freed on close."""

    def __init__(self, hip, host):
        self.hip, self.shape = hip, host.shape
        self.nbytes = host.size * 8
        self.ptr = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.ptr), self.nbytes) == 0
        h = np.ascontiguousarray(host, dtype=np.float64)
        assert hip.hipMemcpy(self.ptr, h.ctypes.data, self.nbytes, H2D) == 0

    def get(self):
        out = np.empty(self.shape)
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, D2H) == 0
        return out

    def close(self):
        if self.ptr:
            self.hip.hipFree(self.ptr)
            self.ptr = C.c_void_p()


STAT_KEYS = ("n", "n_unrecovered", "n_fallback", "n_guess_bad", "iters_mean", "iters_max", "wave_max_mean", "hist")


@pytest.mark.parametrize("guess,n_iter", [("generic", 100), ("warm", 100), ("analytic", 100), ("generic", 3)])
def test_noalloc_equals_handle_solve(guess, n_iter):
    hip = _hip()
    with Con2PrimBatch(1 << 18, G) as cp:
        for n, seed in ((1 << 18, 5), (3001, 6)):
            progress(f"no-alloc solve guess={guess} n_iter={n_iter} n={n}")
            U, P = c4_states(n, seed=seed, gamma=G)
            Pprev = warm_guess(P, seed=seed + 1)
            cp.upload(U, Pprev)
            ref = cp.solve(guess, n_iter=n_iter)
            Pref = cp.download()
            bufs = [DevArray(hip, U.T), DevArray(hip, Pprev.T), DevArray(hip, np.full((5, n), np.nan))]
            try:
                Ud, Ppd, Pd = bufs
                st = solve_noalloc(n, Ud.ptr.value, Pd.ptr.value, G, guess, P_prev_dev=Ppd.ptr.value,
                                   n_iter=n_iter, stream=None)
                Pg = Pd.get().T
            finally:
                for b in bufs:
                    b.close()
            same = np.array_equal(Pg, Pref, equal_nan=True)
            progress(f"  fallback {st['n_fallback']} (handle {ref['n_fallback']}), unrecovered "
                     f"{st['n_unrecovered']}, bit-identical {same}")
            assert same
            for k in STAT_KEYS:
                assert st[k] == ref[k], k
            if n_iter == 3:
                assert st["n_fallback"] > 0  # the rescue list was used


def test_noalloc_rejects_what_it_cannot_hold():
    hip = _hip()
    buf = DevArray(hip, np.zeros((5, 1)))
    try:
        for n, guess in ((MAX_CELLS + 1, "analytic"), (0, "analytic"), (1, "learned")):
            with pytest.raises(VeronaHIPError) as e:
                solve_noalloc(n, buf.ptr.value, buf.ptr.value, G, guess)
            progress(f"  n={n} guess={guess}: {e.value}")
        with pytest.raises(VeronaHIPError):
            solve_noalloc(1, buf.ptr.value, buf.ptr.value, G, "warm", P_prev_dev=None)
        assert (buf.get() == 0).all()  # nothing was written
    finally:
        buf.close()
