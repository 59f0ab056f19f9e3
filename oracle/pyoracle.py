"""ctypes bindings of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: a checker for the HIP library.  Only tests/, the
smoke check and bench.py's CPU-baseline leg import this module; the product
package never does.

Arrays follow Verona's layouts: 3D AoS ``(NZ, NY, NX, 5)`` in C order (which
is ``arr[5, NX, NY, NZ]`` column-major in Julia), 2D ``(NY, NX, 4)``, 1D SoA.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

BC_NONE, BC_PERIODIC, BC_OUTFLOW, BC_NOZZLE = 0, 1, 2, 3
REC = {"minmod": 0, "ppm": 1, "wenoz": 2}
RIEMANN = {"hll": 0, "hllc": 1}


class Mode(C.Structure):
    _fields_ = [
        ("ref", C.c_int), ("rec", C.c_int), ("riemann", C.c_int), ("integrator", C.c_int),
        ("n_iter", C.c_int), ("tol", C.c_double), ("gamma", C.c_double), ("floor_", C.c_double),
        ("dt", C.c_double), ("cfl", C.c_double), ("dx", C.c_double), ("dy", C.c_double),
        ("dz", C.c_double),
    ]


class BC3(C.Structure):
    _fields_ = [
        ("type", C.c_int * 6), ("dx", C.c_double), ("dy", C.c_double),
        ("noz_radius", C.c_double), ("noz_rho", C.c_double), ("noz_u", C.c_double),
        ("noz_umag", C.c_double), ("noz_eps", C.c_double), ("noz_phi", C.c_double),
        ("phi_seq", C.POINTER(C.c_double)),
    ]


_SIG = None


def _signatures():
    d = C.c_double
    p = C.POINTER(C.c_double)
    i = C.c_int
    L = C.c_long
    return {
        "orc_wenoz_v": (None, [p, L, p, p]),
        "orc_ppm_v": (None, [p, L, p, p]),
        "orc_minmod_v": (None, [p, L, p, p]),
        "orc3_p2u_v": (None, [p, L, d, p]),
        "orc3_p2f_v": (None, [p, L, d, i, p]),
        "orc3_face_flux_v": (None, [p, p, L, d, i, i, p, p, p]),
        "orc1_face_flux_v": (None, [p, p, L, d, i, p]),
        "orc3_ptou": (None, [p, p, i, i, i, d]),
        "orc3_limit": (None, [p, d, i, i, i]),
        "orc3_fluxes": (None, [p, d, d, p, i, i, i, i, i, i]),
        "orc3_update": (None, [p, p, d, d, d, d, p, p, p, i, i, i]),
        "orc3_f1dw": (None, [p, d, d, p, p]),
        "orc3_c2p_cell": (i, [p, p, d, i, d, C.POINTER(i)]),
        "orc3_c2p": (L, [p, p, i, i, i, d, i, d, C.POINTER(L)]),
        "orc3_c2p_batch": (L, [p, p, L, d, i, d, C.POINTER(i), C.POINTER(i)]),
        "orc3_fill_periodic": (None, [p, i, i, i, i]),
        "orc3_fill_outflow": (None, [p, i, i, i, i, i]),
        "orc3_fill_nozzle": (None, [p, i, i, i, d, d, d, d, d, d, d, d]),
        "orc3_max_rate": (d, [p, i, i, i, d, d, d, d]),
        "orc3_run": (L, [p, i, i, i, C.POINTER(Mode), C.POINTER(BC3), i, p, p]),
        "orc2_ptou": (None, [p, p, i, i, d]),
        "orc2_limit": (None, [p, d, i, i]),
        "orc2_fluxes": (None, [p, d, d, p, i, i, i, i, i]),
        "orc2_update": (None, [p, p, d, d, d, p, p, i, i]),
        "orc2_c2p": (L, [p, p, i, i, d, i, d, C.POINTER(L)]),
        "orc2_fill_periodic": (None, [p, i, i, i]),
        "orc2_fill_outflow": (None, [p, i, i, i, i]),
        "orc2_max_rate": (d, [p, i, i, d, d, d]),
        "orc2_run": (L, [p, i, i, C.POINTER(Mode), C.POINTER(i), i, p, p]),
        "orc1_f_ptou": (None, [p, d, p]),
        "orc1_jacobian": (None, [p, d, p]),
        "orc1_solve3": (None, [p, p, p]),
        "orc1_c2p": (None, [p, p, p, p, p, p, L, d, i, d]),
        "orc1_run": (i, [p, p, p, L, d, d, d, i, d, i, i, i, d, p]),
        "orc4_w0_from_prims": (None, [p, L, d, p]),
        "orc4_solve": (L, [p, p, L, d, i, d, p, C.POINTER(i), C.POINTER(C.c_uint8)]),
        "orc4_features": (None, [p, L, C.POINTER(C.c_float), p, p, C.POINTER(C.c_uint8)]),
    }


def load(path):
    """Load an oracle build from `path` with every signature bound (the real
    oracle, or one of the planted-mistake builds the mutant tests make)."""
    lb = C.CDLL(path)
    for name, (res, args) in _signatures().items():
        f = getattr(lb, name)
        f.restype = res
        f.argtypes = args
    return lb


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build())")
        _LIB = load(path)
    return _LIB


class use:
    """Context manager: route every function of this module through another
    oracle build (a loaded library from load())."""

    def __init__(self, other):
        self.other = other

    def __enter__(self):
        global _LIB
        lib()
        self.saved = _LIB
        _LIB = self.other
        return self

    def __exit__(self, *exc):
        global _LIB
        _LIB = self.saved
        return False


def _p(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], (a.dtype, a.flags)
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ---------------------------------------------------------------- point functions
def wenoz(q):
    q = _f64(q).reshape(-1, 5)
    lo = np.empty(len(q))
    hi = np.empty(len(q))
    lib().orc_wenoz_v(_p(q), len(q), _p(lo), _p(hi))
    return lo, hi


def ppm(q):
    q = _f64(q).reshape(-1, 5)
    lo = np.empty(len(q))
    hi = np.empty(len(q))
    lib().orc_ppm_v(_p(q), len(q), _p(lo), _p(hi))
    return lo, hi


def minmod(q):
    q = _f64(q).reshape(-1, 3)
    lo = np.empty(len(q))
    hi = np.empty(len(q))
    lib().orc_minmod_v(_p(q), len(q), _p(lo), _p(hi))
    return lo, hi


def p2u3(P, gamma):
    P = _f64(P).reshape(-1, 5)
    U = np.empty_like(P)
    lib().orc3_p2u_v(_p(P), len(P), gamma, _p(U))
    return U


def p2f3(P, gamma, dim):
    P = _f64(P).reshape(-1, 5)
    F = np.empty_like(P)
    lib().orc3_p2f_v(_p(P), len(P), gamma, dim, _p(F))
    return F


def face_flux3(PL, PR, gamma, dim, riemann="hll"):
    PL = _f64(PL).reshape(-1, 5)
    PR = _f64(PR).reshape(-1, 5)
    F = np.empty_like(PL)
    cmax = np.empty(len(PL))
    cmin = np.empty(len(PL))
    lib().orc3_face_flux_v(_p(PL), _p(PR), len(PL), gamma, dim, RIEMANN[riemann], _p(F), _p(cmax), _p(cmin))
    return F, cmax, cmin


def face_flux1(PL, PR, gamma, riemann="hll"):
    PL = _f64(PL).reshape(-1, 3)
    PR = _f64(PR).reshape(-1, 3)
    F = np.empty_like(PL)
    lib().orc1_face_flux_v(_p(PL), _p(PR), len(PL), gamma, RIEMANN[riemann], _p(F))
    return F


def f1dw3(U, W, gamma):
    """1DW residual f(W) and the iteration's analytic f'(W) for one conserved state."""
    U = _f64(U).reshape(5)
    f = C.c_double(0)
    fp = C.c_double(0)
    lib().orc3_f1dw(_p(U), float(W), gamma, C.byref(f), C.byref(fp))
    return f.value, fp.value


def f_ptou1(x, gamma):
    """F_ptoU of the 1D Newton: conserved (rho Gamma, T^t_t, T^t_x) of P = (rho, u, u^x)."""
    x = _f64(x).reshape(3)
    f = np.empty(3)
    lib().orc1_f_ptou(_p(x), gamma, _p(f))
    return f


def jacobian1(x, gamma):
    """Analytic Jacobian of F_ptoU as the 1D Newton forms it, J[i, j] = dF_i/dx_j."""
    x = _f64(x).reshape(3)
    J = np.empty(9)
    lib().orc1_jacobian(_p(x), gamma, _p(J))
    return J.reshape(3, 3).T.copy()  # stored column-major


def solve3(A, b):
    """The 1D Newton's 3x3 solve of A x = b (A given as a 3x3 row-major array)."""
    Acm = np.ascontiguousarray(np.asarray(A, dtype=np.float64).T).reshape(9)
    b = _f64(b).reshape(3)
    x = np.empty(3)
    lib().orc1_solve3(_p(Acm), _p(b), _p(x))
    return x


# ---------------------------------------------------------------- 3D
def dims3(P):
    NZ, NY, NX, nv = P.shape
    assert nv == 5
    return NX, NY, NZ


def ptou3(P, gamma):
    U = np.zeros_like(P)
    lib().orc3_ptou(_p(P), _p(U), *dims3(P), gamma)
    return U


def limit3(P, floor):
    lib().orc3_limit(_p(P), floor, *dims3(P))


def fluxes3(P, gamma, floor, dim, rec="wenoz", riemann="hll", F=None):
    if F is None:
        F = np.zeros_like(P)
    lib().orc3_fluxes(_p(P), gamma, floor, _p(F), dim, *dims3(P), REC[rec], RIEMANN[riemann])
    return F


def update3(U, dt, dx, dy, dz, Fx, Fy, Fz, Ub=None):
    if Ub is None:
        Ub = np.zeros_like(U)
    lib().orc3_update(_p(U), _p(Ub), dt, dx, dy, dz, _p(Fx), _p(Fy), _p(Fz), *dims3(U))
    return Ub


def c2p3(U, P, gamma, n_iter, tol):
    """In place on P (guess in, result out).  Returns (n_unrecovered, method_counts)."""
    counts = (C.c_long * 5)()
    bad = lib().orc3_c2p(_p(U), _p(P), *dims3(P), gamma, n_iter, tol, counts)
    return bad, list(counts)


def c2p3_cell(U, P, gamma, n_iter, tol):
    U = _f64(U).copy()
    P = _f64(P).copy()
    it = C.c_int(0)
    m = lib().orc3_c2p_cell(_p(U), _p(P), gamma, n_iter, tol, C.byref(it))
    return P, m, it.value


def c2p3_batch(U, P, gamma, n_iter, tol):
    U = _f64(U).reshape(-1, 5)
    P = _f64(P).reshape(-1, 5).copy()
    n = len(U)
    meth = np.zeros(n, np.int32)
    its = np.zeros(n, np.int32)
    bad = lib().orc3_c2p_batch(_p(U), _p(P), n, gamma, n_iter, tol,
                               meth.ctypes.data_as(C.POINTER(C.c_int)),
                               its.ctypes.data_as(C.POINTER(C.c_int)))
    return P, bad, meth, its


def fill_periodic3(P, mask=7):
    lib().orc3_fill_periodic(_p(P), *dims3(P), mask)


def fill_outflow3(P, lo=7, hi=7):
    lib().orc3_fill_outflow(_p(P), *dims3(P), lo, hi)


def max_rate3(P, gamma, dx, dy, dz):
    return lib().orc3_max_rate(_p(P), *dims3(P), gamma, dx, dy, dz)


def make_mode(ref=True, rec="wenoz", riemann="hll", integrator="rk2", n_iter=50, tol=1e-10,
              gamma=4.0 / 3.0, floor=1e-7, dt=0.0, cfl=0.0, dx=1.0, dy=1.0, dz=1.0):
    return Mode(int(ref), REC[rec], RIEMANN[riemann], 1 if integrator == "rk3" else 0, n_iter, tol,
                gamma, floor, dt, cfl, dx, dy, dz)


def make_bc3(types, dx=1.0, dy=1.0, nozzle=None, phi_seq=None):
    bc = BC3()
    for k, t in enumerate(types):
        bc.type[k] = t
    bc.dx, bc.dy = dx, dy
    keep = None
    if nozzle is not None:
        bc.noz_radius, bc.noz_rho, bc.noz_u, bc.noz_umag, bc.noz_eps, bc.noz_phi = nozzle
    if phi_seq is not None:
        keep = _f64(phi_seq)
        bc.phi_seq = _p(keep)
    return bc, keep


def run3(P, mode, bc, nsteps):
    """In place on P.  Returns (n_unrecovered, t, dt_last)."""
    t = C.c_double(0)
    dtl = C.c_double(0)
    bad = lib().orc3_run(_p(P), *dims3(P), C.byref(mode), C.byref(bc), nsteps, C.byref(t), C.byref(dtl))
    return bad, t.value, dtl.value


# ---------------------------------------------------------------- 2D
def dims2(P):
    NY, NX, nv = P.shape
    assert nv == 4
    return NX, NY


def ptou2(P, gamma):
    U = np.zeros_like(P)
    lib().orc2_ptou(_p(P), _p(U), *dims2(P), gamma)
    return U


def limit2(P, floor):
    lib().orc2_limit(_p(P), floor, *dims2(P))


def fluxes2(P, gamma, floor, dim, rec="wenoz", riemann="hll", F=None):
    if F is None:
        F = np.zeros_like(P)
    lib().orc2_fluxes(_p(P), gamma, floor, _p(F), dim, *dims2(P), REC[rec], RIEMANN[riemann])
    return F


def update2(U, dt, dx, dy, Fx, Fy, Ub=None):
    if Ub is None:
        Ub = np.zeros_like(U)
    lib().orc2_update(_p(U), _p(Ub), dt, dx, dy, _p(Fx), _p(Fy), *dims2(U))
    return Ub


def c2p2(U, P, gamma, n_iter, tol):
    ns = C.c_long(0)
    lib().orc2_c2p(_p(U), _p(P), *dims2(P), gamma, n_iter, tol, C.byref(ns))
    return ns.value


def fill_periodic2(P, mask=3):
    lib().orc2_fill_periodic(_p(P), *dims2(P), mask)


def max_rate2(P, gamma, dx, dy):
    return lib().orc2_max_rate(_p(P), *dims2(P), gamma, dx, dy)


def run2(P, mode, bc_types, nsteps):
    bc = (C.c_int * 4)(*bc_types)
    t = C.c_double(0)
    dtl = C.c_double(0)
    bad = lib().orc2_run(_p(P), *dims2(P), C.byref(mode), bc, nsteps, C.byref(t), C.byref(dtl))
    return bad, t.value, dtl.value


# ---------------------------------------------------------------- 1D
def c2p1(U, P, gamma, n_iter, tol):
    """U, P: (3, N) SoA; P updated in place."""
    lib().orc1_c2p(_p(U[0]), _p(U[1]), _p(U[2]), _p(P[0]), _p(P[1]), _p(P[2]), P.shape[1], gamma, n_iter, tol)


def run1(P, dt, dx, T, gamma, rec="minmod", riemann="hll", n_iter=40, tol=1e-6, max_steps=0):
    """P: (3, N) float64 SoA, updated in place.  Returns (steps, t)."""
    assert P.shape[0] == 3 and P.flags["C_CONTIGUOUS"]
    t = C.c_double(0)
    steps = lib().orc1_run(_p(P[0]), _p(P[1]), _p(P[2]), P.shape[1], dt, dx, T, max_steps, gamma,
                           REC[rec], RIEMANN[riemann], n_iter, tol, C.byref(t))
    return steps, t.value


# ---------------------------------------------------------------- batched con2prim (C4)
def c2p4_w0_from_prims(Pg, gamma):
    """Initial W of primitive guesses Pg [n][5] as UtoP forms it."""
    Pg = _f64(Pg)
    n = Pg.shape[0]
    W0 = np.empty(n)
    lib().orc4_w0_from_prims(_p(Pg), n, gamma, _p(W0))
    return W0


def c2p4_solve(U, W0, gamma, n_iter, rtol):
    """1DW Newton (relative stop, no Gamma cap) from initial W0.
    Returns (P [n][5], iters [n], converged [n] bool)."""
    U = _f64(U)
    W0 = _f64(W0)
    n = U.shape[0]
    P = np.full((n, 5), np.nan)
    it = np.zeros(n, dtype=np.int32)
    cv = np.zeros(n, dtype=np.uint8)
    lib().orc4_solve(_p(U), _p(W0), n, gamma, n_iter, rtol, _p(P), it.ctypes.data_as(C.POINTER(C.c_int)),
                     cv.ctypes.data_as(C.POINTER(C.c_uint8)))
    return P, it, cv.astype(bool)


def c2p4_features(U):
    """Features X [n][4] (float32), E, z and the domain flag of the learned guess."""
    U = _f64(U)
    n = U.shape[0]
    X = np.empty((n, 4), dtype=np.float32)
    E = np.empty(n)
    z = np.empty(n)
    ok = np.zeros(n, dtype=np.uint8)
    lib().orc4_features(_p(U), n, X.ctypes.data_as(C.POINTER(C.c_float)), _p(E), _p(z),
                        ok.ctypes.data_as(C.POINTER(C.c_uint8)))
    return X, E, z, ok.astype(bool)
