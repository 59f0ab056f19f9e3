"""Shared test helpers: seeded initial states and error metrics."""
import sys

import numpy as np


def rel_l1(a, b):
    """Relative L1 difference sum|a-b| / sum|b| (0 when both are zero)."""
    den = np.abs(b).sum()
    num = np.abs(a - b).sum()
    return num / den if den > 0 else num


def interior(P):
    return P[(slice(None),) + (slice(3, -3),) * (P.ndim - 1)]


def per_var_rel_l1(a, b):
    return [rel_l1(a[v], b[v]) for v in range(a.shape[0])]


def progress(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def state3(nx, ny, nz, seed=0, jump=True, umax=0.5, gamma=4.0 / 3.0):
    """Smooth 3D state plus an optional oblique density/pressure jump,
    including ghost cells; Verona layout (5, nx+6, ny+6, nz+6), Fortran order."""
    rng = np.random.default_rng(seed)
    NX, NY, NZ = nx + 6, ny + 6, nz + 6
    x = (np.arange(NX) - 3 + 0.5) / nx
    y = (np.arange(NY) - 3 + 0.5) / ny
    z = (np.arange(NZ) - 3 + 0.5) / nz
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    rho = 1.0 + 0.3 * np.sin(2 * np.pi * X) * np.cos(2 * np.pi * Y) + 0.2 * np.sin(2 * np.pi * Z)
    u = 1.0 + 0.4 * np.cos(2 * np.pi * (X + Z)) * np.sin(2 * np.pi * Y)
    ux = umax * np.sin(2 * np.pi * Y) * np.cos(2 * np.pi * Z)
    uy = umax * np.sin(2 * np.pi * Z + 0.3) * np.cos(2 * np.pi * X)
    uz = umax * np.sin(2 * np.pi * X + 0.7) * np.cos(2 * np.pi * Y)
    if jump:
        m = (X + 0.5 * Y + 0.25 * Z) > 0.8
        rho = np.where(m, 0.2 * rho, rho)
        u = np.where(m, 0.05 * u, u)
    P = np.zeros((5, NX, NY, NZ), order="F")
    P[0] = rho * (1 + 1e-3 * rng.standard_normal(rho.shape))
    P[1] = u * (1 + 1e-3 * rng.standard_normal(u.shape))
    P[2], P[3], P[4] = ux, uy, uz
    return P


def state2(nx, ny, seed=0, jump=True, umax=0.5):
    rng = np.random.default_rng(seed)
    NX, NY = nx + 6, ny + 6
    x = (np.arange(NX) - 3 + 0.5) / nx
    y = (np.arange(NY) - 3 + 0.5) / ny
    X, Y = np.meshgrid(x, y, indexing="ij")
    rho = 1.0 + 0.3 * np.sin(2 * np.pi * X) * np.cos(2 * np.pi * Y)
    u = 1.0 + 0.4 * np.cos(2 * np.pi * (X + Y))
    ux = umax * np.sin(2 * np.pi * Y)
    uy = umax * np.cos(2 * np.pi * X + 0.4)
    if jump:
        m = (X + 0.5 * Y) > 0.8
        rho = np.where(m, 0.2 * rho, rho)
        u = np.where(m, 0.05 * u, u)
    P = np.zeros((4, NX, NY), order="F")
    P[0] = rho * (1 + 1e-3 * rng.standard_normal(rho.shape))
    P[1] = u * (1 + 1e-3 * rng.standard_normal(u.shape))
    P[2], P[3] = ux, uy
    return P


def to_oracle3(P):
    """Verona (5, NX, NY, NZ) Fortran array -> a fresh oracle (NZ, NY, NX, 5) C array
    (same byte order; always a copy, so in-place oracle runs leave P alone)."""
    return np.array(np.asfortranarray(P).T, order="C", copy=True)


def from_oracle3(A):
    return np.asfortranarray(A.T)


to_oracle2 = to_oracle3
from_oracle2 = from_oracle3
