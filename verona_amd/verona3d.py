"""Verona3D on one MI355X: the names and call conventions of src/3dim/ over libverona_hip.

    P = Verona3D.ParVector3D(Nx, Ny, Nz)          # src/3dim/structs.jl:28-41
    ... fill P.arr[v, i, j, k] (0-based, ghosts included) ...
    CuP = Verona3D.CuParVector3D(P)               # src/3dim/structs.jl:43-56
    Verona3D.HARM_HLL(comm, CuP, (1, 1, 1), (4, 4, 4), dt, dx, dy, dz, Tmax,
                      eos, drops, floor, out_dir, verbose, n_iter, tol, fun_bound)
"""
import numpy as np

from . import driver
from .device import DeviceGrid, host_array


class ParVector3D:
    """Host array arr[5, Nx+6, Ny+6, Nz+6] (Fortran order, Verona byte order)."""

    def __init__(self, Nx, Ny=None, Nz=None):
        if isinstance(Nx, (ParVector3D, CuParVector3D)):
            src = Nx
            self.arr = np.array(src.arr, order="F", copy=True)
        else:
            self.arr = host_array(3, Nx, Ny, Nz)
        self.size_X, self.size_Y, self.size_Z = self.arr.shape[1:]


class CuParVector3D:
    """Device-resident primitives: a DeviceGrid holding P (src/3dim/structs.jl:43-56)."""

    def __init__(self, P, Ny=None, Nz=None, gamma=4.0 / 3.0):
        if isinstance(P, ParVector3D):
            arr = P.arr
        elif isinstance(P, int):
            arr = host_array(3, P, Ny, Nz)
        else:
            arr = np.asarray(P)
        self.size_X, self.size_Y, self.size_Z = arr.shape[1:]
        self.dev = DeviceGrid(3, self.size_X - 6, self.size_Y - 6, self.size_Z - 6, gamma)
        self.dev.upload("P", arr)

    @property
    def arr(self):
        """Host copy of the device array (Array(P.arr) in the reference)."""
        return self.dev.download("P")


def VectorLike(X):
    """src/3dim/structs.jl:58-64"""
    if isinstance(X, CuParVector3D):
        return CuParVector3D(X.size_X - 6, X.size_Y - 6, X.size_Z - 6)
    return ParVector3D(X.size_X - 6, X.size_Y - 6, X.size_Z - 6)


def HARM_HLL(comm, P, MPI_dims, Size, dt, dx, dy, dz, Tmax, eos, drops, floor=1e-7, out_dir=".",
             verbose=False, *kwargs, **opts):
    """src/3dim/algos.jl:203-362.  kwargs = (n_iter, tol[, fun_bound]) as in the
    reference.  MPI_dims must be (1, 1, 1) (one GPU); Size (the KA workgroup)
    is accepted and ignored -- launch shapes are chosen by the library.
    Extra keyword options select the extended method (rec, riemann,
    integrator, cfl, precision, ref_mode) and max_steps."""
    if tuple(MPI_dims) != (1, 1, 1):
        raise ValueError("one GPU: MPI_dims must be (1, 1, 1)")
    n_iter = kwargs[0] if len(kwargs) > 0 else 50
    tol = kwargs[1] if len(kwargs) > 1 else 1e-10
    fun_bound = kwargs[2] if len(kwargs) > 2 else None
    if isinstance(P, ParVector3D):
        P = CuParVector3D(P, gamma=eos.gamma)
    if abs(P.dev.gamma - eos.gamma) > 0:
        P.dev.close()
        raise ValueError("CuParVector3D was created with a different gamma; pass gamma=eos.gamma")
    return driver.harm_hll(P.dev, 3, comm, None, dt, (dx, dy, dz), Tmax, eos, drops, floor, out_dir,
                           verbose, n_iter, tol, fun_bound, **opts)
