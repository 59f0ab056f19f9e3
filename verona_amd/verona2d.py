"""Verona2D on one MI355X: the names and call conventions of src/2dim/ over libverona_hip."""
import numpy as np

from . import driver
from .device import DeviceGrid, host_array


class ParVector2D:
    """Host array arr[4, Nx+6, Ny+6] (src/2dim/structs.jl:26-38)."""

    def __init__(self, Nx, Ny=None):
        if isinstance(Nx, (ParVector2D, CuParVector2D)):
            self.arr = np.array(Nx.arr, order="F", copy=True)
        else:
            self.arr = host_array(2, Nx, Ny)
        self.size_X, self.size_Y = self.arr.shape[1:]


class CuParVector2D:
    """Device-resident primitives (src/2dim/structs.jl:40-52)."""

    def __init__(self, P, Ny=None, gamma=4.0 / 3.0):
        if isinstance(P, ParVector2D):
            arr = P.arr
        elif isinstance(P, int):
            arr = host_array(2, P, Ny)
        else:
            arr = np.asarray(P)
        self.size_X, self.size_Y = arr.shape[1:]
        self.dev = DeviceGrid(2, self.size_X - 6, self.size_Y - 6, 1, gamma)
        self.dev.upload("P", arr)

    @property
    def arr(self):
        return self.dev.download("P")


def HARM_HLL(comm, P, XMPI, YMPI, SizeX, SizeY, dt, dx, dy, Tmax, eos, drops, floor=1e-7,
             out_dir=".", *kwargs, **opts):
    """src/2dim/algos.jl:192-324 with kwargs = (n_iter, tol[, fun_bound])."""
    if (XMPI, YMPI) != (1, 1):
        raise ValueError("one GPU: XMPI = YMPI = 1")
    n_iter = kwargs[0] if len(kwargs) > 0 else 50
    tol = kwargs[1] if len(kwargs) > 1 else 1e-10
    fun_bound = kwargs[2] if len(kwargs) > 2 else None
    if isinstance(P, ParVector2D):
        P = CuParVector2D(P, gamma=eos.gamma)
    verbose = opts.pop("verbose", False)
    return driver.harm_hll(P.dev, 2, comm, None, dt, (dx, dy), Tmax, eos, drops, floor, out_dir,
                           verbose, n_iter, tol, fun_bound, **opts)
