"""Verona1D on one MI355X: the names and call conventions of src/1dim/.

    P = Verona1D.ParVector1D(N)                          # src/1dim/structs.jl:16-28
    out = Verona1D.HARM_HLL(P, N, dt, dx, T, eos, drops, "wenoz", n_iter, tol)

HARM_HLL (src/1dim/algos.jl:58-315) runs the tube on the GPU through the
batched-tube kernel (a batch of one) and returns the snapshot list like the
reference: P at t = 0 and after every step where t passes the next multiple
of drops.
"""
import numpy as np

from .tubes import Tubes


class ParVector1D:
    """SoA primitives without ghost cells: arr1 = rho, arr2 = u, arr3 = u^x."""

    def __init__(self, N):
        self.arr1 = np.zeros(N)
        self.arr2 = np.zeros(N)
        self.arr3 = np.zeros(N)
        self.size = N

    def copy(self):
        q = ParVector1D(self.size)
        q.arr1[:], q.arr2[:], q.arr3[:] = self.arr1, self.arr2, self.arr3
        return q


def _rec_name(rec):
    name = str(rec).lstrip(":").lower()
    if name not in ("minmod", "ppm", "wenoz"):
        raise ValueError(f"reconstruction not supported: {rec!r}")
    return name


def HARM_HLL(P, N, dt, dx, T, eos, drops, rec, *kwargs, riemann="hll"):
    n_iter = int(kwargs[0]) if len(kwargs) > 0 else 50
    tol = float(kwargs[1]) if len(kwargs) > 1 else 1e-10
    if N != P.size:
        raise ValueError("N must equal P.size")
    out = [P.copy()]
    # one workgroup as wide as the tube (up to 1024 lanes): every stage is a
    # single chunk with 4 waves per SIMD (1024-cell shock tube: 57 -> 23 ms
    # against 128 lanes, the same bits)
    block = min(1024, max(64, (N + 63) // 64 * 64))
    with Tubes([N], gamma=eos.gamma) as tb:
        tb.upload(P.arr1, P.arr2, P.arr3, [dt], [dx])
        tb.set_method(_rec_name(rec), riemann, n_iter, tol)
        t = 0.0
        thres = drops
        while t < T:
            # advance to the next snapshot (or the end) in one launch, with the
            # reference's sequential FP64 accumulation of t
            k = 0
            while t < T:
                t += dt
                k += 1
                if t > thres:
                    break
            tb.run(k, block=block, grid=1)
            if t > thres:
                r, u, ux = tb.download()
                q = ParVector1D(N)
                q.arr1[:], q.arr2[:], q.arr3[:] = r, u, ux
                out.append(q)
                thres += drops
        r, u, ux = tb.download()
    P.arr1[:], P.arr2[:], P.arr3[:] = r, u, ux
    return out
