"""Boundary specifications for the one-GPU solver.

The reference has no boundary-condition layer: periodicity comes from a
periodic MPI Cartesian communicator (src/3dim/sync_CUDA.jl:3-129, a
self-exchange on one rank) and everything else from a user callback
``fun_bound(P, t[, coords, dims])`` that mutates the ghost cells of P once per
step (src/3dim/algos.jl:275-277).  On one GPU the callback becomes a
declarative per-face rule evaluated by an on-device ghost kernel; a plain
Python callable is still accepted and bridged through host memory.
"""

FACES = ("xlo", "xhi", "ylo", "yhi", "zlo", "zhi")
KINDS = ("none", "periodic", "outflow", "nozzle")


class Cart:
    """Stand-in for the MPI Cartesian communicator of a one-rank run.

    Only its periodicity matters on one GPU: periodic axes get the halo
    self-exchange after every stage; non-periodic axes see MPI.PROC_NULL and
    keep their ghost cells.
    """

    def __init__(self, periodic=(False, False, False)):
        self.periodic = tuple(bool(p) for p in periodic)

    def Comm_rank(self):
        return 0

    def Cart_coords(self):
        return (0,) * len(self.periodic)


class Boundary:
    """Declarative fun_bound: a kind per face plus an optional jet nozzle.

    ``Boundary(outflow=("ylo", "yhi", "zlo", "zhi", "xhi"))`` reproduces the
    zero-gradient copies of examples/Jet3D/HARM_jet_breakout.jl:136-157.
    """

    def __init__(self, outflow=(), nozzle=None):
        for f in outflow:
            if f not in FACES:
                raise ValueError(f"unknown face {f!r}")
        self.outflow = tuple(outflow)
        self.nozzle = nozzle  # dict(radius, rho, u, umag, eps, phi_seq) or None

    def kinds(self, periodic):
        out = []
        for k, f in enumerate(FACES):
            axis = k // 2
            if axis < len(periodic) and periodic[axis]:
                out.append("periodic")
            elif f == "zlo" and self.nozzle is not None:
                out.append("nozzle")
            elif f in self.outflow:
                out.append("outflow")
            else:
                out.append("none")
        return out


def splitmix64(seed):
    """SplitMix64 stream (Steele, Lea & Flood 2014): yields uint64 values."""
    x = seed & 0xFFFFFFFFFFFFFFFF
    while True:
        x = (x + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        yield z ^ (z >> 31)


def uniform_stream(seed):
    """Doubles in [0, 1) from the top 53 bits of SplitMix64."""
    for z in splitmix64(seed):
        yield (z >> 11) * (1.0 / 9007199254740992.0)


def uniform_array(seed, count):
    """The first `count` values of uniform_stream(seed), vectorised."""
    import numpy as np
    k = np.arange(1, count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + k * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
