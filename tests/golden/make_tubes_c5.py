"""Write tests/golden/tubes_c5.npz: the oracle's result on the full batched-
tubes workload (4,096 tubes, lengths U{256..8192}, SplitMix64 seed 42, PPM +
HLLC, RK2, dt = 0.4 dx, 500 steps, gamma = 5/3, con2prim n_iter 50, tol 1e-10).

Verona1D's con2prim has no floors (src/1dim/structs.jl:78-105), so tubes
whose states it cannot recover carry NaN markers.  The fixture holds, per
tube: the number of NaN cells in each variable and, over each tube's cells,
the sum of |rho|, |u| and |u^x| with NaN cells left out (a checksum of the
full state at 500 steps).  The GPU test runs the same batch and must mark
exactly the same cells and reproduce the checksums.

Each tube is an independent oracle run; tubes run in threads (the C oracle
releases the GIL).  About 10 minutes on 8 cores.
    python tests/golden/make_tubes_c5.py
"""
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

os.environ.setdefault("OMP_NUM_THREADS", "1")

import numpy as np  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle")]

import pyoracle as O  # noqa: E402
from verona_amd.tubes import shock_tube_batch  # noqa: E402

NTUBES, SEED, STEPS = 4096, 42, 500
GAMMA = 5.0 / 3.0


def tube_summary(P):
    """(NaN cells per variable, sum of |value| over non-NaN cells per variable)."""
    nan = ~np.isfinite(P)
    return nan.sum(axis=1), np.where(nan, 0.0, np.abs(P)).sum(axis=1)


def main():
    lengths, rho, u, ux, dt, dx = shock_tube_batch(NTUBES, 256, 8192, seed=SEED)
    off = np.concatenate([[0], np.cumsum(lengths)])
    nan_cells = np.zeros((NTUBES, 3), np.int64)
    abs_sum = np.zeros((NTUBES, 3))

    def run(t):
        P = np.ascontiguousarray(np.stack([rho[off[t]:off[t + 1]], u[off[t]:off[t + 1]], ux[off[t]:off[t + 1]]]))
        O.run1(P, dt[t], dx[t], 1e30, GAMMA, "ppm", "hllc", 50, 1e-10, max_steps=STEPS)
        nan_cells[t], abs_sum[t] = tube_summary(P)

    O.lib()
    t0 = time.time()
    order = np.argsort(-lengths)  # longest first
    with ThreadPoolExecutor(max_workers=os.cpu_count() or 1) as ex:
        list(ex.map(run, order))
    nan_tubes = np.flatnonzero(nan_cells.sum(axis=1) > 0)
    np.savez_compressed(os.path.join(HERE, "tubes_c5.npz"), lengths=lengths, nan_cells=nan_cells,
                        abs_sum=abs_sum, nan_tubes=nan_tubes, steps=STEPS, seed=SEED, gamma=GAMMA)
    print(f"{NTUBES} tubes, {int(lengths.sum())} cells, {STEPS} steps in {time.time() - t0:.0f} s; "
          f"{len(nan_tubes)} tubes with NaN markers")


if __name__ == "__main__":
    main()
