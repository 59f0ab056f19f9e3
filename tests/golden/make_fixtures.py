"""Write the golden fixtures under tests/golden/ from the CPU oracle.

Run after the oracle pins pass (tests/test_oracle_pins.py): the fixtures are
seeded inputs with the oracle's outputs, small enough to commit, and anchor
regressions of both the oracle and the HIP path (which must reproduce them).
    python tests/golden/make_fixtures.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")]

import pyoracle as O  # noqa: E402
from helpers import state2, state3, to_oracle2, to_oracle3  # noqa: E402

G = 4.0 / 3.0


def main():
    # 3D, Verona semantics: one WENOZ + HLL midpoint-RK2 step, periodic
    P = state3(12, 10, 8, seed=101)
    Po = to_oracle3(P)
    dx, dy, dz = 1 / 12, 1 / 10, 1 / 8
    dt = 0.4 / (1 / dx + 1 / dy + 1 / dz)
    m = O.make_mode(ref=True, gamma=G, floor=1e-7, dt=dt, dx=dx, dy=dy, dz=dz, n_iter=50, tol=1e-10)
    bc, _ = O.make_bc3([1] * 6)
    out = Po.copy()
    O.run3(out, m, bc, 1)
    np.savez_compressed(os.path.join(HERE, "step3d_ref.npz"), P0=P, P1=np.asfortranarray(out.T),
                        dt=dt, dx=[dx, dy, dz], gamma=G)
    # 3D extended: HLLC + SSP-RK3 + CFL, outflow
    out = Po.copy()
    m = O.make_mode(ref=False, riemann="hllc", integrator="rk3", gamma=G, floor=1e-7, cfl=0.4,
                    dx=dx, dy=dy, dz=dz, n_iter=50, tol=1e-10)
    bc, _ = O.make_bc3([2] * 6)
    _, t, _ = O.run3(out, m, bc, 2)
    np.savez_compressed(os.path.join(HERE, "step3d_ext.npz"), P0=P, P2=np.asfortranarray(out.T),
                        t=t, dx=[dx, dy, dz], gamma=G)
    # 2D, Verona semantics
    P2 = state2(16, 12, seed=102)
    P2o = to_oracle2(P2)
    dx2, dy2 = 1 / 16, 1 / 12
    dt2 = 0.4 / (1 / dx2 + 1 / dy2)
    m = O.make_mode(ref=True, gamma=G, floor=1e-7, dt=dt2, dx=dx2, dy=dy2, n_iter=50, tol=1e-10)
    O.run2(P2o, m, [1, 1, 1, 1], 1)
    np.savez_compressed(os.path.join(HERE, "step2d_ref.npz"), P0=P2, P1=np.asfortranarray(P2o.T),
                        dt=dt2, dx=[dx2, dy2], gamma=G)
    # con2prim batch: hard states (high Lorentz factor, 50x cold guesses)
    rng = np.random.default_rng(103)
    n = 256
    Pt = np.empty((n, 5))
    Pt[:, 0] = 10 ** rng.uniform(-4, 2, n)
    Pt[:, 1] = 10 ** rng.uniform(-6, 2, n) / (G - 1)
    lor = 10 ** rng.uniform(0, 1.6, n)
    d = rng.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    Pt[:, 2:] = np.sqrt(lor ** 2 - 1)[:, None] * d
    U = O.p2u3(Pt, G)
    guess = np.tile([1.0, 1.0, 0.0, 0.0, 0.0], (n, 1))
    Pr, bad, meth, its = O.c2p3_batch(U, guess, G, 50, 1e-10)
    np.savez_compressed(os.path.join(HERE, "c2p3_batch.npz"), U=U, guess=guess, P=Pr, method=meth,
                        iters=its, gamma=G)
    # 1D shock tube (Marti & Mueller problem 1) at N = 128 to t = 0.4, minmod + HLL
    g1 = 5.0 / 3.0
    N = 128
    P1 = np.zeros((3, N))
    P1[0, :64], P1[1, :64] = 10.0, 13.33 / (g1 - 1)
    P1[0, 64:], P1[1, 64:] = 1.0, 1e-6 / (g1 - 1)
    P10 = P1.copy()
    steps, t = O.run1(P1, 0.4 / N, 1 / N, 0.4, g1, "minmod", "hll", 50, 1e-10)
    np.savez_compressed(os.path.join(HERE, "tube1d.npz"), P0=P10, P=P1, steps=steps, t=t, gamma=g1)
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    main()
