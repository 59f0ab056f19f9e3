"""The oracle reproduces the committed golden fixtures bit for bit (CPU), and
the HIP path reproduces them within the FP64 tolerance (GPU)."""
import os

import numpy as np
import pytest

from helpers import from_oracle2, from_oracle3, interior, per_var_rel_l1, to_oracle2, to_oracle3

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


def test_oracle_step3d_ref(oracle):
    d = load("step3d_ref.npz")
    Po = to_oracle3(d["P0"])
    dx, dy, dz = d["dx"]
    m = oracle.make_mode(ref=True, gamma=float(d["gamma"]), floor=1e-7, dt=float(d["dt"]), dx=dx, dy=dy,
                         dz=dz, n_iter=50, tol=1e-10)
    bc, _ = oracle.make_bc3([1] * 6)
    oracle.run3(Po, m, bc, 1)
    assert np.array_equal(from_oracle3(Po), d["P1"])


def test_oracle_step3d_ext(oracle):
    d = load("step3d_ext.npz")
    Po = to_oracle3(d["P0"])
    dx, dy, dz = d["dx"]
    m = oracle.make_mode(ref=False, riemann="hllc", integrator="rk3", gamma=float(d["gamma"]), floor=1e-7,
                         cfl=0.4, dx=dx, dy=dy, dz=dz, n_iter=50, tol=1e-10)
    bc, _ = oracle.make_bc3([2] * 6)
    _, t, _ = oracle.run3(Po, m, bc, 2)
    assert np.array_equal(from_oracle3(Po), d["P2"]) and t == float(d["t"])


def test_oracle_step2d_ref(oracle):
    d = load("step2d_ref.npz")
    Po = to_oracle2(d["P0"])
    dx, dy = d["dx"]
    m = oracle.make_mode(ref=True, gamma=float(d["gamma"]), floor=1e-7, dt=float(d["dt"]), dx=dx, dy=dy,
                         n_iter=50, tol=1e-10)
    oracle.run2(Po, m, [1, 1, 1, 1], 1)
    assert np.array_equal(from_oracle2(Po), d["P1"])


def test_oracle_c2p_batch(oracle):
    d = load("c2p3_batch.npz")
    Pr, bad, meth, its = oracle.c2p3_batch(d["U"], d["guess"], float(d["gamma"]), 50, 1e-10)
    assert np.array_equal(Pr, d["P"], equal_nan=True)
    assert np.array_equal(meth, d["method"]) and np.array_equal(its, d["iters"])


def test_oracle_tube1d(oracle):
    d = load("tube1d.npz")
    P = np.ascontiguousarray(d["P0"].copy())
    N = P.shape[1]
    steps, t = oracle.run1(P, 0.4 / N, 1 / N, 0.4, float(d["gamma"]), "minmod", "hll", 50, 1e-10)
    assert steps == int(d["steps"]) and np.array_equal(P, d["P"])


@pytest.mark.gpu
def test_gpu_reproduces_golden_steps():
    from verona_amd import DeviceGrid
    d = load("step3d_ref.npz")
    P0 = d["P0"]
    nx, ny, nz = (s - 6 for s in P0.shape[1:])
    dx, dy, dz = d["dx"]
    with DeviceGrid(3, nx, ny, nz, float(d["gamma"])) as dev:
        dev.upload("P", P0)
        dev.fill_ghosts(["periodic"] * 6)
        dev.prim_to_cons()
        dev.set_step(ref_mode=True, dt=float(d["dt"]), dx=dx, dy=dy, dz=dz, bc=["periodic"] * 6,
                     n_iter=50, tol=1e-10, floor=1e-7)
        dev.step()
        P1 = dev.download("P")
    assert max(per_var_rel_l1(interior(P1), interior(d["P1"]))) <= 1e-12
    d = load("step2d_ref.npz")
    P0 = d["P0"]
    nx, ny = (s - 6 for s in P0.shape[1:])
    dx, dy = d["dx"]
    with DeviceGrid(2, nx, ny, 1, float(d["gamma"])) as dev:
        dev.upload("P", P0)
        dev.fill_ghosts(["periodic"] * 4 + ["none"] * 2)
        dev.prim_to_cons()
        dev.set_step(ref_mode=True, dt=float(d["dt"]), dx=dx, dy=dy, bc=["periodic"] * 4 + ["none"] * 2,
                     n_iter=50, tol=1e-10, floor=1e-7)
        dev.step()
        P1 = dev.download("P")
    assert max(per_var_rel_l1(interior(P1), interior(d["P1"]))) <= 1e-12
