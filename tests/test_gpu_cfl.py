"""The CFL time step always comes from the current state (GPU, through the C ABI).

vrn_step keeps the signal rate that a CFL step's last stage reduces for the
next step; a fixed-dt step, a new state written by another entry point, or new
grid spacings must not leave a stale rate behind.  Each case checks the dt of
the following CFL step against cfl / (max signal rate of the state it starts
from), the rate computed by the oracle.
"""
import numpy as np
import pytest

from helpers import progress, state3, to_oracle3

pytestmark = pytest.mark.gpu

G = 4.0 / 3.0
CFL = 0.4


def _expected_dt(oracle, P, dx, dy, dz):
    return CFL / oracle.max_rate3(to_oracle3(P), G, dx, dy, dz)


def _setup(dev, P, dx, dy, dz, cfl, dt=0.0):
    dev.set_step(ref_mode=False, rec="wenoz", riemann="hllc", integrator="rk3", n_iter=50, tol=1e-10,
                 floor=1e-7, cfl=cfl, dt=dt, dx=dx, dy=dy, dz=dz, bc=["periodic"] * 6)


def test_cfl_after_fixed_dt_step_uses_current_state(oracle):
    from verona_amd import DeviceGrid
    progress("test_cfl_after_fixed_dt_step_uses_current_state 24x20x16")
    nx, ny, nz = 24, 20, 16
    dx, dy, dz = 1 / nx, 1 / ny, 1 / nz
    P = state3(nx, ny, nz, seed=21, umax=0.6)
    with DeviceGrid(3, nx, ny, nz, G) as dev:
        dev.upload("P", P)
        dev.fill_ghosts(["periodic"] * 6)
        dev.prim_to_cons()
        _setup(dev, P, dx, dy, dz, CFL)
        dt1 = dev.step(want_dt=True)
        assert abs(dt1 - _expected_dt(oracle, P, dx, dy, dz)) <= 1e-14 * dt1
        # a fixed-dt step, then CFL again: dt from the state after the fixed step
        _setup(dev, P, dx, dy, dz, 0.0, dt=0.5 * dt1)
        dev.step()
        P2 = dev.download("P")
        _setup(dev, P, dx, dy, dz, CFL)
        dt3 = dev.step(want_dt=True)
        want = _expected_dt(oracle, P2, dx, dy, dz)
        progress(f"  dt after fixed step {dt3!r}, from the current state {want!r}")
        assert abs(dt3 - want) <= 1e-14 * want
        # a CFL step straight after another: the rate its last stage reduced
        P3 = dev.download("P")
        dt4 = dev.step(want_dt=True)
        want = _expected_dt(oracle, P3, dx, dy, dz)
        assert abs(dt4 - want) <= 1e-14 * want


def test_cfl_after_upload_and_new_spacing(oracle):
    from verona_amd import DeviceGrid
    progress("test_cfl_after_upload_and_new_spacing 24x20x16")
    nx, ny, nz = 24, 20, 16
    dx, dy, dz = 1 / nx, 1 / ny, 1 / nz
    P = state3(nx, ny, nz, seed=22, umax=0.3)
    Q = state3(nx, ny, nz, seed=23, umax=0.7)
    with DeviceGrid(3, nx, ny, nz, G) as dev:
        dev.upload("P", P)
        dev.fill_ghosts(["periodic"] * 6)
        dev.prim_to_cons()
        _setup(dev, P, dx, dy, dz, CFL)
        dev.step()
        # a new state replaces P: the next dt comes from it
        dev.upload("P", Q)
        dev.fill_ghosts(["periodic"] * 6)
        dev.prim_to_cons()
        dt = dev.step(want_dt=True)
        want = _expected_dt(oracle, Q, dx, dy, dz)
        progress(f"  dt after upload {dt!r}, want {want!r}")
        assert abs(dt - want) <= 1e-14 * want
        # new spacings between two CFL steps
        Pc = dev.download("P")
        _setup(dev, P, 2 * dx, dy, 0.5 * dz, CFL)
        dt = dev.step(want_dt=True)
        want = _expected_dt(oracle, Pc, 2 * dx, dy, 0.5 * dz)
        progress(f"  dt after new spacings {dt!r}, want {want!r}")
        assert abs(dt - want) <= 1e-14 * want
        # Limit writes P: the rate is recomputed from the limited state
        dev.limit(0.5)
        Pl = dev.download("P")
        dt = dev.step(want_dt=True)
        want = _expected_dt(oracle, Pl, 2 * dx, dy, 0.5 * dz)
        assert abs(dt - want) <= 1e-14 * want
