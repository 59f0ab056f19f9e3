"""The fast FP64 form against the exact form (which equals the oracle bit for
bit) over many steps: its per-step rounding differences grow with the flow,
and this pins how far they get.  Kelvin-Helmholtz 256^2 (periodic, 100 steps)
and the jet with its nozzle at 64^3 (60 steps), WENOZ + HLLC + SSP-RK3, each
form with its own CFL dt.  Measured: KH u^y 1.8e-12 after 100 steps (the
other variables ~1e-14); jet u^x, u^y ~1e-12 after 50 steps at 128^3.  The
bound 2e-11 leaves an order of magnitude of headroom; a lost Newton step or
a wrong reciprocal in the fast form shows up as 1e-8 or worse."""
import math

import numpy as np
import pytest

from helpers import per_var_rel_l1, progress

pytestmark = pytest.mark.gpu
G = 4.0 / 3.0


def _kh(prec, steps, n=256):
    from verona_amd import DeviceGrid
    from verona_amd.problems import kelvin_helmholtz
    dx = 1.0 / n
    bc = ["periodic"] * 4 + ["none", "none"]
    with DeviceGrid(2, n, n, 1, G) as dev:
        dev.upload("P", kelvin_helmholtz(n, gamma=G))
        dev.fill_ghosts(bc)
        dev.prim_to_cons()
        dev.set_step(ref_mode=False, rec="wenoz", riemann="hllc", integrator="rk3", precision=prec,
                     n_iter=50, tol=1e-10, floor=1e-7, cfl=0.4, dx=dx, dy=dx, bc=bc)
        t = sum(dev.step(want_dt=True) for _ in range(steps))
        return dev.download("P")[:, 3:-3, 3:-3], t, dev.counters()[0]


def _jet(prec, steps, n=64):
    from verona_amd import DeviceGrid, host_array
    from verona_amd.bc import uniform_stream
    P = host_array(3, n, n, n)
    P[0] = 1.0
    P[1] = 1e-4 / (G - 1)
    dx = 1 / n
    it = uniform_stream(42)
    phis = [2 * math.pi * next(it) for _ in range(steps)]
    noz = (8 * dx, 0.01, 1e-4 / (G - 1), math.sqrt(99.0), 0.01)
    bc = ["outflow"] * 4 + ["nozzle", "outflow"]
    with DeviceGrid(3, n, n, n, G) as dev:
        dev.upload("P", P)
        dev.set_nozzle(noz[0], dx, dx, *noz[1:])
        dev.fill_ghosts(bc)
        dev.prim_to_cons()
        dev.set_step(ref_mode=False, rec="wenoz", riemann="hllc", integrator="rk3", precision=prec,
                     n_iter=50, tol=1e-10, floor=1e-7, cfl=0.4, dx=dx, dy=dx, dz=dx, bc=bc)
        t = sum(dev.step(phi=phis[s], want_dt=True) for s in range(steps))
        return dev.download("P")[:, 3:-3, 3:-3, 3:-3], t, dev.counters()[0]


@pytest.mark.parametrize("problem,steps", [("kh256", 100), ("jet64", 60)])
def test_fast_form_error_growth_stays_at_rounding_level(problem, steps):
    progress(f"test_fast_form_error_growth_stays_at_rounding_level {problem} {steps} steps")
    run = _kh if problem == "kh256" else _jet
    E, te, ne = run("fp64", steps)
    F, tf, nf = run("fp64fast", steps)
    errs = per_var_rel_l1(F, E)
    progress(f"  rel L1 per variable {['%.1e' % e for e in errs]}, t {tf!r} vs {te!r}")
    assert ne == nf == 0
    assert max(errs) <= 2e-11
    assert abs(tf - te) <= 1e-12 * te
