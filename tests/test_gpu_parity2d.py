"""GPU parity of the 2D hot path (4 variables, 4x4 Newton con2prim) against the CPU oracle."""
import pytest

from helpers import from_oracle2, interior, per_var_rel_l1, progress, rel_l1, state2, to_oracle2

pytestmark = pytest.mark.gpu

G = 4.0 / 3.0
TOL = 1e-12


@pytest.fixture(scope="module")
def grid():
    from verona_amd import DeviceGrid
    dev = DeviceGrid(2, 72, 56, 1, G)
    yield dev
    dev.close()


@pytest.mark.parametrize("rec", ["wenoz", "ppm", "minmod"])
@pytest.mark.parametrize("riemann", ["hll", "hllc"])
def test_fluxes2(grid, oracle, rec, riemann):
    progress(f"test_fluxes2 rec={rec} riemann={riemann}")
    P = state2(72, 56, seed=1)
    grid.upload("P", P)
    Po = to_oracle2(P)
    for dim, slot in enumerate(["Fx", "Fy"]):
        grid.fluxes(dim, 1e-7, rec, riemann)
        F = grid.download(slot)
        Fo = from_oracle2(oracle.fluxes2(Po, G, 1e-7, dim, rec, riemann))
        sl = (slice(None), slice(2, -3), slice(2, -3))
        assert rel_l1(F[sl], Fo[sl]) <= TOL


def test_update_c2p2(grid, oracle):
    progress("test_update_c2p2")
    P = state2(72, 56, seed=2)
    Po = to_oracle2(P)
    dx, dy = 1 / 72, 1 / 56
    dt = 0.4 / (1 / dx + 1 / dy)
    grid.upload("P", P)
    grid.prim_to_cons()
    grid.fluxes(0, 1e-7)
    grid.fluxes(1, 1e-7)
    grid.update("Uhalf", dt / 2, dx, dy)
    grid.cons_to_prim("Uhalf", 50, 1e-10)
    grid.limit(1e-7)
    Pg = grid.download("P")
    Uo = oracle.ptou2(Po, G)
    Fx = oracle.fluxes2(Po, G, 1e-7, 0)
    Fy = oracle.fluxes2(Po, G, 1e-7, 1)
    Uho = oracle.update2(Uo, dt / 2, dx, dy, Fx, Fy)
    oracle.c2p2(Uho, Po, G, 50, 1e-10)
    oracle.limit2(Po, 1e-7)
    errs = per_var_rel_l1(Pg, from_oracle2(Po))
    progress(f"  per-var rel L1 {errs}")
    assert max(errs) <= TOL


@pytest.mark.parametrize("mode", ["ref", "ext"])
def test_fused_step2(oracle, mode):
    from verona_amd import DeviceGrid
    progress(f"test_fused_step2 {mode}")
    nx, ny = 72, 56
    P = state2(nx, ny, seed=3)
    dx, dy = 1 / nx, 1 / ny
    ref = mode == "ref"
    cfl = 0.0 if ref else 0.4
    dt = 0.4 / (1 / dx + 1 / dy) if ref else 0.0
    integ = "rk2" if ref else "rk3"
    riem = "hll" if ref else "hllc"
    bc = ["periodic"] * 4 + ["none", "none"]
    with DeviceGrid(2, nx, ny, 1, G) as dev:
        dev.upload("P", P)
        dev.fill_ghosts(bc)
        dev.prim_to_cons()
        dev.set_step(ref_mode=ref, rec="wenoz", riemann=riem, integrator=integ, n_iter=50,
                     tol=1e-10, floor=1e-7, dt=dt, cfl=cfl, dx=dx, dy=dy, bc=bc)
        for _ in range(3):
            dev.step()
        Pg = dev.download("P")
    Po = to_oracle2(P)
    m = oracle.make_mode(ref=ref, rec="wenoz", riemann=riem, integrator=integ, n_iter=50,
                         tol=1e-10, gamma=G, floor=1e-7, dt=dt, cfl=cfl, dx=dx, dy=dy)
    oracle.run2(Po, m, [1, 1, 1, 1], 3)
    errs = per_var_rel_l1(interior(Pg), interior(from_oracle2(Po)))
    progress(f"  per-var rel L1 {['%.2e' % e for e in errs]}")
    assert max(errs) <= TOL


def test_fast_fp64_form2(oracle):
    from verona_amd import DeviceGrid
    progress("test_fast_fp64_form2")
    nx, ny = 72, 56
    P = state2(nx, ny, seed=4)
    dx, dy = 1 / nx, 1 / ny
    dt = 0.4 / (1 / dx + 1 / dy)
    bc = ["periodic"] * 4 + ["none", "none"]
    with DeviceGrid(2, nx, ny, 1, G) as dev:
        dev.upload("P", P)
        dev.fill_ghosts(bc)
        dev.prim_to_cons()
        dev.set_step(ref_mode=True, precision="fp64fast", n_iter=50, tol=1e-10, floor=1e-7, dt=dt,
                     dx=dx, dy=dy, bc=bc)
        for _ in range(3):
            dev.step()
        Pg = dev.download("P")
    Po = to_oracle2(P)
    m = oracle.make_mode(ref=True, gamma=G, floor=1e-7, dt=dt, dx=dx, dy=dy, n_iter=50, tol=1e-10)
    oracle.run2(Po, m, [1, 1, 1, 1], 3)
    errs = per_var_rel_l1(interior(Pg), interior(from_oracle2(Po)))
    progress(f"  per-var rel L1 {['%.2e' % e for e in errs]}")
    assert max(errs) <= TOL


@pytest.mark.parametrize("precision", ["fp64", "fp64fast", "fp32"])
def test_kelvin_helmholtz(oracle, precision):
    """The 2D Kelvin-Helmholtz problem (periodic, WENOZ + HLLC + SSP-RK3, CFL
    0.4) at 128^2: parity with the FP64 oracle after 20 steps, to the FP64
    tolerance in both FP64 forms (the fast one solves the 4x4 Newton system
    in closed form); the FP32 reconstruction/Riemann form's error against it
    is reported and bounded."""
    from verona_amd import DeviceGrid
    from verona_amd.problems import kelvin_helmholtz
    progress(f"test_kelvin_helmholtz {precision}")
    n, steps = 128, 20
    P = kelvin_helmholtz(n, gamma=G)
    dx = 1.0 / n
    bc = ["periodic"] * 4 + ["none", "none"]
    with DeviceGrid(2, n, n, 1, G) as dev:
        dev.upload("P", P)
        dev.fill_ghosts(bc)
        dev.prim_to_cons()
        dev.set_step(ref_mode=False, rec="wenoz", riemann="hllc", integrator="rk3", precision=precision,
                     n_iter=50, tol=1e-10, floor=1e-7, cfl=0.4, dx=dx, dy=dx, bc=bc)
        for _ in range(steps):
            dev.step()
        Pg = dev.download("P")
        assert dev.counters()[0] == 0
    Po = to_oracle2(P)
    m = oracle.make_mode(ref=False, rec="wenoz", riemann="hllc", integrator="rk3", n_iter=50, tol=1e-10,
                         gamma=G, floor=1e-7, cfl=0.4, dx=dx, dy=dx)
    bad, t, _ = oracle.run2(Po, m, [1, 1, 1, 1], steps)
    assert bad == 0 and t > 0
    errs = per_var_rel_l1(interior(Pg), interior(from_oracle2(Po)))
    progress(f"  per-var rel L1 vs FP64 oracle {['%.2e' % e for e in errs]}")
    if precision != "fp32":
        assert max(errs) <= TOL
    else:
        # u^y is the 1 % transverse perturbation, so its relative error is
        # the largest (1.0e-5 measured on MI355X)
        assert max(errs) <= 1e-4


def ragged_state2(stripe):
    """37 x 53 interior (not a multiple of the wave width); with `stripe`,
    x columns 8..12 of the interior hold rho = 4e-7, u = 3e-7, a few times
    the 1e-7 floor, so the dense gas expands into it."""
    import numpy as np
    P = state2(37, 53, seed=9)
    if stripe:
        x = np.arange(P.shape[1])[:, None]
        low = (x > 10) & (x < 16)
        P[0] = np.where(low, 4e-7, P[0])
        P[1] = np.where(low, 3e-7, P[1])
    return P


@pytest.mark.parametrize("precision,stripe", [("fp64", True), ("fp64fast", False)])
def test_ragged_grid_outflow2(oracle, precision, stripe):
    """2D stage end (y sweep + pointwise RK combine and 4x4 Newton) on a
    ragged grid with outflow x faces and periodic y, 3 steps of WENOZ + HLLC +
    SSP-RK3.  The exact form on the near-vacuum stripe: every finite value
    bit-identical to the oracle, the same cells left NaN (the expanding front
    drives some Newton iterates to |u| > 1e154, where sqrt(1 + |u|^2) must be
    +inf) and the same count of singular systems.  The fast form is checked
    without the stripe: the front is ill-conditioned (a 1e-16 relative input
    perturbation moves the oracle's u^x by up to 1e-3 in relative L1 after
    one step, test_oracle_pins.py), so no rounding-level variant of the
    algorithm can match it to 1e-12 there."""
    import numpy as np
    from verona_amd import DeviceGrid
    progress(f"test_ragged_grid_outflow2 {precision} stripe={stripe}")
    nx, ny = 37, 53
    P = ragged_state2(stripe)
    dx, dy = 1 / nx, 1 / ny
    bc = ["outflow", "outflow", "periodic", "periodic", "none", "none"]
    with DeviceGrid(2, nx, ny, 1, G) as dev:
        dev.upload("P", P)
        dev.fill_ghosts(bc)
        dev.prim_to_cons()
        dev.set_step(ref_mode=False, rec="wenoz", riemann="hllc", integrator="rk3", precision=precision,
                     n_iter=50, tol=1e-10, floor=1e-7, cfl=0.4, dx=dx, dy=dy, bc=bc)
        for _ in range(3):
            dev.step()
        Pg = dev.download("P")
        nb = dev.counters()[0]
    Po = to_oracle2(P)
    m = oracle.make_mode(ref=False, rec="wenoz", riemann="hllc", integrator="rk3", n_iter=50, tol=1e-10,
                         gamma=G, floor=1e-7, cfl=0.4, dx=dx, dy=dy)
    nbo = oracle.run2(Po, m, [2, 2, 1, 1], 3)[0]
    A, B = interior(Pg), interior(from_oracle2(Po))
    ok = np.isfinite(B)
    errs = [np.abs(A[v][ok[v]] - B[v][ok[v]]).sum() / np.abs(B[v][ok[v]]).sum() for v in range(4)]
    progress(f"  singular {nb} (oracle {nbo}), NaN cells {int(np.isnan(A[0]).sum())} "
             f"(oracle {int(np.isnan(B[0]).sum())}), per-var rel L1 {['%.2e' % e for e in errs]}")
    assert nb == nbo
    assert np.array_equal(np.isnan(A), np.isnan(B))
    okc = ok.all(axis=0)
    assert (A[0][okc] >= 1e-7).all() and (A[1][okc] >= 1e-7).all()
    if precision == "fp64":
        assert np.array_equal(A[ok], B[ok])
    assert max(errs) <= TOL
