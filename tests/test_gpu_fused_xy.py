"""The 3D FP64 forms' x direction inside the y march (the XY march) against
the same form with a separate x sweep: every face the x sweep
solves is solved by the march (interior faces in-wave, the faces at its
64-column wave edges in batches of 8 rows, one edge cell per lane) with the
same operations, the fast form forms the stage's RK combination a U0 + b U_b
before the x update in both (the exact form in the z sweep in both), and the
updates are applied in the same order, so the two give the same bits.
Grids with ragged x extents cover partial last waves, a last wave of one
active lane and extents that are multiples of 64; the oracle comparison of
the fused form itself is in test_gpu_parity3d.py (the default of both FP64
forms)."""
import numpy as np
import pytest

from helpers import interior, progress, state3

pytestmark = pytest.mark.gpu
G = 4.0 / 3.0


def _run(P, fused, steps, *, rec="wenoz", riemann="hllc", integrator="rk3", ref=False, bc=None,
         precision="fp64fast"):
    from verona_amd import DeviceGrid
    nx, ny, nz = (s - 6 for s in P.shape[1:])
    dx, dy, dz = 1 / nx, 1 / ny, 1 / nz
    bc = bc or ["outflow", "outflow", "periodic", "periodic", "outflow", "outflow"]
    with DeviceGrid(3, nx, ny, nz, G) as dev:
        dev.set_xy_fusion(fused)
        dev.upload("P", P)
        dev.fill_ghosts(bc)
        dev.prim_to_cons()
        if ref:
            dt = 0.4 / (1 / dx + 1 / dy + 1 / dz)
            dev.set_step(ref_mode=True, rec=rec, riemann=riemann, integrator=integrator, precision=precision,
                         n_iter=50, tol=1e-10, floor=1e-7, dt=dt, dx=dx, dy=dy, dz=dz, bc=bc)
        else:
            dev.set_step(ref_mode=False, rec=rec, riemann=riemann, integrator=integrator, precision=precision,
                         n_iter=50, tol=1e-10, floor=1e-7, cfl=0.4, dx=dx, dy=dy, dz=dz, bc=bc)
        ts = [dev.step(want_dt=not ref) for _ in range(steps)]
        out = dev.download("P")
        nb = dev.counters()[0]
    return out, ts, nb


@pytest.mark.parametrize("nx,ny,nz,precision", [(40, 24, 20, "fp64fast"), (65, 20, 18, "fp64fast"),
                                                (128, 16, 12, "fp64fast"), (70, 33, 17, "fp64fast"),
                                                (64, 9, 11, "fp64fast"), (70, 33, 17, "fp64"),
                                                (40, 24, 20, "fp64")])
def test_fused_xy_equals_separate_x_sweep(nx, ny, nz, precision):
    progress(f"test_fused_xy_equals_separate_x_sweep {nx}x{ny}x{nz} {precision}")
    P = state3(nx, ny, nz, seed=21)
    A, ta, na = _run(P, True, 3, precision=precision)
    B, tb, nbb = _run(P, False, 3, precision=precision)
    diff = np.abs(interior(A) - interior(B)).max()
    progress(f"  max |fused - separate| {diff:.3e}, dt {ta[-1]:.6e} / {tb[-1]:.6e}")
    assert na == nbb == 0
    assert ta == tb
    assert np.array_equal(interior(A), interior(B))


@pytest.mark.parametrize("rec,riemann", [("minmod", "hll"), ("ppm", "hllc"), ("wenoz", "hll")])
def test_fused_xy_all_methods(rec, riemann):
    progress(f"test_fused_xy_all_methods {rec} {riemann}")
    P = state3(70, 20, 16, seed=22)
    A, _, _ = _run(P, True, 2, rec=rec, riemann=riemann, integrator="rk2", ref=True,
                   bc=["periodic"] * 6)
    B, _, _ = _run(P, False, 2, rec=rec, riemann=riemann, integrator="rk2", ref=True,
                   bc=["periodic"] * 6)
    assert np.array_equal(interior(A), interior(B))


def _jet(fused_seq, steps=6, n=40, nzc=32):
    """The jet workload at small size (nozzle inflow on z-low with a per-step
    precession angle, outflow elsewhere, HLLC + SSP-RK3 + CFL dt); step s runs
    with the x-y fusion on or off as fused_seq[s] says.  Returns P with its
    ghost cells and the dt of every step."""
    import math

    from verona_amd import DeviceGrid, host_array
    from verona_amd.bc import uniform_stream
    P = host_array(3, n, n, nzc)
    P[0] = 1.0
    P[1] = 1e-4 / (G - 1)
    dx = 1 / n
    it = uniform_stream(42)
    phis = [2 * math.pi * next(it) for _ in range(steps)]
    noz = (8 * dx, 0.01, 1e-4 / (G - 1), math.sqrt(99.0), 0.01)
    bc = ["outflow"] * 4 + ["nozzle", "outflow"]
    with DeviceGrid(3, n, n, nzc, G) as dev:
        dev.upload("P", P)
        dev.set_nozzle(noz[0], dx, dx, *noz[1:])
        dev.fill_ghosts(bc)
        dev.prim_to_cons()
        dev.set_step(ref_mode=False, rec="wenoz", riemann="hllc", integrator="rk3", precision="fp64fast",
                     n_iter=50, tol=1e-10, floor=1e-7, cfl=0.4, dx=dx, dy=dx, dz=dx, bc=bc)
        ts = []
        for s in range(steps):
            dev.set_xy_fusion(fused_seq[s])
            ts.append(dev.step(phi=phis[s], want_dt=True))
        out = dev.download("P")
        nb = dev.counters()[0]
    return out, ts, nb


def test_fused_xy_jet_nozzle_ghost_cells_read_through_the_rules():
    """With the fused march and rules on both x faces the stage-end ghost
    passes leave the x faces unfilled, across steps too: the fused march reads
    ghost columns through the x rules.  On the jet (nozzle + outflow) this
    equals the separate sweeps with complete ghost passes, bit for bit, ghost
    cells included (the download fills the x faces first); so does switching
    to the separate sweeps half-way, which fills them before its first stage."""
    progress("test_fused_xy_jet_nozzle_ghost_cells_read_through_the_rules 40x40x32")
    A, ta, na = _jet([True] * 6)
    B, tb, nbb = _jet([False] * 6)
    C, tc, nc = _jet([True, True, True, False, False, True])
    progress(f"  max |fused - separate| {np.abs(A - B).max():.3e}, max u_z {A[4].max():.3f}")
    assert na == nbb == nc == 0
    assert ta == tb == tc
    assert np.array_equal(A, B) and np.array_equal(C, B)
    assert A[4, 3:-3, 3:-3, 3].max() > 0.5  # the jet has entered the grid


@pytest.mark.parametrize("entry", ["fluxes", "copy", "fill_ghosts", "limit"])
def test_entry_points_after_fused_steps_see_complete_ghosts(entry):
    """After fused fast-form steps that left P's x faces unfilled, each entry
    point that reads P's ghost cells fills them first: its result equals the
    same call after the separate sweeps (which fill every face), bit for bit."""
    progress(f"test_entry_points_after_fused_steps_see_complete_ghosts {entry}")
    from verona_amd import DeviceGrid
    P = state3(70, 20, 16, seed=23)
    nx, ny, nz = 70, 20, 16
    bc = ["outflow", "outflow", "periodic", "periodic", "outflow", "outflow"]
    out = []
    for fused in (True, False):
        with DeviceGrid(3, nx, ny, nz, G) as dev:
            dev.set_xy_fusion(fused)
            dev.upload("P", P)
            dev.fill_ghosts(bc)
            dev.prim_to_cons()
            dev.set_step(ref_mode=False, rec="wenoz", riemann="hllc", integrator="rk3", precision="fp64fast",
                         n_iter=50, tol=1e-10, floor=1e-7, cfl=0.4, dx=1 / nx, dy=1 / ny, dz=1 / nz, bc=bc)
            for _ in range(2):
                dev.step()
            if entry == "fluxes":
                dev.fluxes(0, 1e-7)
                out.append(dev.download("Fx"))
            elif entry == "copy":
                dev.copy("Uhalf", "P")
                out.append(dev.download("Uhalf"))
            elif entry == "fill_ghosts":
                dev.fill_ghosts(["none"] * 2 + bc[2:])  # x faces kept as they are
                out.append(dev.download("P"))
            else:
                dev.limit(1e-7)
                out.append(dev.download("P"))
    assert np.array_equal(out[0], out[1])
