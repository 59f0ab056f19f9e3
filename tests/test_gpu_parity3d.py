"""GPU parity of the 3D hot path against the CPU oracle (through the C ABI).

Bar: relative L1 per variable <= 1e-12 after one step (the FP64 tolerance of
the target), with bit-exact agreement reported; the reference-boundary kernels
are compared array by array.
"""
import numpy as np
import pytest

from helpers import (from_oracle3, interior, per_var_rel_l1, progress, rel_l1, state3,
                     to_oracle3)

pytestmark = pytest.mark.gpu

G = 4.0 / 3.0
TOL = 1e-12


@pytest.fixture(scope="module")
def grid():
    from verona_amd import DeviceGrid
    nx, ny, nz = 40, 24, 20
    dev = DeviceGrid(3, nx, ny, nz, G)
    yield dev
    dev.close()


def _max_rel(a, b):
    den = np.maximum(np.abs(b), 1e-300)
    return float(np.max(np.abs(a - b) / den))


def test_ptou_matches_oracle(grid, oracle):
    progress("test_ptou_matches_oracle 40x24x20")
    P = state3(40, 24, 20, seed=1)
    grid.upload("P", P)
    grid.prim_to_cons()
    U = grid.download("U")
    Uo = from_oracle3(oracle.ptou3(to_oracle3(P), G))
    assert np.array_equal(U, Uo) or _max_rel(U, Uo) < 1e-15


@pytest.mark.parametrize("rec", ["wenoz", "ppm", "minmod"])
@pytest.mark.parametrize("riemann", ["hll", "hllc"])
def test_fluxes_match_oracle(grid, oracle, rec, riemann):
    progress(f"test_fluxes_match_oracle rec={rec} riemann={riemann}")
    P = state3(40, 24, 20, seed=2)
    grid.upload("P", P)
    Po = to_oracle3(P)
    for dim, slot in enumerate(["Fx", "Fy", "Fz"]):
        grid.upload(slot, np.zeros_like(P))
        grid.fluxes(dim, 1e-7, rec, riemann)
        F = grid.download(slot)
        Fo = from_oracle3(oracle.fluxes3(Po, G, 1e-7, dim, rec, riemann))
        sl = (slice(None), slice(2, -3), slice(2, -3), slice(2, -3))
        a, b = F[sl], Fo[sl]
        assert np.all(np.isfinite(a))
        nexact = np.mean(a == b)
        progress(f"  dim={dim}: bit-exact fraction {nexact:.4f}, max rel {_max_rel(a, b):.2e}")
        assert rel_l1(a, b) <= TOL


def test_update_c2p_limit_match_oracle(grid, oracle):
    progress("test_update_c2p_limit_match_oracle")
    P = state3(40, 24, 20, seed=3)
    Po = to_oracle3(P)
    dx, dy, dz = 1 / 40, 1 / 24, 1 / 20
    dt = 0.4 / (1 / dx + 1 / dy + 1 / dz)
    grid.upload("P", P)
    grid.prim_to_cons()
    for d in range(3):
        grid.fluxes(d, 1e-7, "wenoz", "hll")
    grid.upload("Uhalf", np.zeros_like(P))
    grid.update("Uhalf", dt / 2, dx, dy, dz)
    Uh = grid.download("Uhalf")
    Uo = oracle.ptou3(Po, G)
    F = [oracle.fluxes3(Po, G, 1e-7, d, "wenoz", "hll") for d in range(3)]
    Uho = oracle.update3(Uo, dt / 2, dx, dy, dz, *F)
    assert rel_l1(interior(Uh), interior(from_oracle3(Uho))) <= TOL
    grid.cons_to_prim("Uhalf", 50, 1e-10)
    grid.limit(1e-7)
    Pg = grid.download("P")
    Pc = Po.copy()
    oracle.c2p3(Uho, Pc, G, 50, 1e-10)
    oracle.limit3(Pc, 1e-7)
    Pc = from_oracle3(Pc)
    errs = per_var_rel_l1(Pg, Pc)
    progress(f"  c2p per-var rel L1 {errs}, bit-exact {np.mean(Pg == Pc):.4f}")
    assert max(errs) <= TOL


def test_c2p_cascade_and_failure(grid, oracle):
    """Hard states: high Lorentz factors with a poor guess drive 1DW to the
    cascade; a superluminal state has no root and must come back NaN-marked
    and counted (src/3dim/structs.jl:613-619)."""
    progress("test_c2p_cascade_and_failure")
    rng = np.random.default_rng(7)
    P = state3(40, 24, 20, seed=4, umax=0.3)
    Pt = P.copy()
    # true states with large Lorentz factors in a slab
    Pt[2, 3:-3, 3:10, 3:-3] = rng.uniform(5, 40, size=Pt[2, 3:-3, 3:10, 3:-3].shape)
    Pt[1, 3:-3, 3:10, 3:-3] *= 1e-3
    U = from_oracle3(oracle.ptou3(to_oracle3(Pt), G))
    # one superluminal cell: |S| > tau + D has no physical root
    U[:, 10, 12, 10] = [1.0, -1.1, 5.0, 0.0, 0.0]
    grid.upload("P", P)
    grid.upload("Uhalf", U)
    grid.reset_counters()
    grid.cons_to_prim("Uhalf", 5, 1e-12)
    nbad, nresc = grid.counters()
    Pg = grid.download("P")
    Pc = to_oracle3(P)
    nbad_o, counts = oracle.c2p3(to_oracle3(U), Pc, G, 5, 1e-12)
    Pc = from_oracle3(Pc)
    progress(f"  gpu: unrecovered {nbad} rescued {nresc}; oracle counts {counts}")
    assert nbad == nbad_o >= 1
    assert nresc == sum(counts) - counts[1]
    assert np.all(np.isnan(Pg[:, 10, 12, 10]))
    ok = ~np.isnan(Pc)
    assert np.array_equal(np.isnan(Pg), np.isnan(Pc))
    assert rel_l1(Pg[ok], Pc[ok]) <= TOL


def test_c2p_forced_failure_niter0(grid, oracle):
    """n_iter = 0 makes every method of the cascade fall through, so every
    interior cell gets the NaN marker (src/3dim/structs.jl:613-619)."""
    progress("test_c2p_forced_failure_niter0")
    P = state3(40, 24, 20, seed=12)
    U = from_oracle3(oracle.ptou3(to_oracle3(P), G))
    grid.upload("P", P)
    grid.upload("Uhalf", U)
    grid.reset_counters()
    grid.cons_to_prim("Uhalf", 0, 1e-10)
    nbad, _ = grid.counters()
    Pg = grid.download("P")
    assert nbad == 40 * 24 * 20
    assert np.all(np.isnan(interior(Pg)))
    assert np.array_equal(Pg[:, :3], P[:, :3])


@pytest.mark.parametrize("kind", ["periodic", "outflow"])
def test_ghost_fill(grid, oracle, kind):
    progress(f"test_ghost_fill {kind}")
    P = state3(40, 24, 20, seed=5)
    grid.upload("P", P)
    grid.fill_ghosts([kind] * 6)
    Pg = grid.download("P")
    Po = to_oracle3(P)
    if kind == "periodic":
        oracle.fill_periodic3(Po, 7)
    else:
        oracle.fill_outflow3(Po, 7, 7)
    assert np.array_equal(Pg, from_oracle3(Po))


def _run_both(oracle, P, nsteps, *, ref=True, rec="wenoz", riemann="hll", integrator="rk2",
              cfl=0.0, bc=("periodic",) * 6, precision="fp64", n_iter=50, tol=1e-10):
    from verona_amd import DeviceGrid
    nx, ny, nz = P.shape[1] - 6, P.shape[2] - 6, P.shape[3] - 6
    dx, dy, dz = 1 / nx, 1 / ny, 1 / nz
    dt = 0.0 if cfl > 0 else 0.4 / (1 / dx + 1 / dy + 1 / dz)
    kinds = {"none": 0, "periodic": 1, "outflow": 2}
    with DeviceGrid(3, nx, ny, nz, G) as dev:
        dev.upload("P", P)
        dev.fill_ghosts([b if b == "periodic" else "none" for b in bc])
        if not ref:
            dev.fill_ghosts(bc)
        dev.prim_to_cons()
        dev.set_step(ref_mode=ref, rec=rec, riemann=riemann, integrator=integrator,
                     precision=precision, n_iter=n_iter, tol=tol, floor=1e-7, dt=dt, cfl=cfl,
                     dx=dx, dy=dy, dz=dz, bc=bc)
        tg = 0.0
        for _ in range(nsteps):
            d = dev.step(want_dt=True)
            tg += d
        Pg = dev.download("P")
        nbad, _ = dev.counters()
    Po = to_oracle3(P)
    mode = oracle.make_mode(ref=ref, rec=rec, riemann=riemann, integrator=integrator, n_iter=n_iter,
                            tol=tol, gamma=G, floor=1e-7, dt=dt, cfl=cfl, dx=dx, dy=dy, dz=dz)
    bco, _keep = oracle.make_bc3([kinds[b] for b in bc])
    nbad_o, to, _ = oracle.run3(Po, mode, bco, nsteps)
    return Pg, from_oracle3(Po), tg, to, nbad, nbad_o


@pytest.mark.parametrize("nsteps", [1, 4])
def test_fused_step_ref_mode(oracle, nsteps):
    progress(f"test_fused_step_ref_mode nsteps={nsteps}")
    P = state3(40, 24, 20, seed=6)
    Pg, Po, tg, to, nb, nbo = _run_both(oracle, P, nsteps)
    errs = per_var_rel_l1(interior(Pg), interior(Po))
    progress(f"  per-var rel L1 {['%.2e' % e for e in errs]}, bit-exact {np.mean(Pg == Po):.4f}")
    assert nb == nbo == 0
    assert max(errs) <= TOL
    assert tg == to


def test_fused_step_outflow_ref(oracle):
    progress("test_fused_step_outflow_ref")
    P = state3(32, 24, 20, seed=8)
    Pg, Po, *_ = _run_both(oracle, P, 3, bc=("outflow",) * 6)
    errs = per_var_rel_l1(interior(Pg), interior(Po))
    progress(f"  per-var rel L1 {['%.2e' % e for e in errs]}")
    assert max(errs) <= TOL


@pytest.mark.parametrize("rec", ["wenoz", "ppm"])
def test_fused_step_ext_mode(oracle, rec):
    """HLLC + SSP-RK3 + CFL dt, boundaries after every stage."""
    progress(f"test_fused_step_ext_mode rec={rec}")
    P = state3(40, 24, 20, seed=9)
    Pg, Po, tg, to, nb, nbo = _run_both(oracle, P, 3, ref=False, rec=rec, riemann="hllc",
                                        integrator="rk3", cfl=0.4,
                                        bc=("outflow", "outflow", "periodic", "periodic",
                                            "outflow", "outflow"))
    errs = per_var_rel_l1(interior(Pg), interior(Po))
    progress(f"  per-var rel L1 {['%.2e' % e for e in errs]}, t gpu {tg!r} oracle {to!r}")
    assert max(errs) <= TOL
    assert abs(tg - to) <= 1e-14 * to


def test_fp32_form_error_is_bounded(oracle):
    """FP32 reconstruction + Riemann, FP64 con2prim and accumulation: the error
    against the all-FP64 oracle is reported and bounded (not a parity claim)."""
    progress("test_fp32_form_error_is_bounded")
    P = state3(40, 24, 20, seed=10)
    Pg, Po, *_ = _run_both(oracle, P, 2, precision="fp32")
    errs = per_var_rel_l1(interior(Pg), interior(Po))
    progress(f"  fp32 per-var rel L1 vs fp64 oracle {['%.2e' % e for e in errs]}")
    assert max(errs) < 1e-5


@pytest.mark.parametrize("precision", ["fp64", "fp64fast"])
def test_jet_nozzle_cfl_ext_mode(oracle, precision):
    """The jet workload at small size: nozzle inflow on z-low with a per-step
    precession angle, outflow elsewhere, HLLC + SSP-RK3 + CFL dt, in both FP64
    forms (fp64fast is the form of the benchmark line): relative L1 per
    variable <= 1e-12 after 6 steps and t within 1e-13 of the oracle's."""
    import math

    from verona_amd import DeviceGrid, host_array
    from verona_amd.bc import uniform_stream
    progress(f"test_jet_nozzle_cfl_ext_mode 40x40x32 {precision}")
    n, nzc = 40, 32
    P = host_array(3, n, n, nzc)
    P[0] = 1.0
    P[1] = 1e-4 / (G - 1)
    dx = 1 / n
    it = uniform_stream(42)
    phis = [2 * math.pi * next(it) for _ in range(8)]
    noz = (8 * dx, 0.01, 1e-4 / (G - 1), math.sqrt(99.0), 0.01)
    bc = ["outflow"] * 4 + ["nozzle", "outflow"]
    with DeviceGrid(3, n, n, nzc, G) as dev:
        dev.upload("P", P)
        dev.set_nozzle(noz[0], dx, dx, *noz[1:])
        dev.fill_ghosts(bc)
        dev.prim_to_cons()
        dev.set_step(ref_mode=False, rec="wenoz", riemann="hllc", integrator="rk3", n_iter=50,
                     tol=1e-10, floor=1e-7, cfl=0.4, dx=dx, dy=dx, dz=dx, bc=bc, precision=precision)
        tg = sum(dev.step(phi=phis[s], want_dt=True) for s in range(6))
        Pg = dev.download("P")
        nbad, _ = dev.counters()
    Po = to_oracle3(P)
    mode = oracle.make_mode(ref=False, rec="wenoz", riemann="hllc", integrator="rk3", n_iter=50,
                            tol=1e-10, gamma=G, floor=1e-7, cfl=0.4, dx=dx, dy=dx, dz=dx)
    bco, keep = oracle.make_bc3([2, 2, 2, 2, 3, 2], dx=dx, dy=dx, nozzle=noz + (0.0,), phi_seq=phis)
    nbo, to, _ = oracle.run3(Po, mode, bco, 6)
    Po = from_oracle3(Po)
    errs = per_var_rel_l1(interior(Pg), interior(Po))
    progress(f"  per-var rel L1 {['%.2e' % e for e in errs]}, t {tg!r} vs {to!r}, max u_z {Pg[4].max():.3f}")
    assert nbad == nbo == 0
    assert max(errs) <= TOL
    assert abs(tg - to) <= (1e-14 if precision == "fp64" else 1e-13) * to
    assert Pg[4, 3:-3, 3:-3, 3].max() > 0.5  # the jet has entered the grid


def test_fp32_form_on_jet_has_no_unrecovered_cells(oracle):
    """The FP32 reconstruction/Riemann form on the jet workload (a flat medium
    with a relativistic inflow front, where FP32 WENO-Z weights would overflow
    without rescaling): no unrecovered cells, and its L1 error against the
    all-FP64 oracle is reported."""
    import math

    from verona_amd import DeviceGrid, host_array
    from verona_amd.bc import uniform_stream
    progress("test_fp32_form_on_jet_has_no_unrecovered_cells 32^3")
    n = 32
    P = host_array(3, n, n, n)
    P[0] = 1.0
    P[1] = 1e-4 / (G - 1)
    dx = 1 / n
    it = uniform_stream(42)
    phis = [2 * math.pi * next(it) for _ in range(10)]
    noz = (8 * dx, 0.01, 1e-4 / (G - 1), math.sqrt(99.0), 0.01)
    bc = ["outflow"] * 4 + ["nozzle", "outflow"]
    out = {}
    for prec in ("fp64", "fp32"):
        with DeviceGrid(3, n, n, n, G) as dev:
            dev.upload("P", P)
            dev.set_nozzle(noz[0], dx, dx, *noz[1:])
            dev.fill_ghosts(bc)
            dev.prim_to_cons()
            dev.set_step(ref_mode=False, rec="wenoz", riemann="hllc", integrator="rk3", precision=prec,
                         n_iter=50, tol=1e-10, floor=1e-7, cfl=0.4, dx=dx, dy=dx, dz=dx, bc=bc)
            for s in range(10):
                dev.step(phi=phis[s])
            out[prec] = (dev.download("P"), dev.counters()[0])
    assert out["fp32"][1] == 0 and out["fp64"][1] == 0
    errs = per_var_rel_l1(interior(out["fp32"][0]), interior(out["fp64"][0]))
    progress(f"  fp32 vs fp64 per-var rel L1 after 10 steps: {['%.2e' % e for e in errs]}")
    assert max(errs) < 1e-3


@pytest.mark.parametrize("mode", ["ref", "ext"])
def test_fast_fp64_form_within_tolerance(oracle, mode):
    """FP64 with FMA contraction and reciprocal division: not bit-identical,
    but within the FP64 tolerance (relative L1 <= 1e-12 per variable) of the
    oracle after several steps."""
    progress(f"test_fast_fp64_form_within_tolerance {mode}")
    P = state3(40, 24, 20, seed=13)
    if mode == "ref":
        Pg, Po, tg, to, nb, nbo = _run_both(oracle, P, 4, precision="fp64fast")
    else:
        Pg, Po, tg, to, nb, nbo = _run_both(oracle, P, 3, ref=False, riemann="hllc", integrator="rk3",
                                            cfl=0.4, precision="fp64fast",
                                            bc=("outflow",) * 2 + ("periodic",) * 2 + ("outflow",) * 2)
    errs = per_var_rel_l1(interior(Pg), interior(Po))
    progress(f"  per-var rel L1 {['%.2e' % e for e in errs]}, bit-exact {np.mean(Pg == Po):.4f}")
    assert nb == nbo == 0
    assert max(errs) <= TOL
