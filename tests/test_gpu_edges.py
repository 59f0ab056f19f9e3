"""Edge cases of the GPU path against the oracle: the smallest legal grid, a
strongly anisotropic grid (x rows split into segments), near-vacuum states
where the floors bind, a NaN cell (the reference's failure marker), the
shortest tubes, and ragged / extreme con2prim batches."""
import numpy as np
import pytest

from helpers import interior, per_var_rel_l1, progress, state3
from test_gpu_parity3d import G, TOL, _run_both

pytestmark = pytest.mark.gpu

EXT = dict(ref=False, riemann="hllc", integrator="rk3", cfl=0.4)


def test_smallest_grid(oracle):
    progress("edge: 4^3 interior")
    P = state3(4, 4, 4, seed=21)
    Pg, Po, tg, to, nb, nbo = _run_both(oracle, P, 3, **EXT)
    assert nb == nbo == 0
    assert max(per_var_rel_l1(interior(Pg), interior(Po))) <= TOL
    assert abs(tg - to) <= 1e-14 * to


def test_anisotropic_grid_splits_rows(oracle):
    # 5 x 6 rows of 200 cells: too few rows for the chip, so each row is
    # streamed as several segments
    progress("edge: 200 x 5 x 6")
    P = state3(200, 5, 6, seed=22)
    Pg, Po, tg, to, nb, nbo = _run_both(oracle, P, 2, **EXT,
                                        bc=("outflow", "outflow", "periodic", "periodic", "periodic",
                                            "periodic"))
    assert nb == nbo == 0
    assert max(per_var_rel_l1(interior(Pg), interior(Po))) <= TOL


@pytest.mark.parametrize("ref", [True, False])
def test_near_vacuum_floors(oracle, ref):
    progress(f"edge: near vacuum ref={ref}")
    P = state3(24, 20, 16, seed=23)
    x = np.arange(P.shape[1])[:, None, None]
    low = (x > 12) & (x < 20)
    P[0] = np.where(low, 3e-7, P[0])   # a few times the 1e-7 floor
    P[1] = np.where(low, 2e-7, P[1])
    kw = {} if ref else EXT
    Pg, Po, tg, to, nb, nbo = _run_both(oracle, P, 2, **kw)
    progress(f"  unrecovered cells gpu {nb} oracle {nbo}")
    assert nb == nbo
    A, B = interior(Pg), interior(Po)
    assert np.array_equal(np.isnan(A), np.isnan(B))
    ok = np.isfinite(B).all(axis=0)
    assert (A[0][ok] >= 1e-7).all() and (A[1][ok] >= 1e-7).all()
    for v in range(5):
        a, b = A[v][ok], B[v][ok]
        assert np.abs(a - b).sum() <= TOL * np.abs(b).sum()


def test_nan_cell_follows_reference_marker(oracle):
    progress("edge: NaN cell")
    P = state3(20, 16, 12, seed=24)
    P[1, 10, 8, 6] = np.nan
    Pg, Po, tg, to, nb, nbo = _run_both(oracle, P, 1)
    assert nb == nbo and nb > 0
    assert np.array_equal(np.isnan(interior(Pg)), np.isnan(interior(Po)))
    ok = np.isfinite(interior(Po))
    a, b = interior(Pg)[ok], interior(Po)[ok]
    assert np.abs(a - b).sum() <= TOL * np.abs(b).sum()


def test_shortest_tubes(oracle):
    from test_gpu_tubes import _oracle_tubes
    from verona_amd.tubes import Tubes, shock_tube_batch
    progress("edge: 8-cell tubes")
    lengths, rho, u, ux, dt, dx = shock_tube_batch(20, lmin=8, lmax=12, seed=25)
    with Tubes(lengths) as tb:
        tb.upload(rho, u, ux, dt, dx)
        tb.set_method("ppm", "hllc", 50, 1e-10)
        tb.run(30, block=64, grid=4)
        Gt = np.stack(tb.download())
    O = _oracle_tubes(oracle, lengths, rho, u, ux, dt, dx, 5.0 / 3.0, "ppm", "hllc", 30)
    finite = np.isfinite(O)
    assert np.array_equal(np.isfinite(Gt), finite)
    for v in range(3):
        f = finite[v]
        assert np.abs(Gt[v][f] - O[v][f]).sum() <= 1e-12 * np.abs(O[v][f]).sum()


@pytest.mark.parametrize("n", [1, 63, 1000])
def test_c2p_ragged_batches(oracle, n):
    from test_gpu_c2p import check_counters
    from verona_amd.c2p import Con2PrimBatch, c4_states, warm_guess
    progress(f"edge: con2prim batch of {n}")
    U, P = c4_states(n, seed=26, gamma=G)
    Pp = warm_guess(P)
    with Con2PrimBatch(1024, G) as cp:
        cp.upload(U, Pp)
        for guess in ("generic", "warm", "analytic"):
            st = cp.solve(guess, n_iter=100, rtol=1e-12, record=True)
            Pd, W0, it = cp.download(record=True)
            Po, ito, cv = oracle.c2p4_solve(U, W0, G, 100, 1e-12)
            assert st["n"] == n and st["n_unrecovered"] == 0 and cv.all()
            assert np.array_equal(Pd, Po) and np.array_equal(it.astype(np.int32), ito)
            check_counters(st, it)


def test_c2p_extreme_corners(oracle):
    """Corners of the benchmark law: Gamma = 100 cold thin gas (W within
    5e-5 of |S|, where the W_min clamp matters), Gamma = 1 hot dense gas, and
    mixed."""
    from verona_amd.c2p import Con2PrimBatch, p2u
    progress("edge: con2prim corners")
    rows = []
    for rho in (1e-4, 1e2):
        for p in (1e-6, 1e2):
            for lor in (1.0, 1.0 + 1e-12, 100.0):
                for d in ((1, 0, 0), (0, 0, -1), (0.6, 0.0, 0.8)):
                    ur = np.sqrt(lor * lor - 1.0)
                    rows.append([rho, p / (G - 1.0), ur * d[0], ur * d[1], ur * d[2]])
    P = np.array(rows)
    U = p2u(P, G)
    with Con2PrimBatch(1024, G) as cp:
        cp.upload(U)
        for guess in ("generic", "analytic"):
            st = cp.solve(guess, n_iter=200, rtol=1e-12, record=True)
            Pd, W0, it = cp.download(record=True)
            Po, ito, cv = oracle.c2p4_solve(U, W0, G, 200, 1e-12)
            assert st["n_unrecovered"] == 0 and cv.all()
            assert np.array_equal(Pd, Po)
            # recovered density matches the state the cells were built from
            assert np.allclose(Pd[:, 0], P[:, 0], rtol=1e-6)


def test_partition_changes_no_arithmetic():
    """The sweep work partition (rows per wave, row segments, march segment
    length) only redistributes work: results are bit-identical."""
    from verona_amd import DeviceGrid
    progress("edge: partition invariance")
    P = state3(96, 20, 24, seed=27)
    dx = 1 / 96
    outs = []
    for part in (None, (1, 3, 8), (5, 1, 13)):
        with DeviceGrid(3, 96, 20, 24, G) as dev:
            if part is not None:
                dev.set_partition(*part)
                assert dev.partition() == part
            dev.upload("P", P)
            bc = ("outflow", "outflow", "periodic", "periodic", "outflow", "outflow")
            dev.fill_ghosts(bc)
            dev.prim_to_cons()
            dev.set_step(ref_mode=False, rec="wenoz", riemann="hllc", integrator="rk3", precision="fp64fast",
                         n_iter=50, tol=1e-10, floor=1e-7, cfl=0.4, dx=dx, dy=1 / 20, dz=1 / 24, bc=bc)
            for _ in range(2):
                dev.step()
            outs.append(dev.download("P"))
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])


@pytest.mark.parametrize("floor", [float("nan"), -0.0])
def test_step_rejects_floor_without_a_numeric_bound(floor):
    """Floors are applied as max(floor, x) by a compare-select, which equals
    Julia's max only for a numeric bound other than -0: vrn_step refuses the
    rest with an error instead of producing a different state."""
    from verona_amd import DeviceGrid
    from verona_amd._lib import VeronaHIPError
    progress(f"edge: floor {floor!r} rejected")
    with DeviceGrid(3, 8, 8, 8, G) as dev:
        dev.upload("P", state3(8, 8, 8, seed=28))
        dev.prim_to_cons()
        bc = ("periodic",) * 6
        dev.set_step(ref_mode=False, rec="wenoz", riemann="hllc", integrator="rk3", precision="fp64fast",
                     n_iter=50, tol=1e-10, floor=floor, cfl=0.4, dx=1 / 8, dy=1 / 8, dz=1 / 8, bc=bc)
        with pytest.raises(VeronaHIPError, match="floor"):
            dev.step()


@pytest.mark.parametrize("rec", ["ppm", "minmod", "wenoz"])
@pytest.mark.parametrize("riemann", ["hll", "hllc"])
def test_fluxes_on_special_values(oracle, rec, riemann):
    """The limiters' Julia sign / min / max semantics (PPM's written with
    compares and copysign) on data built to reach every branch: exact zeros
    of both signs, tiny magnitudes, sign flips, flat runs, an infinite
    velocity and a NaN density.  The exact-form face fluxes equal the
    oracle's bit for bit, with NaN in the same places."""
    from helpers import from_oracle3, to_oracle3
    from verona_amd import DeviceGrid
    progress(f"edge: fluxes on special values rec={rec} riemann={riemann}")
    P = state3(24, 12, 10, seed=31)
    rng = np.random.default_rng(31)
    shape = P.shape[1:]
    for v in (2, 3, 4):
        pick = rng.integers(0, 6, shape)
        vals = np.stack([np.zeros(shape), np.full(shape, -0.0), 1e-30 * rng.standard_normal(shape),
                         np.where(rng.standard_normal(shape) > 0, 0.3, -0.3), P[v], np.full(shape, 0.25)])
        P[v] = np.take_along_axis(vals, pick[None], 0)[0]
    P[0][5:12] = 1.0  # flat runs along x
    P[1][5:12] = 2.0
    P[4][15, 6, 5] = np.inf
    P[0][18, 4, 7] = np.nan
    Po = to_oracle3(P)
    sl = (slice(None), slice(2, -3), slice(2, -3), slice(2, -3))
    with DeviceGrid(3, 24, 12, 10, G) as dev:
        dev.upload("P", P)
        for dim, slot in enumerate(["Fx", "Fy", "Fz"]):
            dev.upload(slot, np.zeros_like(P))
            dev.fluxes(dim, 1e-7, rec, riemann)
            a = dev.download(slot)[sl]
            b = from_oracle3(oracle.fluxes3(Po, G, 1e-7, dim, rec, riemann))[sl]
            nan = np.isnan(b)
            progress(f"  dim={dim}: NaN faces {int(nan.any(axis=0).sum())}, "
                     f"bit-exact {np.mean((a == b) | (np.isnan(a) & nan)):.6f}")
            assert np.array_equal(np.isnan(a), nan)
            assert np.array_equal(a[~nan], b[~nan])
