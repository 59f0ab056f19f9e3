"""Batched 1D tubes on the GPU against the CPU oracle's Verona1D.HARM_HLL, per tube."""
import numpy as np
import pytest

from helpers import progress

pytestmark = pytest.mark.gpu


def _oracle_tubes(oracle, lengths, rho, u, ux, dt, dx, gamma, rec, riemann, nsteps):
    out = []
    off = 0
    for k, L in enumerate(lengths):
        P = np.ascontiguousarray(np.stack([rho[off:off + L], u[off:off + L], ux[off:off + L]]))
        oracle.run1(P, dt[k], dx[k], 1e30, gamma, rec, riemann, 50, 1e-10, max_steps=nsteps)
        out.append(P)
        off += L
    return np.concatenate(out, axis=1)


@pytest.mark.parametrize("rec,riemann", [("minmod", "hll"), ("minmod", "hllc")])
def test_shock_tube_1024(oracle, rec, riemann):
    """Marti & Mueller problem 1 at N = 1024 to t = 0.4 (1025 steps)."""
    from verona_amd import EosTypes, Verona1D
    progress(f"test_shock_tube_1024 {rec} {riemann}")
    N = 1024
    g = 5.0 / 3.0
    P = Verona1D.ParVector1D(N)
    P.arr1[:512], P.arr2[:512] = 10.0, 13.33 / (g - 1)
    P.arr1[512:], P.arr2[512:] = 1.0, 1e-6 / (g - 1)
    Po = np.ascontiguousarray(np.stack([P.arr1, P.arr2, P.arr3]))
    dx = 1.0 / N
    out = Verona1D.HARM_HLL(P, N, 0.4 * dx, dx, 0.4, EosTypes.Polytrope(g), 0.2, rec, 50, 1e-10, riemann=riemann)
    steps, t = oracle.run1(Po, 0.4 * dx, dx, 0.4, g, rec, riemann, 50, 1e-10)
    G = np.stack([P.arr1, P.arr2, P.arr3])
    progress(f"  steps {steps}, snapshots {len(out)}, bit-exact {np.mean(G == Po):.4f}, max rel "
             f"{np.max(np.abs(G - Po) / np.maximum(np.abs(Po), 1e-300)):.2e}")
    for v in range(3):
        assert np.abs(G[v] - Po[v]).sum() <= 1e-12 * np.abs(Po[v]).sum() + 1e-300
    assert len(out) == 3


@pytest.mark.parametrize("rec,riemann", [("ppm", "hllc"), ("wenoz", "hll")])
def test_batch_matches_oracle_per_tube(oracle, rec, riemann):
    from verona_amd.tubes import Tubes, shock_tube_batch
    progress(f"test_batch_matches_oracle_per_tube {rec} {riemann}")
    lengths, rho, u, ux, dt, dx = shock_tube_batch(48, lmin=40, lmax=700, seed=7)
    nsteps = 40
    with Tubes(lengths) as tb:
        tb.upload(rho, u, ux, dt, dx)
        tb.set_method(rec, riemann, 50, 1e-10)
        tb.run(nsteps, block=128, grid=16)
        G = np.stack(tb.download())
    O = _oracle_tubes(oracle, lengths, rho, u, ux, dt, dx, 5.0 / 3.0, rec, riemann, nsteps)
    finite = np.isfinite(O)
    assert np.array_equal(np.isfinite(G), finite)
    progress(f"  bit-exact {np.mean(G[finite] == O[finite]):.4f}")
    for v in range(3):
        f = finite[v]
        assert np.abs(G[v][f] - O[v][f]).sum() <= 1e-12 * np.abs(O[v][f]).sum()


@pytest.mark.parametrize("rec,riemann", [("ppm", "hllc"), ("wenoz", "hll")])
def test_fast_form_matches_oracle_per_tube(oracle, rec, riemann):
    """The fast FP64 form (FMA contraction, reciprocal division, squared-norm
    Newton stop) against the oracle to 1e-12 relative L1 on the tubes both
    leave finite.  Then, on the same object, switching back to the exact form
    and replaying a captured graph of the same shape must rebuild the graph:
    the result is bit-identical to the oracle again."""
    from verona_amd.tubes import Tubes, shock_tube_batch
    progress(f"test_fast_form_matches_oracle_per_tube {rec} {riemann}")
    lengths, rho, u, ux, dt, dx = shock_tube_batch(48, lmin=40, lmax=700, seed=7)
    nsteps = 40
    O = _oracle_tubes(oracle, lengths, rho, u, ux, dt, dx, 5.0 / 3.0, rec, riemann, nsteps)
    off = np.concatenate([[0], np.cumsum(lengths)])
    with Tubes(lengths) as tb:
        tb.upload(rho, u, ux, dt, dx)
        tb.set_method(rec, riemann, 50, 1e-10)
        tb.set_precision("fp64fast")
        tb.run(nsteps, block=128, grid=16)
        G = np.stack(tb.download())
        tb.upload(rho, u, ux, dt, dx)
        tb.set_precision("fp64")
        tb.run(nsteps, block=128, grid=16)
        E = np.stack(tb.download())
    fin = lambda X: np.array([np.isfinite(X[:, off[t]:off[t + 1]]).all() for t in range(len(lengths))])
    fo, fg = fin(O), fin(G)
    m = np.repeat(fo & fg, lengths)
    errs = [np.abs(G[v][m] - O[v][m]).sum() / np.abs(O[v][m]).sum() for v in range(3)]
    progress(f"  tubes finite: oracle {int(fo.sum())}, fast {int(fg.sum())}, both {int((fo & fg).sum())}; "
             f"rel L1 {['%.2e' % e for e in errs]}")
    assert (fo & fg).sum() >= 0.9 * fo.sum()
    assert max(errs) <= 1e-12
    fe = np.isfinite(O)
    assert np.array_equal(np.isfinite(E), fe) and np.array_equal(E[fe], O[fe])


def test_tubes_refuse_fp32():
    from verona_amd._lib import VeronaHIPError
    from verona_amd.tubes import Tubes, _check
    progress("test_tubes_refuse_fp32")
    with Tubes(np.full(4, 16)) as tb:
        with pytest.raises(ValueError):
            tb.set_precision("fp32")
        with pytest.raises(VeronaHIPError, match="VRN_PREC_FP64"):
            _check(tb.lib.vrn_tubes_set_precision(tb.h, 1), "vrn_tubes_set_precision")


def test_graph_replay_and_autotune_are_transparent(oracle):
    """Graph replays (several launches in a row) and an autotune pass (which
    checkpoints and restores the state) give the same answer as one eager run."""
    from verona_amd.tubes import BLOCKS, Tubes, shock_tube_batch
    progress("test_graph_replay_and_autotune_are_transparent")
    lengths, rho, u, ux, dt, dx = shock_tube_batch(96, lmin=32, lmax=400, seed=9)
    res = []
    for mode in ("eager", "graph", "tuned"):
        with Tubes(lengths) as tb:
            tb.upload(rho, u, ux, dt, dx)
            tb.set_method("ppm", "hllc", 50, 1e-10)
            if mode == "eager":
                tb.run(24, 64, 8, use_graph=False)
            elif mode == "graph":
                for _ in range(3):
                    tb.run(8, 128, 12, use_graph=True)
            else:
                shape = tb.autotune(probe_steps=1)
                assert shape[0] in BLOCKS and shape[1] >= 1
                tb.run(24, *shape)
            res.append(np.stack(tb.download()))
    assert np.array_equal(res[0], res[1], equal_nan=True)
    assert np.array_equal(res[0], res[2], equal_nan=True)


def test_full_batch_matches_oracle_fixture():
    """The whole batched-tubes workload (4,096 tubes, lengths U{256..8192},
    seed 42, PPM + HLLC, RK2, 500 steps) in the exact FP64 form: the cells
    Verona1D's floor-free Newton leaves unrecovered (NaN markers) are exactly
    the oracle's, tube by tube and variable by variable, and every tube's
    sum of |rho|, |u|, |u^x| over its other cells equals the oracle's
    (tests/golden/tubes_c5.npz, written by tests/golden/make_tubes_c5.py)."""
    import os

    from verona_amd.tubes import Tubes, shock_tube_batch
    fx = np.load(os.path.join(os.path.dirname(__file__), "golden", "tubes_c5.npz"))
    ntubes, steps = len(fx["lengths"]), int(fx["steps"])
    progress(f"test_full_batch_matches_oracle_fixture {ntubes} tubes x {steps} steps")
    lengths, rho, u, ux, dt, dx = shock_tube_batch(ntubes, 256, 8192, seed=int(fx["seed"]))
    assert np.array_equal(lengths, fx["lengths"])
    with Tubes(lengths) as tb:
        tb.upload(rho, u, ux, dt, dx)
        tb.set_method("ppm", "hllc", 50, 1e-10)
        tb.run(steps, 256, 4 * tb.cu_count)
        G = np.stack(tb.download())
    off = np.concatenate([[0], np.cumsum(lengths)])
    nan_cells = np.zeros((ntubes, 3), np.int64)
    abs_sum = np.zeros((ntubes, 3))
    for t in range(ntubes):
        P = G[:, off[t]:off[t + 1]]
        nan = ~np.isfinite(P)
        nan_cells[t], abs_sum[t] = nan.sum(axis=1), np.where(nan, 0.0, np.abs(P)).sum(axis=1)
    nan_tubes = np.flatnonzero(nan_cells.sum(axis=1) > 0)
    rel = np.abs(abs_sum - fx["abs_sum"]) / np.maximum(np.abs(fx["abs_sum"]), 1e-300)
    progress(f"  NaN tubes: GPU {len(nan_tubes)}, oracle {len(fx['nan_tubes'])}; checksums equal "
             f"{int((abs_sum == fx['abs_sum']).all(axis=1).sum())}/{ntubes}, max rel diff {rel.max():.2e}")
    assert np.array_equal(nan_tubes, fx["nan_tubes"])
    assert np.array_equal(nan_cells, fx["nan_cells"])
    assert np.array_equal(abs_sum, fx["abs_sum"])
