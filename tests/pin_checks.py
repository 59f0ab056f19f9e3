"""Pin checks for oracle functions that closed forms alone do not fix.

Each check raises AssertionError when the oracle (the module `pyoracle`,
possibly routed to another build with `pyoracle.use`) breaks the property.
tests/test_oracle_pins.py runs them on the real oracle; tests/
test_oracle_mutants.py runs each on an oracle build carrying one planted
mistake (orc_core.h ORC_MUT_*) and requires it to fail there.

None of these restates the oracle's formula: they compare it with
mathematics that fixes the answer independently -- temporal convergence
order, the textbook fourth-order interface value, central finite
differences, exact rational arithmetic, and the iteration's stopping rule
observed from outside.
"""
import math
from fractions import Fraction

import numpy as np

G3 = 4.0 / 3.0
G1 = 5.0 / 3.0


# ------------------------------------------------------------------ SSP-RK3
def _advection_2d(O, nsteps, T=0.2, n=32):
    """Density wave rho = 1 + 0.2 sin(2 pi x) carried at v = 0.5 through a
    periodic 2D grid (n x 4, p = 1), SSP-RK3 with a fixed dt = T / nsteps."""
    NX, NY = n + 6, 4 + 6
    x = (np.arange(NX) - 3 + 0.5) / n
    P = np.zeros((NY, NX, 4))
    P[:, :, 0] = 1.0 + 0.2 * np.sin(2 * np.pi * x)[None, :]
    P[:, :, 1] = 1.0 / (G3 - 1.0)
    v = 0.5
    P[:, :, 2] = v / math.sqrt(1 - v * v)
    m = O.make_mode(ref=False, rec="wenoz", riemann="hll", integrator="rk3", gamma=G3, floor=1e-9,
                    dt=T / nsteps, dx=1.0 / n, dy=1.0 / 4, n_iter=100, tol=1e-14)
    bad, t, _ = O.run2(P, m, [1, 1, 1, 1], nsteps)
    assert bad == 0
    return P[3:-3, 3:-3, 0]


def _advection_3d(O, nsteps, T=0.2, n=24):
    """The same density wave along z through a periodic 4 x 4 x n grid."""
    NX, NY, NZ = 4 + 6, 4 + 6, n + 6
    z = (np.arange(NZ) - 3 + 0.5) / n
    P = np.zeros((NZ, NY, NX, 5))
    P[..., 0] = 1.0 + 0.2 * np.sin(2 * np.pi * z)[:, None, None]
    P[..., 1] = 1.0 / (G3 - 1.0)
    v = 0.5
    P[..., 4] = v / math.sqrt(1 - v * v)
    m = O.make_mode(ref=False, rec="wenoz", riemann="hll", integrator="rk3", gamma=G3, floor=1e-9,
                    dt=T / nsteps, dx=0.25, dy=0.25, dz=1.0 / n, n_iter=100, tol=1e-14)
    bc, _ = O.make_bc3([1] * 6)
    bad, t, _ = O.run3(P, m, bc, nsteps)
    assert bad == 0
    return P[3:-3, 3:-3, 3:-3, 0]


def rk3_temporal_order(O, dims=2):
    """Third-order convergence in time of the SSP-RK3 integrator on a fixed
    grid: the method-of-lines error against a 16x finer dt falls by ~8x per
    halving of dt (Shu & Osher 1988).  Returns the two observed rates."""
    run = _advection_2d if dims == 2 else _advection_3d
    base = 10
    ref = run(O, base * 16)
    errs = [np.abs(run(O, base * k) - ref).max() for k in (1, 2, 4)]
    rates = [math.log2(errs[i] / errs[i + 1]) for i in range(2)]
    assert min(rates) > 2.7, (errs, rates)
    return rates


# ------------------------------------------------------------------ PPM
def _cos_cell_averages(n, shift):
    """Exact cell averages of cos(2 pi x) on n cells whose centres sit at
    (k + shift)/n, with the two neighbours on each side for the stencil."""
    h = 1.0 / n
    xc = (np.arange(n) + shift) * h
    avg = lambda c: (np.sin(2 * np.pi * (c + h / 2)) - np.sin(2 * np.pi * (c - h / 2))) / (2 * np.pi * h)  # noqa: E731
    return np.stack([avg(xc + k * h) for k in range(-2, 3)], axis=1), xc, h


def ppm_convergence_order(O):
    """Face values of smooth data converge at fourth order (the interface
    values are fourth-order and the limiter leaves smooth data alone)."""
    errs = []
    for n in (16, 32, 64, 128):
        q, xc, h = _cos_cell_averages(n, 0.5)
        lo, hi = O.ppm(q)
        e = max(np.abs(lo - np.cos(2 * np.pi * (xc - h / 2))).max(),
                np.abs(hi - np.cos(2 * np.pi * (xc + h / 2))).max())
        errs.append(e)
    rates = [math.log2(errs[i] / errs[i + 1]) for i in range(3)]
    assert min(rates) > 3.8, (errs, rates)
    return rates


def ppm_smooth_extremum(O):
    """A smooth extremum at (or a quarter cell from) a cell centre, resolved
    by >= 16 cells per wavelength, is left unclipped: every cell returns the
    textbook fourth-order interface values q_{i+1/2} = 7/12 (q_i + q_{i+1})
    - 1/12 (q_{i-1} + q_{i+2}).  That is the design point of the
    Colella-Sekora extremum limiter with C2 = 1.25: neighbouring second
    differences of smooth data differ by O(h) relative, inside the factor
    1.25, so the limited curvature equals the interface curvature."""
    for n in (16, 24, 32, 64):
        for shift in (0.0, 0.25):
            q, xc, h = _cos_cell_averages(n, shift)
            lo, hi = O.ppm(q)
            tl = 7.0 / 12.0 * (q[:, 1] + q[:, 2]) - 1.0 / 12.0 * (q[:, 0] + q[:, 3])
            th = 7.0 / 12.0 * (q[:, 2] + q[:, 3]) - 1.0 / 12.0 * (q[:, 1] + q[:, 4])
            bad = np.nonzero((np.abs(lo - tl) > 1e-14) | (np.abs(hi - th) > 1e-14))[0]
            assert len(bad) == 0, (n, shift, xc[bad])


# ------------------------------------------------------------------ 1D Newton
def _states1(rng, n):
    rho = 10 ** rng.uniform(-2, 1, n)
    u = 10 ** rng.uniform(-3, 1, n)
    ux = rng.uniform(-3, 3, n)
    return np.stack([rho, u, ux], axis=1)


def jacobian1_fd(O, seed=5):
    """The 1D Newton's analytic Jacobian equals central finite differences of
    its residual map F_ptoU (1dim/structs.jl:32-54) to 1e-7 relative."""
    rng = np.random.default_rng(seed)
    worst = 0.0
    for x in _states1(rng, 200):
        J = O.jacobian1(x, G1)
        Jfd = np.empty((3, 3))
        for j in range(3):
            h = 1e-6 * max(abs(x[j]), 1e-3)
            xp, xm = x.copy(), x.copy()
            xp[j] += h
            xm[j] -= h
            Jfd[:, j] = (O.f_ptou1(xp, G1) - O.f_ptou1(xm, G1)) / (xp[j] - xm[j])
        scale = np.abs(J).max()
        worst = max(worst, np.abs(J - Jfd).max() / scale)
    assert worst < 1e-7, worst
    return worst


def solve3_exact(O, seed=6):
    """On small integer systems every product and sum of the adjugate solve
    is exact, so J \\ b must equal the correctly rounded rational solution
    (Cramer's rule in exact arithmetic) bit for bit."""
    rng = np.random.default_rng(seed)
    done = 0
    while done < 300:
        A = rng.integers(-9, 10, (3, 3))
        b = rng.integers(-9, 10, 3)
        M = [[Fraction(int(v)) for v in row] for row in A]
        det = (M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1])
               - M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0])
               + M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]))
        if det == 0:
            continue
        exact = []
        for col in range(3):
            Mc = [row[:] for row in M]
            for r in range(3):
                Mc[r][col] = Fraction(int(b[r]))
            dc = (Mc[0][0] * (Mc[1][1] * Mc[2][2] - Mc[1][2] * Mc[2][1])
                  - Mc[0][1] * (Mc[1][0] * Mc[2][2] - Mc[1][2] * Mc[2][0])
                  + Mc[0][2] * (Mc[1][0] * Mc[2][1] - Mc[1][1] * Mc[2][0]))
            exact.append(float(dc / det))
        got = O.solve3(A.astype(np.float64), b.astype(np.float64))
        assert np.array_equal(got, np.array(exact)), (A, b, got, exact)
        done += 1


def newton1_break_before_update(O, seed=7):
    """UtoP (1dim/structs.jl:86-101) tests the step's norm before applying
    it.  (a) A start within tol of the root: the first step is below tol, so
    P comes back bit-for-bit unchanged.  (b) One iteration from a start whose
    step exceeds tol applies exactly one Newton step x - J(x)^-1 (F(x) - U)
    (checked against numpy's LU solve of the same Jacobian)."""
    rng = np.random.default_rng(seed)
    Pt = _states1(rng, 64)
    U = np.stack([O.f_ptou1(p, G1) for p in Pt])
    Us = np.ascontiguousarray(U.T)
    # (a) start 1e-8 relative from the root, tol 1e-3
    P0 = Pt * (1 + 1e-8 * rng.standard_normal(Pt.shape))
    got = np.ascontiguousarray(P0.T.copy())
    O.c2p1(Us, got, G1, 20, 1e-3)
    assert np.array_equal(got, P0.T), "a step below tol was applied"
    # (b) start 5 % away, one iteration, tol 1e-30
    P0 = Pt * (1 + 0.05 * rng.standard_normal(Pt.shape))
    got = np.ascontiguousarray(P0.T.copy())
    O.c2p1(Us, got, G1, 1, 1e-30)
    for i in range(len(Pt)):
        x = P0[i]
        step = np.linalg.solve(O.jacobian1(x, G1), O.f_ptou1(x, G1) - U[i])
        want = x - step
        assert np.allclose(got[:, i], want, rtol=1e-11, atol=0), (i, got[:, i], want)


# ------------------------------------------------------------------ 1DW derivative
def f1dw_derivative_fd(O, seed=8):
    """The analytic f'(W) of the 1DW iteration (3dim/structs.jl:318-321)
    equals central finite differences of its residual f(W) (:315-317) with
    v^2 = S^2/W^2, on conserved states with Lorentz factors 1-20 and W from
    the lower bound |S| up to 10x the root."""
    rng = np.random.default_rng(seed)
    worst = 0.0
    for _ in range(200):
        rho = 10 ** rng.uniform(-3, 1)
        p = 10 ** rng.uniform(-4, 1)
        lor = 10 ** rng.uniform(0, math.log10(20))
        d = rng.standard_normal(3)
        ui = math.sqrt(lor * lor - 1) * d / np.linalg.norm(d)
        Pc = np.array([rho, p / (G3 - 1), *ui])
        U = O.p2u3(Pc, G3)[0]
        S = math.sqrt((U[2:] ** 2).sum())
        Wroot = (rho + G3 / (G3 - 1) * p) * lor * lor
        for W in (Wroot, Wroot * rng.uniform(1.01, 10), S * rng.uniform(1.2, 2.0)):
            if S * S / (W * W) > 1 - 1e-6:
                continue
            h = 1e-6 * W
            fp_fd = (O.f1dw3(U, W + h, G3)[0] - O.f1dw3(U, W - h, G3)[0]) / ((W + h) - (W - h))
            fp = O.f1dw3(U, W, G3)[1]
            worst = max(worst, abs(fp - fp_fd) / max(abs(fp), 1e-300))
    assert worst < 1e-6, worst
    return worst
