"""Run-loop drop-in (Verona3D/2D.HARM_HLL over the fused step) and restart from
a dump: a run continued from dump<i>.npz reaches the same final state as the
uninterrupted run.  The dump holds P only (as the reference's does), so the
restarted run rebuilds U = PtoU(P); con2prim's round trip is not bit-exact,
which bounds the agreement at rounding level rather than bitwise."""
import glob
import os

import numpy as np
import pytest

from helpers import per_var_rel_l1, progress, state3

pytestmark = pytest.mark.gpu

G = 4.0 / 3.0


def _run(P, out_dir, Tmax, drops, restart=None):
    from verona_amd import Cart, EosTypes, Verona3D
    nx, ny, nz = (s - 6 for s in P.shape[1:])
    Pv = Verona3D.ParVector3D(nx, ny, nz)
    Pv.arr[...] = P
    CuP = Verona3D.CuParVector3D(Pv, gamma=G)
    try:
        n = Verona3D.HARM_HLL(Cart((True, True, True)), CuP, (1, 1, 1), (4, 4, 4), 0.0, 1 / nx, 1 / ny,
                              1 / nz, Tmax, EosTypes.Polytrope(G), drops, 1e-7, out_dir, False, 50, 1e-10,
                              rec="wenoz", riemann="hllc", integrator="rk3", cfl=0.4, ref_mode=False,
                              precision="fp64", restart=restart)
        final = CuP.arr
    finally:
        CuP.dev.close()
    return n, final


def test_restart_from_dump_continues_the_run(tmp_path):
    progress("restart from dump")
    P = state3(24, 20, 16, seed=31)
    Tmax, drops = 0.05, 0.0125
    a, b = str(tmp_path / "a"), str(tmp_path / "b")
    n_a, Pa = _run(P, a, Tmax, drops)
    dumps = sorted(glob.glob(os.path.join(a, "dump*.npz")), key=lambda p: int(p.rsplit("dump", 1)[1][:-4]))
    assert len(dumps) == n_a + 1 >= 4
    mid = dumps[2]
    n_b, Pb = _run(P, b, Tmax, drops, restart=mid)
    assert n_b == n_a
    with np.load(dumps[-1]) as za, np.load(os.path.join(b, os.path.basename(dumps[-1]))) as zb:
        assert float(za["T"]) == float(zb["T"]) and int(za["step"]) == int(zb["step"])
        errs = per_var_rel_l1(zb["data"], za["data"])
    progress(f"  final dump after restart vs uninterrupted: per-var rel L1 {['%.1e' % e for e in errs]}")
    assert max(errs) <= 1e-12
    # dumps before the restart point were not rewritten
    assert not os.path.exists(os.path.join(b, "dump0.npz"))


def test_verona2d_run_loop_matches_oracle(oracle, tmp_path):
    """Verona2D.HARM_HLL (src/2dim/algos.jl:192-324) as a drop-in: periodic
    Kelvin-Helmholtz at 64^2 in the reference's own method (WENOZ + HLL,
    midpoint RK2, fixed dt); the last dump equals the oracle after the same
    number of steps, and the dump cadence follows `t > thres`."""
    from helpers import from_oracle2, to_oracle2
    from verona_amd import Cart, EosTypes, Verona2D
    from verona_amd.problems import kelvin_helmholtz
    progress("Verona2D run loop")
    n = 64
    dx = 1.0 / n
    dt = 0.4 / (1 / dx + 1 / dx)
    Tmax, drops = 40 * dt, 10 * dt
    P = Verona2D.ParVector2D(n, n)
    P.arr[...] = kelvin_helmholtz(n, gamma=G)
    P0 = P.arr.copy()
    out = str(tmp_path / "kh")
    ndump = Verona2D.HARM_HLL(Cart((True, True)), P, 1, 1, 4, 4, dt, dx, dx, Tmax, EosTypes.Polytrope(G), drops,
                              1e-7, out, 50, 1e-10)
    # the reference's loop: t += dt while t < Tmax, dump when t > thres
    t, steps, thres, i = 0.0, 0, drops, 0
    while t < Tmax:
        t += dt
        steps += 1
        if t > thres:
            i += 1
            thres += drops
    assert ndump == i
    Po = to_oracle2(P0)
    m = oracle.make_mode(ref=True, gamma=G, floor=1e-7, dt=dt, dx=dx, dy=dx, n_iter=50, tol=1e-10)
    bad, _, _ = oracle.run2(Po, m, [1, 1, 1, 1], steps)
    assert bad == 0
    with np.load(os.path.join(out, f"dump{ndump}.npz")) as z:
        data, T = z["data"], float(z["T"])
    ref = from_oracle2(Po)[:, 3:-3, 3:-3]
    assert abs(T - t) == 0.0
    errs = per_var_rel_l1(data, ref)
    progress(f"  {steps} steps, {ndump} dumps, per-var rel L1 {['%.1e' % e for e in errs]}")
    assert max(errs) <= 1e-12


def test_host_callback_boundary_equals_declarative(tmp_path):
    """fun_bound as a plain callable (run on a host copy of P at the start of
    every step, as src/3dim/algos.jl:276 calls it) with the zero-gradient
    copies of examples/Jet3D/HARM_jet_breakout.jl:136-157 gives the same
    result as the declarative Boundary(outflow=...) the library applies on
    the device."""
    from verona_amd import Boundary, Cart, EosTypes, Verona3D
    progress("fun_bound callback vs declarative outflow")
    P = state3(20, 16, 12, seed=33)
    nx, ny, nz = 20, 16, 12
    dt = 0.4 / (nx + ny + nz)
    faces = ("xlo", "xhi", "ylo", "yhi", "zlo", "zhi")

    def outflow(Ph, t, coords, dims):
        # whole-plane copies, x then y then z
        Ph[:, 0:3] = Ph[:, 3:4]
        Ph[:, -3:] = Ph[:, -4:-3]
        Ph[:, :, 0:3] = Ph[:, :, 3:4]
        Ph[:, :, -3:] = Ph[:, :, -4:-3]
        Ph[:, :, :, 0:3] = Ph[:, :, :, 3:4]
        Ph[:, :, :, -3:] = Ph[:, :, :, -4:-3]

    finals = []
    for fb in (outflow, Boundary(outflow=faces)):
        Pv = Verona3D.ParVector3D(nx, ny, nz)
        Pv.arr[...] = P
        CuP = Verona3D.CuParVector3D(Pv, gamma=G)
        try:
            Verona3D.HARM_HLL(Cart((False, False, False)), CuP, (1, 1, 1), (4, 4, 4), dt, 1 / nx, 1 / ny, 1 / nz,
                              6 * dt, EosTypes.Polytrope(G), 10.0, 1e-7, None, False, 50, 1e-10, fb)
            finals.append(CuP.arr[:, 3:-3, 3:-3, 3:-3])
        finally:
            CuP.dev.close()
    assert np.array_equal(finals[0], finals[1])
