"""The drop-in paths against the oracle (not against the GPU itself).

* julia/VeronaHIP.jl's HARM_HLL loop body (src/3dim/algos.jl:274-329), read
  from the shim file and executed call for call through the C entry points:
  bit-identical to the oracle's run of the reference sequence.
* Verona3D.HARM_HLL (the Python mirror of src/3dim/algos.jl:203-362 over
  the fused step): every dump, its time T and the dump cadence equal the
  oracle after the same number of steps.
* Restart from a dump: the continued run equals an oracle run that stops
  at the dump, rebuilds U = PtoU(P) from the dumped primitives (as the
  reference does at the start of a run, src/3dim/algos.jl:259) and goes on.
* A host fun_bound callback (src/3dim/algos.jl:276) equals the oracle with
  the examples' zero-gradient boundaries applied at each step start.
"""
import glob
import os

import numpy as np
import pytest

import julia_shim
from helpers import from_oracle3, interior, per_var_rel_l1, progress, state3, to_oracle3

pytestmark = pytest.mark.gpu

G = 4.0 / 3.0
FLOOR, N_ITER, TOL = 1e-7, 50, 1e-10


def _steps_and_dumps(dt, Tmax, drops):
    """The reference's loop counters: (steps, [(dump index, steps so far, t)])."""
    t, steps, thres, i, dumps = 0.0, 0, drops, 0, [(0, 0, 0.0)]
    while t < Tmax:
        t += dt
        steps += 1
        if t > thres:
            i += 1
            thres += drops
            dumps.append((i, steps, t))
    return steps, dumps


def _oracle_ref(P, nsteps, dt, dx, dy, dz, bc_types, oracle):
    Po = to_oracle3(P)
    m = oracle.make_mode(ref=True, gamma=G, floor=FLOOR, dt=dt, dx=dx, dy=dy, dz=dz, n_iter=N_ITER, tol=TOL)
    bc, _ = oracle.make_bc3(bc_types)
    bad, t, _ = oracle.run3(Po, m, bc, nsteps)
    assert bad == 0
    return from_oracle3(Po), t


def test_julia_shim_kernel_sequence_matches_oracle(oracle):
    from verona_amd import DeviceGrid
    nx, ny, nz, nsteps = 24, 20, 16, 4
    progress(f"julia shim kernel sequence {nx}x{ny}x{nz}, {nsteps} steps")
    seq = julia_shim.step_sequence()
    P = state3(nx, ny, nz, seed=41)
    dx, dy, dz = 1 / nx, 1 / ny, 1 / nz
    dt = 0.4 / (1 / dx + 1 / dy + 1 / dz)
    env = {"dt": dt, "dx": dx, "dy": dy, "dz": dz, "floor": FLOOR, "n_iter": N_ITER, "tol": TOL,
           "SLOT_UHALF": "Uhalf", "SLOT_U": "U", "bc": ["periodic"] * 6}
    val = lambda e: eval(e, {"__builtins__": {}}, env)  # noqa: E731
    with DeviceGrid(3, nx, ny, nz, G) as dev:
        # prelude of HARM_HLL: upload!, fill_ghosts, PtoU
        dev.upload("P", P)
        dev.fill_ghosts(env["bc"])
        dev.prim_to_cons()
        for _ in range(nsteps):
            for name, args in seq:
                if name == "Fluxes":
                    dev.fluxes(int(args[2]) - 1, val(args[1]), "wenoz", "hll")
                elif name == "Update":
                    dev.update(val(args[1]), *[val(a) for a in args[2:6]])
                elif name == "UtoP":
                    dev.cons_to_prim(val(args[1]), val(args[2]), val(args[3]))
                elif name == "Limit":
                    dev.limit(val(args[1]))
                elif name == "fill_ghosts":
                    dev.fill_ghosts(val(args[1]))
                else:
                    raise AssertionError(f"unknown call {name} in the shim's loop")
        Pg = dev.download("P")
        nbad, _ = dev.counters()
    Po, _ = _oracle_ref(P, nsteps, dt, dx, dy, dz, [1] * 6, oracle)
    a, b = interior(Pg), interior(Po)
    same = float(np.mean(a == b))
    progress(f"  {len(seq)} calls per step; bit-identical fraction {same:.6f}, "
             f"per-var rel L1 {['%.1e' % e for e in per_var_rel_l1(a, b)]}")
    assert nbad == 0
    assert np.array_equal(a, b)


def test_verona3d_run_loop_matches_oracle(oracle, tmp_path):
    """Verona3D.HARM_HLL in the reference's own method (WENOZ + HLL, midpoint
    RK2, fixed dt) on a periodic grid: dump cadence (t > thres), each dump's T
    and data equal the oracle after the same number of steps."""
    from verona_amd import Cart, EosTypes, Verona3D
    nx, ny, nz = 24, 20, 16
    progress(f"Verona3D run loop {nx}x{ny}x{nz}")
    P = state3(nx, ny, nz, seed=43)
    dx, dy, dz = 1 / nx, 1 / ny, 1 / nz
    dt = 0.4 / (1 / dx + 1 / dy + 1 / dz)
    Tmax, drops = 10 * dt, 3 * dt
    Pv = Verona3D.ParVector3D(nx, ny, nz)
    Pv.arr[...] = P
    out = str(tmp_path / "run")
    ndump = Verona3D.HARM_HLL(Cart((True, True, True)), Pv, (1, 1, 1), (4, 4, 4), dt, dx, dy, dz, Tmax,
                              EosTypes.Polytrope(G), drops, FLOOR, out, False, N_ITER, TOL)
    steps, dumps = _steps_and_dumps(dt, Tmax, drops)
    assert ndump == dumps[-1][0]
    files = glob.glob(os.path.join(out, "dump*.npz"))
    assert len(files) == ndump + 1
    for i, k, t in dumps:
        with np.load(os.path.join(out, f"dump{i}.npz")) as z:
            data, T = z["data"], float(z["T"])
        ref = interior(P) if k == 0 else interior(_oracle_ref(P, k, dt, dx, dy, dz, [1] * 6, oracle)[0])
        progress(f"  dump{i}: {k} steps, T {T!r}, bit-identical {np.array_equal(data, ref)}")
        assert T == t
        assert np.array_equal(data, ref)


def test_restart_matches_oracle(oracle, tmp_path):
    """A run restarted from dump<i>.npz (the extended method: HLLC, SSP-RK3,
    CFL dt, exact FP64) equals the oracle stopped after the dump's step
    count, restarted from the dumped interior primitives with U = PtoU(P)
    rebuilt, and run for the remaining steps."""
    from verona_amd import Cart, EosTypes, Verona3D
    nx, ny, nz = 24, 20, 16
    progress("restart vs oracle")
    P = state3(nx, ny, nz, seed=31)
    dx, dy, dz = 1 / nx, 1 / ny, 1 / nz
    Tmax, drops = 0.05, 0.0125
    kw = dict(rec="wenoz", riemann="hllc", integrator="rk3", cfl=0.4, ref_mode=False, precision="fp64")

    def run(out_dir, restart=None):
        Pv = Verona3D.ParVector3D(nx, ny, nz)
        Pv.arr[...] = P
        CuP = Verona3D.CuParVector3D(Pv, gamma=G)
        try:
            return Verona3D.HARM_HLL(Cart((True, True, True)), CuP, (1, 1, 1), (4, 4, 4), 0.0, dx, dy, dz, Tmax,
                                     EosTypes.Polytrope(G), drops, FLOOR, out_dir, False, N_ITER, TOL,
                                     restart=restart, **kw)
        finally:
            CuP.dev.close()

    a, b = str(tmp_path / "a"), str(tmp_path / "b")
    n_a = run(a)
    assert n_a >= 3
    mid = os.path.join(a, "dump2.npz")
    n_b = run(b, restart=mid)
    assert n_b == n_a
    with np.load(mid) as z:
        k, Tk, Pk_data = int(z["step"]), float(z["T"]), z["data"]
    with np.load(os.path.join(b, f"dump{n_b}.npz")) as z:
        n_tot, T_end, data_end = int(z["step"]), float(z["T"]), z["data"]

    m = oracle.make_mode(ref=False, rec="wenoz", riemann="hllc", integrator="rk3", gamma=G, floor=FLOOR,
                         cfl=0.4, dx=dx, dy=dy, dz=dz, n_iter=N_ITER, tol=TOL)
    bc, _ = oracle.make_bc3([1] * 6)
    # the oracle up to the dump, then its dumped interior as a fresh start
    Po = to_oracle3(P)
    bad, t1, _ = oracle.run3(Po, m, bc, k)
    assert bad == 0
    Pk = from_oracle3(Po)
    assert np.array_equal(interior(Pk), Pk_data) and t1 == Tk
    Pr = P.copy(order="F")
    Pr[:, 3:-3, 3:-3, 3:-3] = Pk_data
    Po = to_oracle3(Pr)
    bad, t2, _ = oracle.run3(Po, m, bc, n_tot - k)
    assert bad == 0
    ref = interior(from_oracle3(Po))
    errs = per_var_rel_l1(data_end, ref)
    progress(f"  restart at step {k} of {n_tot}: bit-identical {np.array_equal(data_end, ref)}, "
             f"per-var rel L1 {['%.1e' % e for e in errs]}, T {T_end!r} vs {Tk + t2!r}")
    assert np.array_equal(data_end, ref)
    assert abs(T_end - (Tk + t2)) <= 1e-15 * T_end


def test_host_callback_boundary_matches_oracle():
    """fun_bound as a host callable with the zero-gradient copies of
    examples/Jet3D/HARM_jet_breakout.jl:136-157, called at every step start
    (src/3dim/algos.jl:276): the run equals the oracle with outflow faces
    applied at the step start (the reference's callback timing)."""
    import pyoracle as oracle

    from verona_amd import Cart, EosTypes, Verona3D
    nx, ny, nz = 20, 16, 12
    progress("fun_bound callback vs oracle")
    P = state3(nx, ny, nz, seed=33)
    dx, dy, dz = 1 / nx, 1 / ny, 1 / nz
    dt = 0.4 / (nx + ny + nz)
    nsteps = 6

    def outflow(Ph, t, coords, dims):
        Ph[:, 0:3] = Ph[:, 3:4]
        Ph[:, -3:] = Ph[:, -4:-3]
        Ph[:, :, 0:3] = Ph[:, :, 3:4]
        Ph[:, :, -3:] = Ph[:, :, -4:-3]
        Ph[:, :, :, 0:3] = Ph[:, :, :, 3:4]
        Ph[:, :, :, -3:] = Ph[:, :, :, -4:-3]

    Pv = Verona3D.ParVector3D(nx, ny, nz)
    Pv.arr[...] = P
    CuP = Verona3D.CuParVector3D(Pv, gamma=G)
    try:
        Verona3D.HARM_HLL(Cart((False, False, False)), CuP, (1, 1, 1), (4, 4, 4), dt, dx, dy, dz,
                          nsteps * dt, EosTypes.Polytrope(G), 10.0, FLOOR, None, False, N_ITER, TOL, outflow)
        got = interior(CuP.arr)
    finally:
        CuP.dev.close()
    steps, _ = _steps_and_dumps(dt, nsteps * dt, 10.0)
    ref, _ = _oracle_ref(P, steps, dt, dx, dy, dz, [2] * 6, oracle)
    progress(f"  {steps} steps, bit-identical {np.array_equal(got, interior(ref))}")
    assert np.array_equal(got, interior(ref))
