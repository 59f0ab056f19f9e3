"""Benchmark: zone-updates/s of the 3D relativistic-jet workload on one MI355X.

Workload (fits one GPU): 384^3 interior cells on [0,1]^3, gamma = 4/3, ambient
rho = 1, p = 1e-4 at rest; jet nozzle of radius 8 cells on the z-low face
(rho = 0.01, p = 1e-4, Lorentz factor 10 along e_z with a 1% precession angle
drawn per step from SplitMix64 seed 42); outflow elsewhere.  WENOZ + HLLC +
SSP-RK3, CFL 0.4, all FP64.  One step = one full SSP-RK3 time step of the
whole grid (inputs resident in HBM); value = interior cells x steps / s.

Also timed, same JSON line: the reference-semantics form (WENOZ + HLL +
midpoint RK2, fixed dt) and the FP32 reconstruction/Riemann form, plus the
CPU oracle on a bounded sample of the same workload.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 384]
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
GAMMA = 4.0 / 3.0
# compulsory HBM bytes per interior zone-update (FP64 P and U, 5 variables):
# SSP-RK3 = 14 array passes, midpoint RK2 = 8 array passes
BYTES_PER_ZU = {"rk3": 14 * 40, "rk2": 8 * 40}


def jet_state(n):
    """Ambient medium at rest, Verona layout (5, n+6, n+6, n+6)."""
    from verona_amd import host_array
    P = host_array(3, n, n, n)
    P[0] = 1.0
    P[1] = 1e-4 / (GAMMA - 1.0)
    return P


def nozzle(n):
    dx = 1.0 / n
    return dict(radius=8 * dx, rho=0.01, u=1e-4 / (GAMMA - 1.0), umag=math.sqrt(99.0), eps=0.01)


def phi_stream(count, seed=42):
    from verona_amd.bc import uniform_stream
    it = uniform_stream(seed)
    return [2.0 * math.pi * next(it) for _ in range(count)]


BC_JET = ["outflow", "outflow", "outflow", "outflow", "nozzle", "outflow"]


def run_gpu(n, steps, warmup, *, ref, precision="fp64fast", timing=True):
    from verona_amd import DeviceGrid
    dx = 1.0 / n
    nz = nozzle(n)
    phis = phi_stream(steps + warmup)
    dev = DeviceGrid(3, n, n, n, GAMMA)
    try:
        dev.upload("P", jet_state(n))
        dev.set_nozzle(nz["radius"], dx, dx, nz["rho"], nz["u"], nz["umag"], nz["eps"])
        if ref:
            dt = 0.4 / (3.0 / dx)
            dev.set_step(ref_mode=True, rec="wenoz", riemann="hll", integrator="rk2",
                         precision=precision, n_iter=50, tol=1e-10, floor=1e-7, dt=dt,
                         dx=dx, dy=dx, dz=dx, bc=BC_JET)
        else:
            dev.fill_ghosts(BC_JET)
            dev.set_step(ref_mode=False, rec="wenoz", riemann="hllc", integrator="rk3",
                         precision=precision, n_iter=50, tol=1e-10, floor=1e-7, cfl=0.4,
                         dx=dx, dy=dx, dz=dx, bc=BC_JET)
        dev.prim_to_cons()
        for s in range(warmup):
            dev.step(phi=phis[s])
        dev.sync()
        if timing:
            # HIP events around the directional sweeps' launches only: every
            # event record holds the stream a little, and bracketing all
            # kernels (ghost fill, rescue, dt) costs 0.7 % of the step
            dev.timing(2)
        t0 = time.perf_counter()
        for s in range(steps):
            dev.step(phi=phis[warmup + s])
        dev.sync()
        el = time.perf_counter() - t0
        rep = dev.timing_report() if timing else {}
        if timing:
            dev.timing(False)
        nbad, nresc = dev.counters()
    finally:
        dev.close()
    return el, rep, nbad, nresc


def run_kh(n=2048, steps=2000, warmup=3, precision="fp64fast", fp32_error=True):
    """2D relativistic Kelvin-Helmholtz, n^2 periodic, WENOZ + HLLC + SSP-RK3, CFL 0.4,
    for the configured 2000 steps: the rate over all of them (after `warmup`
    untimed steps from the same initial state), unrecovered cells, and the FP32
    form's relative L1 error per variable against the exact FP64 form after
    the same number of steps."""
    from verona_amd import DeviceGrid
    from verona_amd.problems import kelvin_helmholtz
    dx = 1.0 / n
    bc = ["periodic"] * 4 + ["none", "none"]
    P0 = kelvin_helmholtz(n, gamma=GAMMA)

    def run(prec, nsteps, warm, timed):
        with DeviceGrid(2, n, n, 1, GAMMA) as dev:
            dev.upload("P", P0)
            dev.fill_ghosts(bc)
            dev.prim_to_cons()
            dev.set_step(ref_mode=False, rec="wenoz", riemann="hllc", integrator="rk3", precision=prec,
                         n_iter=50, tol=1e-10, floor=1e-7, cfl=0.4, dx=dx, dy=dx, bc=bc)
            for _ in range(warm):
                dev.step()
            if warm:
                # the timed run starts from the initial state again
                dev.upload("P", P0)
                dev.fill_ghosts(bc)
                dev.prim_to_cons()
            dev.sync()
            t0 = time.perf_counter()
            t = 0.0
            for _ in range(nsteps):
                if timed:
                    dev.step()
                else:
                    t += dev.step(want_dt=True)
            dev.sync()
            el = time.perf_counter() - t0
            nbad = dev.counters()[0]
            Pn = None if timed else dev.download("P")[:, 3:-3, 3:-3]
        return el, nbad, Pn, t

    el, nbad, _, _ = run(precision, steps, warmup, True)
    out = {"what": f"2D Kelvin-Helmholtz {n}^2 periodic, WENOZ+HLLC+SSP-RK3, CFL 0.4, {precision}, "
                   f"all {steps} configured steps",
           "value": n * n * steps / el, "unit": "zone-updates/s", "steps": steps,
           "ms_per_step": el / steps * 1e3, "n_unrecovered": nbad}
    if fp32_error:
        _, nb64, A, t64 = run("fp64", steps, 0, False)
        _, nb32, B, t32 = run("fp32", steps, 0, False)
        out["fp32_error_at_t_end"] = {
            "what": f"FP32 recon/Riemann vs exact FP64 after {steps} steps (each form's own CFL dt)",
            "t_end_fp64": t64, "t_end_fp32": t32, "n_unrecovered": [nb64, nb32],
            "rel_l1_per_var": dict(zip(["rho", "u", "u^x", "u^y"],
                                       [float(np.abs(B[v] - A[v]).sum() / np.abs(A[v]).sum()) for v in range(4)]))}
    return out


def run_jet_developed(n=384, advance=700, steps=20, precision="fp64fast", log=None):
    """The jet after `advance` untimed steps (the beam, its bow shock and
    cocoon fill a third of the grid: shocked cells at Gamma up to 10), then
    `steps` timed steps; the early window (steps 6-25 from rest) is the
    headline's.  The method (WENO-Z + HLLC + SSP-RK3, CFL 0.4, no positivity
    limiter, as Verona's) produces an unrecoverable state on the jet axis
    after ~830 steps at 384^3, so the window stays before that."""
    say = log or (lambda m: None)
    from verona_amd import DeviceGrid
    dx = 1.0 / n
    nz = nozzle(n)
    phis = phi_stream(advance + steps)
    with DeviceGrid(3, n, n, n, GAMMA) as dev:
        dev.upload("P", jet_state(n))
        dev.set_nozzle(nz["radius"], dx, dx, nz["rho"], nz["u"], nz["umag"], nz["eps"])
        dev.fill_ghosts(BC_JET)
        dev.set_step(ref_mode=False, rec="wenoz", riemann="hllc", integrator="rk3", precision=precision,
                     n_iter=50, tol=1e-10, floor=1e-7, cfl=0.4, dx=dx, dy=dx, dz=dx, bc=BC_JET)
        dev.prim_to_cons()
        t = 0.0
        for s in range(advance):
            t += dev.step(phi=phis[s], want_dt=True)
            if (s + 1) % 100 == 0:
                say(f"jet developed: step {s + 1}, t {t:.4f}")
        dev.sync()
        resc0 = dev.counters()[1]
        dev.timing(True)
        t0 = time.perf_counter()
        for s in range(steps):
            dev.step(phi=phis[advance + s])
        dev.sync()
        el = time.perf_counter() - t0
        rep = dev.timing_report()
        dev.timing(False)
        nbad, nresc = dev.counters()
        Pz = dev.download("P")[4, 3:-3, 3:-3, 3:-3]
    lor = np.sqrt(1.0 + Pz * Pz)  # longitudinal part of the Lorentz factor
    head = np.flatnonzero((lor > 1.5).any(axis=(0, 1)))
    return {"what": f"jet {n}^3, {precision}, {steps} timed steps after {advance} untimed steps",
            "value": n ** 3 * steps / el, "unit": "zone-updates/s", "ms_per_step": el / steps * 1e3,
            "t_at_window": t, "jet_head_z_cells": int(head.max()) + 1 if head.size else 0,
            "n_unrecovered": nbad, "n_rescued_in_window": nresc - resc0, "n_rescued_total": nresc,
            "kernels_ms": {k: {"launches": c, "mean_ms": m} for k, (c, m) in rep.items()}}


def _tube_subset(sel, lengths, rho, u, ux, dt, dx):
    """The inputs of the tubes sel (indices into the batch), as their own batch."""
    off = np.concatenate([[0], np.cumsum(lengths)])
    idx = np.concatenate([np.arange(off[t], off[t + 1]) for t in sel])
    return lengths[sel], rho[idx], u[idx], ux[idx], dt[sel], dx[sel]


def run_tubes(ntubes=4096, steps=500, precision="fp64", keep=False, subset=None):
    """Batched uneven 1D shock tubes: 4096 tubes, lengths U{256..8192} (SplitMix64 seed
    42), PPM + HLLC, RK2, dt = 0.4 dx per tube, one persistent hipGraph launch with the
    autotuned launch shape.  value = cells x steps / s of the timed launch.  subset:
    run only these tubes of the batch, as a batch of their own.  The autotuner
    probes the state half-way through the run (steps / 2 untimed steps first)."""
    from verona_amd.tubes import Tubes, shock_tube_batch
    batch = shock_tube_batch(ntubes, 256, 8192, seed=42)
    if subset is not None:
        batch = _tube_subset(subset, *batch)
    lengths, rho, u, ux, dt, dx = batch
    with Tubes(lengths) as tb:
        tb.upload(rho, u, ux, dt, dx)
        tb.set_method("ppm", "hllc", 50, 1e-10)
        tb.set_precision(precision)
        tb.run(steps // 2, 256, 4 * tb.cu_count)  # a developed state for the autotuner
        block, grid = tb.autotune(probe_steps=2)
        tb.upload(rho, u, ux, dt, dx)
        tb.run(2, block, grid)  # graph capture + warm-up
        tb.upload(rho, u, ux, dt, dx)
        ms = tb.run(steps, block, grid)
        r, uu, vx = tb.download()
    cells = int(lengths.sum())
    off = np.concatenate([[0], np.cumsum(lengths)])
    # Verona1D's con2prim has no floors: tubes whose random states it cannot
    # recover end with NaN markers (the oracle marks the same cells)
    finite = np.array([bool(np.isfinite(r[off[t]:off[t + 1]]).all() and np.isfinite(uu[off[t]:off[t + 1]]).all()
                            and np.isfinite(vx[off[t]:off[t + 1]]).all()) for t in range(len(lengths))])
    out = {"what": f"{len(lengths)} uneven 1D tubes ({cells} cells), PPM+HLLC, RK2, {steps} steps, "
                   "one persistent hipGraph launch", "form": precision, "value": cells * steps / (ms * 1e-3),
           "unit": "cell-steps/s", "ms": ms, "launch_shape": [block, grid],
           "tubes_with_nan_markers": int((~finite).sum()), "_finite": finite}
    if keep:
        out["_state"] = (off, np.stack([r, uu, vx]))
    return out


def run_tubes_both(ntubes=4096, steps=500):
    """The tubes leg in the exact FP64 form and in the fast FP64 form.  The
    line's value is the rate over the tubes that stay finite (the batch minus
    the tubes Verona1D's floor-free Newton leaves with NaN markers, re-run as a
    batch of their own, so no cell skips its con2prim); the all-tubes rates
    are kept beside it.  The fast form's relative L1 difference from the exact
    one is taken over the tubes both leave finite."""
    ex = run_tubes(ntubes, steps, "fp64", keep=True)
    fa = run_tubes(ntubes, steps, "fp64fast", keep=True)
    off, A = ex.pop("_state")
    _, B = fa.pop("_state")
    fin_ex, fin_fa = ex.pop("_finite"), fa.pop("_finite")
    ok = fin_ex & fin_fa
    m = np.repeat(ok, np.diff(off))
    rel = [float(np.abs(B[v][m] - A[v][m]).sum() / np.abs(A[v][m]).sum()) for v in range(3)]
    sel = np.flatnonzero(fin_ex)
    exf = run_tubes(ntubes, steps, "fp64", subset=sel)
    faf = run_tubes(ntubes, steps, "fp64fast", subset=sel)
    for d in (exf, faf):
        d.pop("_finite")
    keys = ("value", "unit", "ms", "launch_shape", "tubes_with_nan_markers")
    out = {"what": f"{len(sel)} of {ntubes} uneven 1D tubes (those Verona1D's con2prim keeps finite), "
                   f"PPM+HLLC, RK2, {steps} steps, one persistent hipGraph launch",
           "form": "fp64", "value": exf["value"], "unit": "cell-steps/s", "ms": exf["ms"],
           "launch_shape": exf["launch_shape"], "tubes": int(len(sel)),
           "tubes_with_nan_markers": exf["tubes_with_nan_markers"],
           "all_tubes": {"what": ex["what"], **{k: ex[k] for k in keys}},
           "fast_form": {"value": faf["value"], "unit": "cell-steps/s", "ms": faf["ms"],
                         "launch_shape": faf["launch_shape"],
                         "tubes_with_nan_markers": faf["tubes_with_nan_markers"],
                         "all_tubes": {k: fa[k] for k in keys},
                         "rel_l1_vs_exact": rel, "tubes_compared": int(ok.sum())}}
    return out


def fp32_error(n, steps):
    """Relative L1 per variable of the FP32 reconstruction/Riemann form against the
    all-FP64 (exact) form after `steps` steps of the jet from the same state."""
    out = {}
    for prec in ("fp64", "fp32"):
        from verona_amd import DeviceGrid
        dx = 1.0 / n
        nz = nozzle(n)
        phis = phi_stream(steps)
        with DeviceGrid(3, n, n, n, GAMMA) as dev:
            dev.upload("P", jet_state(n))
            dev.set_nozzle(nz["radius"], dx, dx, nz["rho"], nz["u"], nz["umag"], nz["eps"])
            dev.fill_ghosts(BC_JET)
            dev.set_step(ref_mode=False, rec="wenoz", riemann="hllc", integrator="rk3", precision=prec,
                         n_iter=50, tol=1e-10, floor=1e-7, cfl=0.4, dx=dx, dy=dx, dz=dx, bc=BC_JET)
            dev.prim_to_cons()
            for s in range(steps):
                dev.step(phi=phis[s])
            out[prec] = dev.download("P")[:, 3:-3, 3:-3, 3:-3]
    a, b = out["fp32"], out["fp64"]
    err = [float(np.abs(a[v] - b[v]).sum() / np.abs(b[v]).sum()) for v in range(5)]
    return {"what": f"FP32 recon/Riemann vs all-FP64, jet {n}^3 after {steps} SSP-RK3 steps",
            "rel_l1_per_var": dict(zip(["rho", "u", "u^x", "u^y", "u^z"], err))}


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(budget_s=12.0, budget_ref_s=6.0, n=128, all_configs=True):
    """CPU oracle (plain-C restatement of Verona's algorithm, OpenMP) on a
    bounded sample of the same workload: an n^3 grid, steps until about
    budget_s seconds of wall time have elapsed -- the benchmark's method
    (WENOZ + HLLC + SSP-RK3, CFL), then Verona's own (WENOZ + HLL, midpoint RK2,
    fixed dt) for budget_ref_s."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    os.environ.setdefault("OMP_NUM_THREADS", str(cores))
    dx = 1.0 / n
    nz = nozzle(n)
    phis = phi_stream(4096)
    bc, keep = O.make_bc3([2, 2, 2, 2, 3, 2], dx=dx, dy=dx,
                          nozzle=(nz["radius"], nz["rho"], nz["u"], nz["umag"], nz["eps"], 0.0),
                          phi_seq=phis)

    def timed(mode, budget):
        Po = np.ascontiguousarray(jet_state(n).T)
        steps = 0
        t0 = time.perf_counter()
        while True:
            O.run3(Po, mode, bc, 2)
            steps += 2
            el = time.perf_counter() - t0
            if el > budget or steps >= 400:
                return steps, el

    ext = O.make_mode(ref=False, rec="wenoz", riemann="hllc", integrator="rk3", n_iter=50,
                      tol=1e-10, gamma=GAMMA, floor=1e-7, cfl=0.4, dx=dx, dy=dx, dz=dx)
    steps, el = timed(ext, budget_s)
    ref = O.make_mode(ref=True, rec="wenoz", riemann="hll", integrator="rk2", n_iter=50,
                      tol=1e-10, gamma=GAMMA, floor=1e-7, dt=0.4 / (3.0 / dx), dx=dx, dy=dx, dz=dx)
    steps_r, el_r = timed(ref, budget_ref_s)
    configs = cpu_baselines_by_config(O, cores) if all_configs else None
    return {"value": n ** 3 * steps / el, "unit": "zone-updates/s", "cores": cores, "kind": "port",
            "sample": f"CPU oracle (C restatement of Verona's algorithm, OpenMP) on the same jet "
                      f"workload at {n}^3, {steps} SSP-RK3 WENOZ+HLLC steps, {el:.1f} s",
            "ref_semantics_value": n ** 3 * steps_r / el_r,
            "ref_semantics_sample": f"same, Verona's method (WENOZ+HLL, midpoint RK2, fixed dt), "
                                    f"{steps_r} steps, {el_r:.1f} s",
            "cpu_model": _cpu_model(), "nproc": os.cpu_count(),
            "what": "CPU restatement of Verona's algorithm (the oracle), not Verona itself",
            "configs": configs}


def run_c1_gpu():
    """The 1D relativistic shock tube on the GPU (Verona1D.HARM_HLL over one
    tube): N = 1024, rho 10 / 1, p 13.33 / 1e-6, minmod + HLLC, RK2,
    dt = 0.4 dx, outflow (frozen) edges, to t = 0.4 (1025 steps)."""
    from verona_amd.problems import shock_tube
    from verona_amd.tubes import Tubes
    n = 1024
    rho, u, ux = shock_tube(n)
    dx = 1.0 / n
    steps, t = 0, 0.0
    while t < 0.4:
        t += 0.4 * dx
        steps += 1
    with Tubes([n]) as tb:
        tb.upload(rho, u, ux, [0.4 * dx], [dx])
        tb.set_method("minmod", "hllc", 50, 1e-10)
        # one workgroup of 1024 lanes: each stage is one 1024-cell chunk
        tb.run(steps, 1024, 1)  # graph capture + warm-up
        tb.upload(rho, u, ux, [0.4 * dx], [dx])
        ms = tb.run(steps, 1024, 1)
    return {"what": "1D shock tube N=1024, minmod+HLLC, RK2, to t=0.4, one tube (one workgroup)",
            "steps": steps, "ms": ms, "value": n * steps / (ms * 1e-3), "unit": "cell-steps/s"}


def cpu_baselines_by_config(O, cores, budget_s=4.0):
    """The CPU oracle (C restatement of Verona's algorithm) on the other
    configurations, each on a bounded sample, all host cores: C1 in full, C2
    (Kelvin-Helmholtz 2048^2) for the steps that fit the budget, C4 (2^24
    con2prim cells) from the generic and the warm start, C5 (the 4,096
    tubes) for a fixed number of steps with the tubes spread over the cores."""
    import ctypes
    from concurrent.futures import ThreadPoolExecutor

    from verona_amd.c2p import c4_states, warm_guess
    from verona_amd.problems import kelvin_helmholtz, shock_tube
    from verona_amd.tubes import shock_tube_batch
    out = {}
    print("cpu baseline C1", file=sys.stderr, flush=True)
    # C1: the whole run (the 1D oracle is Verona1D's loop, threaded inside con2prim)
    n = 1024
    rho, u, ux = shock_tube(n)
    P1 = np.ascontiguousarray(np.stack([rho, u, ux]))
    t0 = time.perf_counter()
    steps, _ = O.run1(P1, 0.4 / n, 1.0 / n, 0.4, 5.0 / 3.0, "minmod", "hllc", 50, 1e-10)
    el = time.perf_counter() - t0
    out["C1_shock_tube"] = {"value": n * steps / el, "unit": "cell-steps/s", "seconds": el,
                            "sample": f"the whole run: 1024 cells, minmod+HLLC, RK2, {steps} steps to t=0.4"}
    print("cpu baseline C2", file=sys.stderr, flush=True)
    # C2: Kelvin-Helmholtz 2048^2, the benchmark method, steps within the budget
    n = 2048
    P2 = np.ascontiguousarray(kelvin_helmholtz(n, gamma=GAMMA).T)
    m2 = O.make_mode(ref=False, rec="wenoz", riemann="hllc", integrator="rk3", n_iter=50, tol=1e-10,
                     gamma=GAMMA, floor=1e-7, cfl=0.4, dx=1.0 / n, dy=1.0 / n)
    steps, t0 = 0, time.perf_counter()
    while True:
        O.run2(P2, m2, [1, 1, 1, 1], 1)
        steps += 1
        el = time.perf_counter() - t0
        if el > budget_s:
            break
    out["C2_kelvin_helmholtz"] = {"value": n * n * steps / el, "unit": "zone-updates/s", "seconds": el,
                                  "sample": f"2048^2 periodic, WENOZ+HLLC+SSP-RK3, CFL 0.4, {steps} steps "
                                            "from the initial state"}
    del P2
    print("cpu baseline C4", file=sys.stderr, flush=True)
    # C4: 2^24 held-out cells, the 1DW Newton from the generic and the warm start
    n = 1 << 24
    U, P = c4_states(n, seed=42, gamma=GAMMA)
    res = {}
    for name, Pg in (("generic", np.tile([1.0, 1.0, 0.0, 0.0, 0.0], (n, 1))), ("warm", warm_guess(P))):
        W0 = O.c2p4_w0_from_prims(Pg, GAMMA)
        t0 = time.perf_counter()
        _, it, cv = O.c2p4_solve(U, W0, GAMMA, 100, 1e-12)
        el = time.perf_counter() - t0
        res[name] = {"value": n / el, "unit": "cells/s", "seconds": el, "iters_mean": float(it.mean()),
                     "not_converged": int((~cv).sum())}
    del U, P
    out["C4_con2prim"] = {**res, "sample": "2^24 cells of the benchmark law (seed 42), 1DW Newton to "
                                           "|dW| < 1e-12 W, solve only"}
    print("cpu baseline C5", file=sys.stderr, flush=True)
    # C5: every tube, a fixed number of steps, tubes spread over the cores
    # (one OpenMP thread per tube run)
    lengths, rho, u, ux, dt, dx = shock_tube_batch(4096, 256, 8192, seed=42)
    off = np.concatenate([[0], np.cumsum(lengths)])
    nst = 4
    gomp = ctypes.CDLL("libgomp.so.1")

    def one(t):
        gomp.omp_set_num_threads(1)
        Pt = np.ascontiguousarray(np.stack([rho[off[t]:off[t + 1]], u[off[t]:off[t + 1]], ux[off[t]:off[t + 1]]]))
        O.run1(Pt, dt[t], dx[t], 1e30, 5.0 / 3.0, "ppm", "hllc", 50, 1e-10, max_steps=nst)

    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=cores) as ex:
        list(ex.map(one, np.argsort(-lengths)))
    el = time.perf_counter() - t0
    out["C5_tubes"] = {"value": int(lengths.sum()) * nst / el, "unit": "cell-steps/s", "seconds": el,
                       "sample": f"all 4096 tubes ({int(lengths.sum())} cells), PPM+HLLC, RK2, {nst} steps"}
    return out


def run_c2p(n_cells=1 << 24, n_train=1 << 22, train_steps=20000, reps=5, log=None, quick=False):
    """Learned-guess con2prim vs plain Newton on held-out cells of the benchmark
    law (seed 42): every guess finished by the same FP64 1DW Newton to
    |dW| < 1e-12 W.  Training pairs: the solver's own dumps (2D Kelvin-Helmholtz,
    1D blast wave) plus synthetic cells drawn with seed 7."""
    from verona_amd.c2p import Con2PrimBatch, c4_states, solver_pairs, training_set, warm_guess
    say = log or (lambda m: None)
    if quick:
        # profiling sequence: short training on synthetic pairs, each guess solved twice
        n_train, train_steps, reps, dumps = 1 << 18, 800, 1, None
    else:
        say("c2p: solver dumps")
        dumps = solver_pairs()
    Ut, Pt = training_set(n_train, seed=7, gamma=GAMMA, dumps=dumps)
    say("c2p: held-out cells")
    U, P = c4_states(n_cells, seed=42, gamma=GAMMA)
    Pprev = warm_guess(P)
    out = {"n_cells": n_cells, "rtol": 1e-12, "n_iter": 100, "gamma": GAMMA,
           "held_out": "rho ~ logU[1e-4,1e2], p ~ logU[1e-6,1e2], Gamma ~ logU[1,100], isotropic, seed 42",
           "warm": "previous primitives = true P x (1 + 0.01 N(0,1)) per component, seed 43",
           "generic": "P = (1, 1, 0, 0, 0)", "analytic": "W0 = E = -T^t_t"}
    with Con2PrimBatch(n_cells, GAMMA) as cp:
        say("c2p: training")
        curve, tms = cp.train(Ut, Pt, steps=train_steps, batch=8192, lr=3e-3, seed=7)
        out["train"] = {"pairs": int(len(Ut)),
                        "solver_dump_pairs": int(min(len(dumps[0]), n_train // 4)) if dumps else 0,
                        "steps": train_steps, "batch": 8192, "ms": tms,
                        "loss_first": float(curve[:10].mean()), "loss_last": float(curve[-100:].mean())}
        cp.upload(U, Pprev)
        res = {}
        for g in ("generic", "warm", "analytic", "learned"):
            say(f"c2p: solve {g}")
            cp.solve(g)
            runs = [cp.solve(g) for _ in range(reps)]
            st = min(runs, key=lambda r: r["ms"])
            st.pop("hist")
            st["ms_median"] = float(np.median([r["ms"] for r in runs]))
            st["cells_per_s"] = n_cells / (st["ms_median"] * 1e-3)
            res[g] = st
        if quick:
            return {"guesses": res}
        Pl = cp.download()
        cp.solve("analytic")
        Pa = cp.download()
        relL1 = (np.abs(Pl - Pa).sum(0) / np.abs(Pa).sum(0)).tolist()
        # cells above the reference's Gamma <= 50 cap, on their own
        lor = np.sqrt(1.0 + (P[:, 2:] ** 2).sum(1))
        hi = lor > 50.0
        cp.upload(U[hi], Pprev[hi])
        sub = {}
        for g in ("generic", "warm", "analytic", "learned"):
            r = cp.solve(g)
            sub[g] = {k: r[k] for k in ("n", "n_unrecovered", "n_fallback", "iters_mean", "iters_p99",
                                        "iters_max", "wave_max_mean")}
    for g in res:
        res[g]["speedup_learned"] = res["learned"]["cells_per_s"] / res[g]["cells_per_s"]
    out["guesses"] = res
    out["learned_vs_analytic_rel_l1"] = relL1
    out["gamma_above_50"] = sub
    bytes_cell = 80
    t_s = res["learned"]["ms_median"] * 1e-3
    ach = bytes_cell * n_cells / t_s / 1e9
    # The solve is VALU-issue bound, not HBM bound: rocprofv3 counts 700 VALU
    # wave-instructions per 64-cell chunk for the learned guess at 4.0 wave-max
    # Newton iterations (quick network) and 103 per iteration, so ~600 at the
    # full network's 3.0 (profiles/c2p_solve_2e24_counters.md).  A wave64
    # instruction holds a 16-lane SIMD for 4 cycles: 256 CUs x 4 SIMDs x
    # 2.4 GHz / 4 per second.
    valu_chunk = 700.0 - 103.0 * (4.0 - res["learned"]["wave_max_mean"])
    issue = valu_chunk * (n_cells / 64) / t_s / (256 * 4 * 2.4e9 / 4)
    out["roofline"] = {"bound": "valu_issue", "kernel": "k_solve<learned>", "bytes_per_cell": bytes_cell,
                       "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                       "valu_issue_frac_est": issue,
                       "source": "profiles/c2p_solve_2e24_counters.md"}
    return out


def copy_bandwidth(n, reps=5):
    """Device-to-device copy rate of one n^3 state slot (read + write bytes per
    second): the practical streaming ceiling the sweeps are compared with,
    measured on the same GPU in the same run."""
    from verona_amd import DeviceGrid
    with DeviceGrid(3, n, n, n, GAMMA) as dev:
        nbytes = dev.slot_bytes()
        dev.copy("U", "P")
        dev.sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            dev.copy("U", "P")
        dev.sync()
        el = time.perf_counter() - t0
    return 2.0 * nbytes * reps / el / 1e9


def _traffic_profile(n, precision):
    """The committed rocprofv3 FETCH_SIZE/WRITE_SIZE passes (copy-calibrated) of the
    jet's three sweeps, when they were taken on this workload and form; else None."""
    path = os.path.join(ROOT, "profiles", "jet384_rk3_hbm_traffic.json")
    if n != 384 or precision != "fp64fast" or not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("kernels", {})


def measured_traffic(kernel, n, precision):
    """HBM bytes per launch of `kernel` from the committed counter passes, or None."""
    k = (_traffic_profile(n, precision) or {}).get(kernel)
    return None if k is None else float(k["bytes"])


# The kernels of one RK stage of the fused step, by arithmetic form: both
# FP64 forms run the x direction inside the y march.
STAGE_KERNELS = {"fp64fast": ("xysweep", "zsweep_c2p"), "fp64": ("xysweep", "zsweep_c2p")}


def measured_step_traffic(n, precision):
    """Measured HBM bytes per zone-update of the whole step: the stage kernels'
    calibrated bytes per interior cell per launch, each launched once per RK
    stage, times 3 stages (the rescue and ghost passes are < 1 % of it)."""
    kern = _traffic_profile(n, precision)
    ks = STAGE_KERNELS.get(precision, ())
    if not kern or not all(k in kern for k in ks):
        return None
    return 3.0 * sum(kern[k]["per_cell"] for k in ks)


def valu_issue_ceiling(n, precision):
    """Zone-updates/s the step's instruction count allows at full VALU issue:
    the stage kernels' VALU lane-instructions per interior cell per launch
    (committed rocprofv3 SQ_INSTS_VALU pass), three launches each per step,
    against 256 CUs x 4 SIMDs x 2.4 GHz / 4 cycles per wave64 instruction."""
    path = os.path.join(ROOT, "profiles", "jet384_sweeps_instruction_mix.json")
    if n != 384 or precision != "fp64fast" or not os.path.exists(path):
        return None
    with open(path) as f:
        mix = json.load(f)["per_cell_lane_instructions"]
    ks = STAGE_KERNELS[precision]
    if not all(k in mix for k in ks):
        return None
    lane_ops = 3.0 * sum(mix[k]["SQ_INSTS_VALU"] for k in ks)
    return {"valu_lane_ops_per_zone_update": lane_ops,
            "zone_updates_per_s": 256 * 4 * 2.4e9 / 4 * 64 / lane_ops}


def fp64_roofline(kernel, cells, mean_ms, n, precision):
    """FP64 VALU view of the dominant kernel: FLOPs per cell (2 FMA + MUL + ADD +
    transcendental, lane counts from the committed rocprofv3 SQ_INSTS_* pass) over its
    event-measured launch time, against MI355X's FP64 vector peak, and the VALU issue
    fraction (a wave64 instruction occupies a 16-lane SIMD for 4 cycles)."""
    path = os.path.join(ROOT, "profiles", "jet384_sweeps_instruction_mix.json")
    if n != 384 or precision != "fp64fast" or not os.path.exists(path):
        return None
    with open(path) as f:
        mix = json.load(f)["per_cell_lane_instructions"].get(kernel)
    if mix is None:
        return None
    flop = (2 * mix["SQ_INSTS_VALU_FMA_F64"] + mix["SQ_INSTS_VALU_MUL_F64"] + mix["SQ_INSTS_VALU_ADD_F64"]
            + mix["SQ_INSTS_VALU_TRANS_F64"])
    tf = flop * cells / (mean_ms * 1e-3) / 1e12
    issue_peak = 256 * 4 * 2.4e9 / 4  # wave instructions per second
    issue = mix["SQ_INSTS_VALU"] / 64 * cells / (mean_ms * 1e-3) / issue_peak
    return {"bound": "fp64_valu", "kernel": kernel, "flop_per_cell": flop, "achieved": tf,
            "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": tf / FP64_PEAK_TFLOPS,
            "valu_issue_frac": issue, "source": "profiles/jet384_sweeps_instruction_mix.md"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=384)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--no-c2p", action="store_true", help="skip the learned con2prim leg")
    ap.add_argument("--c2p-only", action="store_true", help="only the learned con2prim leg")
    ap.add_argument("--c2p-quick", action="store_true",
                    help="with --c2p-only: the short profiling sequence (800 training steps, 2 solves per guess)")
    ap.add_argument("--tubes-only", action="store_true", help="only the batched 1D tubes leg")
    ap.add_argument("--kh-only", action="store_true", help="only the 2D Kelvin-Helmholtz 2048^2 leg")
    ap.add_argument("--tubes-steps", type=int, default=500, help="steps of the batched tubes leg")
    ap.add_argument("--kh-steps", type=int, default=40, help="steps of the --kh-only leg")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 3 device copies of one 384^3 slot (known bytes) for counter calibration")
    ap.add_argument("--precision", default="fp64fast", choices=["fp64", "fp64fast"],
                    help="FP64 form of the headline run (both agree with the oracle to <= 1e-12)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: there is no multi-GPU path (the grid would shard by "
              "Cartesian domain decomposition with 3-cell halos, one exchange per RK stage).",
              file=sys.stderr)
        sys.exit(2)
    n, K, W = args.n, args.steps, args.warmup
    if args.copy_calibration:
        from verona_amd import DeviceGrid
        with DeviceGrid(3, n, n, n, GAMMA) as dev:
            for _ in range(3):
                dev.copy("Fx", "P")
            dev.sync()
            print(json.dumps({"copy_bytes_read": dev.slot_bytes(), "copy_bytes_written": dev.slot_bytes(),
                              "copies": 3}), flush=True)
        return
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    if args.c2p_only:
        print(json.dumps({"con2prim": run_c2p(log=log, quick=args.c2p_quick)}), flush=True)
        return
    if args.tubes_only:
        print(json.dumps({"batched_tubes": run_tubes_both(steps=args.tubes_steps)}), flush=True)
        return
    if args.kh_only:
        print(json.dumps({"kelvin_helmholtz": run_kh(steps=args.kh_steps, fp32_error=False)}), flush=True)
        return
    cells = n ** 3
    el, rep, nbad, nresc = run_gpu(n, K, W, ref=False, precision=args.precision)
    value = cells * K / el
    ms_step = el / K * 1e3
    # dominant kernel by total time in the timed region (HIP events on the library stream)
    tot = {k: c * m for k, (c, m) in rep.items()}
    dom = max(tot, key=tot.get)
    cnt, mean_ms = rep[dom]
    # Algorithmic bytes per launch of the dominant kernel, per interior cell,
    # averaged over the 3 RK stages.  Exact form: the z sweep closes each
    # stage and forms the RK combination, so it carries the stage's
    # compulsory traffic: 560 B per zone-update over 3 stages (stage 1: read
    # P, U^n, write U, P = 160 B; stages 2-3 also read U^n = 200 B; mean
    # 560/3 B); its x-y march (or the x and y sweeps) read P and U_b/acc and
    # write acc (120 B, plus the edge stencils' 3.75 B).
    # Fast form: the fused x-y march reads P and U_b, in stages 2-3 also U^n
    # (it forms a U^n + b U_b), and writes acc (120 B + 2/3 x 40 B), plus P of
    # the 3 cells either side of each 64-column wave that its edge-face
    # stencils reach (6 cells x 40 B per 64 cells = 3.75 B); the z sweep
    # reads P and acc and writes U and P (160 B).
    if args.precision == "fp64fast":
        per_cell = {"zsweep_c2p": 4 * 40, "xysweep": 3 * 40 + 2 * 40 / 3 + 6 * 40 / 64}.get(dom, 4 * 40)
    else:
        per_cell = {"zsweep_c2p": BYTES_PER_ZU["rk3"] / 3.0, "xsweep": 3 * 40, "ysweep": 3 * 40,
                    "xysweep": 3 * 40 + 6 * 40 / 64}.get(dom, 4 * 40)
    ach = per_cell * cells / (mean_ms * 1e-3) / 1e9
    step_gbs = BYTES_PER_ZU["rk3"] * value / 1e9
    step_traffic = measured_step_traffic(n, args.precision)
    ceiling = valu_issue_ceiling(n, args.precision)
    copy_gbs = copy_bandwidth(n)
    fp64v = fp64_roofline(dom, cells, mean_ms, n, args.precision)
    # which of the two the dominant kernel sits closer to: its VALU issue
    # fraction (instruction-mix profile) or its HBM rate over the device copy
    limited_by = ("valu_issue" if fp64v is not None and copy_gbs and fp64v["valu_issue_frac"] > ach / copy_gbs
                  else "hbm")
    out = {
        "metric": "zone-updates/s (3D WENOZ+HLLC, 1 GPU); con2prim cells/s learned vs Newton",
        "value": value,
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "value_reference_formula": (n + 6) ** 3 * K / el,  # ghosts counted, as 3dim/algos.jl:347 prints
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"3D relativistic jet {n}^3, WENOZ+HLLC+SSP-RK3, CFL 0.4, FP64",
                   "fp64_form": args.precision,
                   "grid": [n, n, n], "gamma": GAMMA, "nozzle_radius_cells": 8,
                   "lorentz": 10.0, "bc": "nozzle z-low, outflow elsewhere"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": ach, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                     "measured_copy_gbs": copy_gbs, "frac_of_copy": ach / copy_gbs,
                     "traffic": measured_traffic(dom, n, args.precision),
                     "traffic_source": "profiles/jet384_rk3_hbm_traffic.md (FETCH_SIZE x2.0 + WRITE_SIZE x1.0, copy-calibrated)",
                     "limited_by": limited_by,
                     "valu_issue_frac": None if fp64v is None else fp64v["valu_issue_frac"],
                     "bytes_per_cell_per_launch": per_cell, "mean_launch_ms": mean_ms,
                     "launches": cnt, "algorithmic_bytes": per_cell * cells,
                     # the metric itself: the whole step at 560 B per zone-update
                     "step": {"bytes_per_zone_update": BYTES_PER_ZU["rk3"], "achieved": step_gbs,
                              "frac": step_gbs / HBM_PEAK_GBS,
                              "measured_traffic_bytes_per_zone_update": step_traffic,
                              "traffic_over_algorithmic": (None if step_traffic is None
                                                           else step_traffic / BYTES_PER_ZU["rk3"]),
                              "hbm_ceiling_zone_updates_per_s": HBM_PEAK_GBS * 1e9 / BYTES_PER_ZU["rk3"],
                              "valu_issue_ceiling": ceiling}},
        "roofline_fp64": fp64v,
        "roofline_step": {"bytes_per_zone_update": BYTES_PER_ZU["rk3"], "achieved": step_gbs,
                          "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": step_gbs / HBM_PEAK_GBS,
                          "measured_traffic_bytes_per_zone_update": step_traffic,
                          "valu_issue_ceiling": ceiling},
        "kernels_ms": {k: {"launches": c, "mean_ms": m} for k, (c, m) in rep.items()},
        "n_unrecovered": nbad,
        "n_rescued": nresc,
    }
    if not args.no_extra:
        other = "fp64" if args.precision == "fp64fast" else "fp64fast"
        el1, _, nb1, _ = run_gpu(n, max(K // 2, 4), 2, ref=False, precision=other, timing=False)
        out["other_fp64_form"] = {"form": other, "value": cells * max(K // 2, 4) / el1,
                                  "unit": "zone-updates/s", "n_unrecovered": nb1}
        el2, rep2, nb2, _ = run_gpu(n, max(K // 2, 4), 2, ref=True, timing=False)
        out["ref_form"] = {"what": "Verona semantics: WENOZ+HLL, midpoint RK2, fixed dt", "form": "fp64fast",
                           "value": cells * max(K // 2, 4) / el2, "unit": "zone-updates/s",
                           "n_unrecovered": nb2}
        el2x, _, nb2x, _ = run_gpu(n, max(K // 2, 4), 2, ref=True, precision="fp64", timing=False)
        out["ref_form_exact"] = {"what": "Verona semantics, FP64 bit-identical to the oracle", "form": "fp64",
                                 "value": cells * max(K // 2, 4) / el2x, "unit": "zone-updates/s",
                                 "n_unrecovered": nb2x}
        el3, _, nb3, _ = run_gpu(n, max(K // 2, 4), 2, ref=False, precision="fp32", timing=False)
        out["fp32_form"] = {"what": "reconstruction + Riemann in FP32, con2prim + accumulation FP64",
                            "value": cells * max(K // 2, 4) / el3, "unit": "zone-updates/s",
                            "n_unrecovered": nb3}
        log("fp32 error")
        out["fp32_form"]["error_vs_fp64"] = fp32_error(n, 10)
        log("kelvin-helmholtz")
        out["kelvin_helmholtz"] = run_kh()
        log("jet, developed flow")
        out["jet_developed"] = run_jet_developed(n, precision=args.precision, log=log)
        log("shock tube C1")
        out["shock_tube_c1"] = run_c1_gpu()
        log("batched tubes")
        out["batched_tubes"] = run_tubes_both()
    if not args.no_c2p:
        out["con2prim"] = run_c2p(log=log)
    out["cpu_baseline"] = None if args.no_cpu else cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
