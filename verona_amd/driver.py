"""Run loop shared by the 2D and 3D drivers (Verona2D/3D.HARM_HLL).

Follows src/3dim/algos.jl:203-362 (2D: src/2dim/algos.jl:192-324): halo
exchange and PtoU at the start, dump 0, then ``while t < Tmax``: boundary
callback, one fused time step on the GPU, ``t += dt``, and a dump whenever
``t`` passes the next multiple of ``drops``.  Returns the dump count.
"""
import os
import time

import numpy as np

from .bc import Boundary, Cart


def _kinds(comm, fun_bound, ndim):
    periodic = getattr(comm, "periodic", (False,) * ndim) if comm is not None else (False,) * ndim
    periodic = tuple(periodic) + (False,) * (3 - len(periodic))
    if isinstance(fun_bound, Boundary):
        return fun_bound.kinds(periodic), None
    kinds = ["periodic" if periodic[k // 2] else "none" for k in range(6)]
    return kinds, (fun_bound if callable(fun_bound) else None)


def save_dump(path, Phost, t, grid, step=0):
    """Interior primitives as in SaveHDF5Parallel (src/3dim/saving.jl:57-146):
    dataset ``data`` (Julia shape (nvar, Nx, Ny[, Nz])), ``T`` and ``grid``;
    raises on NaN like src/3dim/saving.jl:70-72.  ``step`` (the number of
    steps taken) is added so that a run can restart from the dump."""
    sl = (slice(None),) + (slice(3, -3),) * (Phost.ndim - 1)
    data = np.ascontiguousarray(Phost[sl])
    if np.isnan(data).any():
        raise RuntimeError("Nan in matrix")
    np.savez(path, data=data, T=np.float64(t), grid=np.asarray(grid, dtype=np.float64),
             step=np.int64(step))


def load_dump(path):
    """(data, T, grid, step, index) of a dump written by save_dump; the index
    is the i of ``dump<i>.npz``."""
    with np.load(path, allow_pickle=False) as z:
        data, T, grid = z["data"], float(z["T"]), tuple(float(g) for g in z["grid"])
        step = int(z["step"]) if "step" in z.files else None
    base = os.path.basename(path)
    if not (base.startswith("dump") and base.endswith(".npz")):
        raise ValueError(f"not a dump file name: {base}")
    return data, T, grid, step, int(base[4:-4])


def _h5py():
    try:
        import h5py
    except ImportError as e:  # not in every environment; the npz dumps need nothing
        raise ImportError("HDF5 dumps need h5py, which is not installed here; the run loop "
                          "writes dump<i>.npz (same datasets) and dump_to_hdf5 converts them "
                          "where h5py is available") from e
    return h5py


def dump_to_hdf5(npz_path, h5_path=None):
    """Write a dump<i>.npz as the HDF5 file SaveHDF5Parallel writes
    (src/3dim/saving.jl:57-146) and SaveHDF5Gather (src/2dim/saving.jl:2-45)
    write: dataset ``data`` with the
    dimensions reversed as HDF5.jl stores a Julia (nvar, Nx, Ny[, Nz]) array
    (C shape (Nz, Ny, Nx, nvar), variable fastest: the same bytes as Julia's
    column-major array), scalar ``T`` and ``grid`` ([dx, dy(, dz)]); ``step``
    is kept for restarts.  Returns the path written."""
    h5py = _h5py()
    data, T, grid, step, _ = load_dump(npz_path)
    h5_path = h5_path or npz_path[:-4] + ".h5"
    with h5py.File(h5_path, "w") as f:
        f.create_dataset("data", data=np.ascontiguousarray(data.T))
        f.create_dataset("T", data=np.float64(T))
        f.create_dataset("grid", data=np.asarray(grid, dtype=np.float64))
        if step is not None:
            f.create_dataset("step", data=np.int64(step))
    return h5_path


def load_hdf5_dump(h5_path):
    """(data, T, grid) of an HDF5 dump in the reference's layout (see
    dump_to_hdf5), ``data`` back in Julia's (nvar, Nx, Ny[, Nz]) index order."""
    h5py = _h5py()
    with h5py.File(h5_path, "r") as f:
        data = np.ascontiguousarray(np.asarray(f["data"]).T)
        T = float(np.asarray(f["T"]))
        grid = tuple(float(g) for g in np.asarray(f["grid"]))
    return data, T, grid


def harm_hll(dev, ndim, comm, P_host_view, dt, grid, Tmax, eos, drops, floor, out_dir, verbose,
             n_iter, tol, fun_bound, *, rec="wenoz", riemann="hll", integrator="rk2", cfl=0.0,
             precision="fp64", ref_mode=True, max_steps=None, restart=None):
    """restart: path of a dump<i>.npz written by an earlier run of the same
    problem; the run continues from its state, time, dump index and step
    count (the reference has no restart; this is the natural one for its
    dump format)."""
    kinds, callback = _kinds(comm if comm is not None else Cart(), fun_bound, ndim)
    t = 0.0
    i = 0
    steps = 0
    if restart is not None:
        data, t, g0, step0, i = load_dump(restart)
        if step0 is None:
            raise ValueError("dump has no step count; it was not written by this driver")
        if any(abs(a - b) > 0 for a, b in zip(g0, grid)):
            raise ValueError(f"dump grid {g0} differs from the run's {tuple(grid)}")
        Ph = dev.download("P")
        sl = (slice(None),) + (slice(3, -3),) * ndim
        if Ph[sl].shape != data.shape:
            raise ValueError(f"dump shape {data.shape} differs from the grid's {Ph[sl].shape}")
        Ph[sl] = data
        dev.upload("P", Ph)
        steps = step0
    nozzle = getattr(fun_bound, "nozzle", None)
    phi_seq = None
    if nozzle is not None:
        dev.set_nozzle(nozzle["radius"], grid[0], grid[1], nozzle["rho"], nozzle["u"],
                       nozzle["umag"], nozzle["eps"])
        phi_seq = nozzle.get("phi_seq")
    dx, dy = grid[0], grid[1]
    dz = grid[2] if ndim == 3 else 1.0
    # start: halo exchange (periodic axes) [+ boundaries in the extended form], PtoU
    dev.fill_ghosts([k if k == "periodic" else "none" for k in kinds])
    if not ref_mode:
        dev.fill_ghosts(kinds)
    dev.prim_to_cons()
    dev.set_step(ref_mode=ref_mode, rec=rec, riemann=riemann, integrator=integrator,
                 precision=precision, n_iter=n_iter, tol=tol, floor=floor, dt=dt, cfl=cfl,
                 dx=dx, dy=dy, dz=dz,
                 bc=[("none" if (callback is not None and k != "periodic") else k) for k in kinds])
    coords = (0,) * ndim
    dims = (1,) * ndim
    if out_dir is not None:
        os.makedirs(out_dir, exist_ok=True)
        if restart is None:
            save_dump(os.path.join(out_dir, f"dump{i}.npz"), dev.download("P"), t, grid)
    # the threshold sequence drops, 2 drops, ... accumulated in FP64 as the
    # straight run accumulates it
    thres = drops
    for _ in range(i):
        thres += drops
    t0 = time.time()
    steps_at_print = steps
    ncell = int(np.prod([s for s in dev.shape[1:]]))
    while t < Tmax:
        if max_steps is not None and steps >= max_steps:
            break
        if callback is not None:
            Ph = dev.download("P")
            if ndim == 3:
                callback(Ph, t, coords, dims)
            else:
                callback(Ph, t)
            dev.upload("P", Ph)
        phi = phi_seq[steps] if phi_seq is not None else 0.0
        dt_used = dev.step(phi=phi, want_dt=cfl > 0)
        t += dt_used if cfl > 0 else dt
        steps += 1
        if t > thres:
            i += 1
            thres += drops
            if verbose:
                dev.sync()
                el = time.time() - t0
                print(f"{t:.3g} elapsed: {el:.3g} s")
                # the reference prints cells * drops / dt / elapsed (src/3dim/algos.jl:344-352),
                # i.e. cells x steps taken since the last print for a fixed dt; the step
                # count is what it means, and is also defined in CFL mode (dt = 0)
                print(f"speed: {ncell * (steps - steps_at_print) / el * 1e-6:.3g} zones [10^6/s]")
                t0 = time.time()
                steps_at_print = steps
            if out_dir is not None:
                save_dump(os.path.join(out_dir, f"dump{i}.npz"), dev.download("P"), t, grid, steps)
    dev.sync()
    return i
