# VeronaHIP.jl -- Verona's device work on one MI355X through libverona_hip.so.
#
# Drop-in for the GPU run loop of Verona3D (src/3dim/algos.jl:203-362): the
# problem setup, the `while t < Tmax` loop, the dump cadence and the HDF5
# output stay Verona's; the five KernelAbstractions kernel objects built at
# src/3dim/algos.jl:251-257 and the halo exchange of src/3dim/sync_CUDA.jl
# become `ccall`s into the C ABI declared in include/verona_hip.h.
#
#   using Verona, MPI
#   include("julia/VeronaHIP.jl")
#   P = Verona3D.ParVector3D{Float64}(Nx, Ny, Nz)          # host array, filled by the problem setup
#   n = VeronaHIP.HARM_HLL(comm, P, (1, 1, 1), (8, 8, 4), dt, dx, dy, dz, Tmax,
#                          Verona.EosTypes.Polytrope{Float64}(4/3), drops, 1e-7, "out", true,
#                          50, 1e-10, fun_bound)
#
# `HARM_HLL` issues the reference's per-kernel sequence call for call;
# `HARM_HLL_fused` runs the same loop with one fused `vrn_step` per time step.
# One rank per GPU: MPI_dims must be (1, 1, 1) and periodic faces come from
# the Cartesian communicator's `periods`, as the reference's single-rank
# self-exchange does.
module VeronaHIP

using MPI
using ..Verona: EosTypes, Verona3D

const lib = joinpath(@__DIR__, "..", "verona_amd", "libverona_hip.so")

# include/verona_hip.h
const SLOT_P, SLOT_U, SLOT_UHALF, SLOT_FX, SLOT_FY, SLOT_FZ = 0, 1, 2, 3, 4, 5
const REC_MINMOD, REC_PPM, REC_WENOZ = 0, 1, 2
const RIEMANN_HLL, RIEMANN_HLLC = 0, 1
const INT_RK2, INT_RK3 = 0, 1
const PREC_FP64, PREC_FP32, PREC_FP64_FAST = 0, 1, 2
const BC_NONE, BC_PERIODIC, BC_OUTFLOW, BC_NOZZLE = 0, 1, 2, 3

last_error() = unsafe_string(ccall((:vrn_last_error, lib), Cstring, ()))
check(rc, what) = rc == 0 || error(what, ": ", last_error())

mutable struct Ctx
    h::Ptr{Cvoid}
end

# replaces VectorLike(P) x 5 and the 12 halo buffers (src/3dim/algos.jl:224-243)
function Ctx(Nx::Integer, Ny::Integer, Nz::Integer, eos::EosTypes.Polytrope{Float64})
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:vrn_create, lib), Cint, (Cint, Cint, Cint, Cint, Cdouble, Ref{Ptr{Cvoid}}),
                3, Nx, Ny, Nz, eos.gamma, h), "vrn_create")
    c = Ctx(h[])
    finalizer(c) do c
        c.h == C_NULL || ccall((:vrn_destroy, lib), Cint, (Ptr{Cvoid},), c.h)
        c.h = C_NULL
    end
    return c
end

upload!(c::Ctx, slot, A::Array{Float64,4}) =
    check(ccall((:vrn_upload, lib), Cint, (Ptr{Cvoid}, Cint, Ptr{Float64}), c.h, slot, A), "vrn_upload")
download!(c::Ctx, slot, A::Array{Float64,4}) =
    check(ccall((:vrn_download, lib), Cint, (Ptr{Cvoid}, Cint, Ptr{Float64}), c.h, slot, A), "vrn_download")
sync(c::Ctx) = check(ccall((:vrn_sync, lib), Cint, (Ptr{Cvoid},), c.h), "vrn_sync")

# ---- the five kernel objects of HARM_HLL, same argument meaning
# PtoU(P.arr, U.arr, eos)                                     src/3dim/algos.jl:259
PtoU(c::Ctx) = check(ccall((:vrn_prim_to_cons, lib), Cint, (Ptr{Cvoid},), c.h), "PtoU")
# Fluxes(P.arr, eos, floor, F.arr, dim), dim = 1, 2, 3 (x, y, z)   :280-282
Fluxes(c::Ctx, floor, dim; rec = REC_WENOZ, riemann = RIEMANN_HLL) =
    check(ccall((:vrn_fluxes, lib), Cint, (Ptr{Cvoid}, Cint, Cdouble, Cint, Cint),
                c.h, dim - 1, floor, rec, riemann), "Fluxes")
# Update(U.arr, Ubuff.arr, dt, dx, dy, dz, Fx.arr, Fy.arr, Fz.arr), Ubuff = SLOT_UHALF or SLOT_U   :285, :311
Update(c::Ctx, dst, dt, dx, dy, dz) =
    check(ccall((:vrn_update, lib), Cint, (Ptr{Cvoid}, Cint, Cdouble, Cdouble, Cdouble, Cdouble),
                c.h, dst, dt, dx, dy, dz), "Update")
# UtoP(src.arr, P.arr, eos, n_iter, tol), warm-started from the P in the context   :290, :318
UtoP(c::Ctx, src, n_iter, tol) =
    check(ccall((:vrn_cons_to_prim, lib), Cint, (Ptr{Cvoid}, Cint, Cint, Cdouble), c.h, src, n_iter, tol), "UtoP")
# Limit(P.arr, floor)                                         :292, :320
Limit(c::Ctx, floor) = check(ccall((:vrn_limit, lib), Cint, (Ptr{Cvoid}, Cdouble), c.h, floor), "Limit")
# SendBoundaryX/Y/Z + WaitForBoundary on one rank (src/3dim/sync_CUDA.jl:3-129)
fill_ghosts(c::Ctx, bc::NTuple{6,Cint}) =
    check(ccall((:vrn_fill_ghosts, lib), Cint, (Ptr{Cvoid}, Ref{NTuple{6,Cint}}), c.h, bc), "fill_ghosts")

function counters(c::Ctx)
    bad, resc = Ref{Int64}(0), Ref{Int64}(0)
    check(ccall((:vrn_get_counters, lib), Cint, (Ptr{Cvoid}, Ref{Int64}, Ref{Int64}), c.h, bad, resc),
          "vrn_get_counters")
    return bad[], resc[]
end

# vrn_step_params, same field order and types
struct StepParams
    ref_mode::Cint
    rec::Cint
    riemann::Cint
    integrator::Cint
    precision::Cint
    n_iter::Cint
    tol::Cdouble
    floor_::Cdouble
    dt::Cdouble
    cfl::Cdouble
    dx::Cdouble
    dy::Cdouble
    dz::Cdouble
    bc::NTuple{6,Cint}
    nozzle_phi::Cdouble
end

function step!(c::Ctx, p::StepParams)
    dt_used = Ref{Cdouble}(0.0)
    check(ccall((:vrn_step, lib), Cint, (Ptr{Cvoid}, Ref{StepParams}, Ref{Cdouble}), c.h, p, dt_used),
          "vrn_step")
    return dt_used[]
end

# periodic faces of the Cartesian communicator: the single-rank form of the
# reference's neighbour exchange (a non-periodic face has MPI.PROC_NULL)
function periodic_bc(comm)
    _, periods, _ = MPI.Cart_get(comm)
    b(k) = periods[k] != 0 ? Cint(BC_PERIODIC) : Cint(BC_NONE)
    return (b(1), b(1), b(2), b(2), b(3), b(3))
end

# fun_bound(P, t, coords, dims) mutates a host array; it runs on a host copy
# of P (download, call, upload) -- drop-in, PCIe-bound.  Zero-gradient
# callbacks are faster as declarative VRN_BC_OUTFLOW faces in HARM_HLL_fused.
function host_callback!(c::Ctx, Phost, fun_bound, t, idx_mpi, MPI_dims)
    download!(c, SLOT_P, Phost.arr)
    fun_bound(Phost, t, idx_mpi, MPI_dims)
    upload!(c, SLOT_P, Phost.arr)
end

function dump!(c::Ctx, comm, Phost, MPI_dims, out_dir, i, t, dx, dy, dz)
    download!(c, SLOT_P, Phost.arr)
    to_save = Dict("T" => t, "grid" => [dx, dy, dz])
    name = out_dir * "/dump" * string(i) * ".h5"
    Verona3D.SaveHDF5Parallel(comm, Phost, MPI_dims[1], MPI_dims[2], MPI_dims[3], name, to_save)
end

"""
    HARM_HLL(comm, P, MPI_dims, Size, dt, dx, dy, dz, Tmax, eos, drops, floor, out_dir, verbose, n_iter, tol[, fun_bound])

src/3dim/algos.jl:203-362 with every kernel launch and halo exchange as a
call into libverona_hip.so.  `P` is the host `ParVector3D` (its array is
uploaded once; dumps download it).  `Size` (the KA workgroup) is accepted and
ignored.  Returns the dump count, as the reference does.
"""
function HARM_HLL(comm, P::Verona3D.ParVector3D{Float64}, MPI_dims::Tuple{Int64,Int64,Int64},
                  Size::Tuple{Int64,Int64,Int64}, dt::Float64, dx::Float64, dy::Float64, dz::Float64,
                  Tmax::Float64, eos::EosTypes.Polytrope{Float64}, drops::Float64, floor::Float64 = 1e-7,
                  out_dir::String = ".", verbose::Bool = false, kwargs...)
    MPI_dims == (1, 1, 1) || error("VeronaHIP: one rank per GPU, MPI_dims must be (1, 1, 1)")
    n_iter, tol = kwargs[1], kwargs[2]
    fun_bound = length(kwargs) > 2 ? kwargs[3] : nothing
    bc = periodic_bc(comm)
    c = Ctx(P.size_X - 6, P.size_Y - 6, P.size_Z - 6, eos)
    upload!(c, SLOT_P, P.arr)
    fill_ghosts(c, bc)                                         # :240-243
    PtoU(c)                                                    # :259
    t = 0.0
    thres_to_dump = drops
    i = 0
    t0 = time()
    dump!(c, comm, P, MPI_dims, out_dir, i, t, dx, dy, dz)     # :268-270
    idx_mpi = MPI.Cart_coords(comm)
    while t < Tmax
        fun_bound === nothing || host_callback!(c, P, fun_bound, t, idx_mpi, MPI_dims)   # :276
        for d in 1:3
            Fluxes(c, floor, d)                                # :280-282
        end
        Update(c, SLOT_UHALF, dt / 2, dx, dy, dz)              # :285
        UtoP(c, SLOT_UHALF, n_iter, tol)                       # :290
        Limit(c, floor)                                        # :292
        fill_ghosts(c, bc)                                     # :297-300
        for d in 1:3
            Fluxes(c, floor, d)                                # :307-309
        end
        Update(c, SLOT_U, dt, dx, dy, dz)                      # :311
        UtoP(c, SLOT_U, n_iter, tol)                           # :318
        Limit(c, floor)                                        # :320
        fill_ghosts(c, bc)                                     # :324-327
        t += dt
        if t > thres_to_dump
            i += 1
            thres_to_dump += drops
            sync(c)
            if verbose
                elapsed = time() - t0
                println(round(t, sigdigits = 3), " elapsed: ", round(elapsed, sigdigits = 3), " s")
                println("speed: ", round(P.size_X * P.size_Y * P.size_Z * drops / dt / elapsed * 1e-6,
                                         sigdigits = 3), " zones [10^6/s]")
            end
            t0 = time()
            dump!(c, comm, P, MPI_dims, out_dir, i, t, dx, dy, dz)
        end
    end
    download!(c, SLOT_P, P.arr)
    finalize(c)
    return i
end

"""
    HARM_HLL_fused(...; rec, riemann, integrator, precision, cfl, bc)

The same loop with one `vrn_step` per time step: ref_mode = 1 with the
defaults reproduces HARM_HLL (WENOZ, HLL, midpoint RK2, fixed dt, the
boundary callback at the step start); the keywords select the extended
method (HLLC, SSP-RK3, CFL dt, declarative faces, the FP64-fast or FP32
arithmetic).
"""
function HARM_HLL_fused(comm, P::Verona3D.ParVector3D{Float64}, MPI_dims, Size, dt, dx, dy, dz, Tmax,
                        eos::EosTypes.Polytrope{Float64}, drops, floor = 1e-7, out_dir = ".",
                        verbose = false, kwargs...; ref_mode = true, rec = REC_WENOZ,
                        riemann = RIEMANN_HLL, integrator = INT_RK2, precision = PREC_FP64,
                        cfl = 0.0, bc = nothing)
    MPI_dims == (1, 1, 1) || error("VeronaHIP: one rank per GPU, MPI_dims must be (1, 1, 1)")
    n_iter, tol = kwargs[1], kwargs[2]
    fun_bound = length(kwargs) > 2 ? kwargs[3] : nothing
    # declarative faces (BC_OUTFLOW, BC_NOZZLE, ...) or the communicator's
    # periodic faces; with a host callback the non-periodic faces are its own
    faces = bc === nothing ? periodic_bc(comm) : NTuple{6,Cint}(bc)
    c = Ctx(P.size_X - 6, P.size_Y - 6, P.size_Z - 6, eos)
    upload!(c, SLOT_P, P.arr)
    fill_ghosts(c, periodic_bc(comm))
    ref_mode || fill_ghosts(c, faces)
    PtoU(c)
    t, thres_to_dump, i = 0.0, drops, 0
    dump!(c, comm, P, MPI_dims, out_dir, i, t, dx, dy, dz)
    idx_mpi = MPI.Cart_coords(comm)
    while t < Tmax
        fun_bound === nothing || host_callback!(c, P, fun_bound, t, idx_mpi, MPI_dims)
        p = StepParams(ref_mode ? 1 : 0, rec, riemann, integrator, precision, n_iter, tol, floor, dt, cfl,
                       dx, dy, dz, faces, 0.0)
        t += step!(c, p)
        if t > thres_to_dump
            i += 1
            thres_to_dump += drops
            verbose && println(round(t, sigdigits = 3))
            dump!(c, comm, P, MPI_dims, out_dir, i, t, dx, dy, dz)
        end
    end
    download!(c, SLOT_P, P.arr)
    finalize(c)
    return i
end

end # module
