"""julia/VeronaHIP.jl against include/verona_hip.h (CPU only, no Julia).

Every ccall in the shim names an exported symbol with the header's arity and
argument types, the shim's StepParams has vrn_step_params's field order and
types, and HARM_HLL's loop body issues the reference's kernel sequence
(src/3dim/algos.jl:274-329): Fluxes x3, Update(Uhalf, dt/2), UtoP(Uhalf),
Limit, halo, Fluxes x3, Update(U, dt), UtoP(U), Limit, halo.
"""
import ctypes

import julia_shim
from verona_amd import _lib

# Julia ccall type -> the C parameter types it may bind
JL_TO_C = {
    "Cint": {"int"},
    "Cdouble": {"double"},
    "Int64": {"int64_t"},
    "Ptr{Cvoid}": {"vrn_ctx *ctx", "vrn_ctx*"},
    "Ref{Ptr{Cvoid}}": {"vrn_ctx **out"},
    "Ptr{Float64}": {"const double *host", "double *host"},
    "Ref{Int64}": {"int64_t *n_unrecovered", "int64_t *n_rescued", "int64_t *bytes"},
    "Ref{Cdouble}": {"double *dt_used"},
    "Ref{NTuple{6,Cint}}": {"const int bc[6]"},
    "Ref{StepParams}": {"const vrn_step_params *p"},
    "Cstring": {"const char *"},
}


def _c_type_of(decl):
    """'double dt' -> 'double'; pointer and array declarators are kept whole."""
    parts = decl.split()
    if "*" in decl or "[" in decl:
        return decl
    return " ".join(parts[:-1])


def test_every_ccall_matches_the_header():
    protos = julia_shim.header_prototypes()
    calls = julia_shim.ccalls()
    assert len(calls) >= 12
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name, ret, args in calls:
        assert name in protos, f"{name} is not declared in include/verona_hip.h"
        assert hasattr(lib, name), f"{name} is not exported by libverona_hip.so"
        cret, cargs = protos[name]
        assert (ret, cret) in {("Cint", "int"), ("Cstring", "const char *")}, (name, ret, cret)
        assert len(args) == len(cargs), (name, args, cargs)
        for jl, c in zip(args, cargs):
            ct = _c_type_of(c)
            assert ct in JL_TO_C.get(jl, set()), f"{name}: Julia {jl} cannot bind C '{c}'"


def test_step_params_layout_matches_the_header():
    cfields = julia_shim.step_params_fields()
    jfields = julia_shim.julia_struct_fields("StepParams")
    jl_to_c = {"Cint": "int", "Cdouble": "double", "NTuple{6,Cint}": "int[6]"}
    assert [n for _, n in jfields] == [n for _, n in cfields]
    assert [jl_to_c[t] for t, _ in jfields] == [t for t, _ in cfields]


def test_run_loop_issues_the_reference_sequence():
    seq = julia_shim.step_sequence()
    names = [f for f, _ in seq]
    stage = ["Fluxes"] * 3 + ["Update", "UtoP", "Limit", "fill_ghosts"]
    assert names == stage + stage
    assert [a[2] for f, a in seq if f == "Fluxes"] == ["1", "2", "3", "1", "2", "3"]
    upd = [a for f, a in seq if f == "Update"]
    assert [u[1] for u in upd] == ["SLOT_UHALF", "SLOT_U"]
    assert [u[2].replace(" ", "") for u in upd] == ["dt/2", "dt"]
    assert [a[1] for f, a in seq if f == "UtoP"] == ["SLOT_UHALF", "SLOT_U"]
