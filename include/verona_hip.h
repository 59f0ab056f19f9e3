/*
 * verona_hip.h -- C ABI of libverona_hip.so, the MI355X (gfx950) hot path of
 * the Verona SRHD solver.
 *
 * The reference has no plugin API; its operator boundary is the five
 * KernelAbstractions kernel objects that Verona3D.HARM_HLL builds and calls
 * (src/3dim/algos.jl:251-257; 2D: src/2dim/algos.jl:235-239):
 *
 *   Fluxes(P, eos, floor, F, dim)            src/3dim/algos.jl:280-282
 *   Update(U, Ubuff, dt, dx, dy, dz, Fx..)   src/3dim/algos.jl:285, 311
 *   UtoP(U, P, eos, n_iter, tol)             src/3dim/algos.jl:290, 318
 *   PtoU(P, U, eos)                          src/3dim/algos.jl:259
 *   Limit(P, floor)                          src/3dim/algos.jl:292, 320
 *
 * plus the halo exchange of src/3dim/sync_CUDA.jl:3-129, which on one GPU
 * becomes an on-device ghost fill.  Each is exported below with the same
 * argument meaning; vrn_step() is the fused path (one call per time step).
 *
 * Conventions
 *   - All functions return 0 on success and a negative VRN_E* code on error;
 *     vrn_last_error() gives the message (thread-local).  Nothing aborts and
 *     no C++ exception crosses this boundary.
 *   - Host arrays use Verona's byte order: Julia arr[var, i, j, k] column
 *     major, variable fastest, 3 ghost cells per side (3D: 5 x (nx+6) x
 *     (ny+6) x (nz+6) doubles; 2D: 4 x (nx+6) x (ny+6)).  Host pointers are
 *     only borrowed for the duration of the call.
 *   - A context owns its device buffers and one HIP stream; calls enqueue
 *     work and return, except downloads, vrn_sync and counter reads, which
 *     block.  A context is not re-entrant; distinct contexts are independent.
 *     Per-step calls allocate nothing.
 */
#ifndef VERONA_HIP_H
#define VERONA_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vrn_ctx vrn_ctx;

/* error codes */
#define VRN_OK 0
#define VRN_E_ARG (-1)
#define VRN_E_HIP (-2)
#define VRN_E_STATE (-3)
#define VRN_E_NOMEM (-4)

/* reconstruction: src/fluxlimiter.jl MINMOD :1-13, PPM :74-180, WENOZ :19-71 */
#define VRN_REC_MINMOD 0
#define VRN_REC_PPM 1
#define VRN_REC_WENOZ 2
/* Riemann solver: HLL (src/3dim/algos.jl:146-174) or HLLC with HLL fallback */
#define VRN_RIEMANN_HLL 0
#define VRN_RIEMANN_HLLC 1
/* time integrator: midpoint RK2 (src/3dim/algos.jl:279-327) or SSP-RK3 */
#define VRN_INT_RK2 0
#define VRN_INT_RK3 1
/* arithmetic of reconstruction + Riemann solve (con2prim and the conserved
 * accumulation are always FP64) */
#define VRN_PREC_FP64 0      /* operation order of the reference, bit-reproducible */
#define VRN_PREC_FP32 1      /* FP32 reconstruction + Riemann solve */
#define VRN_PREC_FP64_FAST 2 /* FP64 with FMA contraction and reciprocal division */
/* boundary kinds, per face */
#define VRN_BC_NONE 0     /* ghosts left as they are (MPI.PROC_NULL neighbour) */
#define VRN_BC_PERIODIC 1 /* single-rank periodic halo exchange */
#define VRN_BC_OUTFLOW 2  /* zero-gradient copy (examples/Jet3D TurnOff) */
#define VRN_BC_NOZZLE 3   /* z-low only: outflow plus jet nozzle inflow */
/* buffer slots of a context (the arrays HARM_HLL allocates, src/3dim/algos.jl:224-228);
 * the flux slots are allocated (zeroed) by the first call that uses them --
 * vrn_fluxes, vrn_update, or an upload/download/copy naming them -- since
 * vrn_step never does */
#define VRN_SLOT_P 0
#define VRN_SLOT_U 1
#define VRN_SLOT_UHALF 2
#define VRN_SLOT_FX 3
#define VRN_SLOT_FY 4
#define VRN_SLOT_FZ 5

/* Create a solver context for an ndim (2 or 3) grid of nx*ny(*nz) interior
 * cells (nz ignored for 2D) and adiabatic index gamma
 * (src/eos.jl:7-9 Polytrope).  Replaces VectorLike(P) x5 + the halo buffers
 * of src/3dim/algos.jl:224-243. */
int vrn_create(int ndim, int nx, int ny, int nz, double gamma, vrn_ctx **out);
int vrn_destroy(vrn_ctx *ctx);
/* Block until all work enqueued on the context has finished. */
int vrn_sync(vrn_ctx *ctx);
const char *vrn_last_error(void);

/* Copy a whole array (ghosts included, Verona byte order) to/from a slot.
 * Replaces CuParVector3D{T}(P) (src/3dim/structs.jl:49-51) and Array(view)
 * (src/3dim/saving.jl:68). */
int vrn_upload(vrn_ctx *ctx, int slot, const double *host);
int vrn_download(vrn_ctx *ctx, int slot, double *host);

/* Device-to-device copy of a whole slot (no reference counterpart: a
 * snapshot utility, also the known-size copy used to calibrate HBM byte
 * counters); vrn_slot_bytes gives the bytes one slot occupies on the device. */
int vrn_copy(vrn_ctx *ctx, int dst_slot, int src_slot);
int vrn_slot_bytes(vrn_ctx *ctx, int64_t *bytes);

/* PtoU(P, U, eos): U = conserved(P) over the whole array,
 * src/3dim/structs.jl:66-81 (2D: src/2dim/structs.jl:62-76). */
int vrn_prim_to_cons(vrn_ctx *ctx);

/* Fluxes(P, eos, floor, F_dim, dim): reconstructed, floored face states and
 * the Riemann flux at face i+1/2 along dim (0, 1, 2) into slot FX/FY/FZ,
 * src/3dim/algos.jl:14-176 (2D: src/2dim/algos.jl:11-168).  The reference
 * hard-codes rec = WENOZ, riemann = HLL. */
int vrn_fluxes(vrn_ctx *ctx, int dim, double floor_, int rec, int riemann);

/* Update(U, dst, dt, dx, dy, dz, Fx, Fy, Fz): dst = U - dt/dx dFx - dt/dy dFy
 * - dt/dz dFz on interior cells, src/3dim/algos.jl:178-200; dst_slot is
 * VRN_SLOT_UHALF or VRN_SLOT_U (in place, as at :311). */
int vrn_update(vrn_ctx *ctx, int dst_slot, double dt, double dx, double dy, double dz);

/* UtoP(src, P, eos, n_iter, tol): con2prim on interior cells, warm-started
 * from the P already in the context.  3D: the 1DW -> 2D -> 5D -> bisection
 * cascade of src/3dim/structs.jl:218-628; 2D: the 4x4 Newton of
 * src/2dim/structs.jl:249-330.  Unrecovered cells get NaN (as the
 * reference) and are counted, see vrn_get_counters. */
int vrn_cons_to_prim(vrn_ctx *ctx, int src_slot, int n_iter, double tol);

/* Limit(P, floor): rho, u = max(., floor) over the whole array,
 * src/3dim/algos.jl:5-12. */
int vrn_limit(vrn_ctx *ctx, double floor_);

/* Ghost fill of slot P with the per-face boundary kinds bc[6] = {xlo, xhi,
 * ylo, yhi, zlo, zhi}, applied axis by axis (x, y, z).  Periodic faces are
 * the single-rank form of SendBoundaryX/Y/Z + WaitForBoundary
 * (src/3dim/sync_CUDA.jl:3-129); OUTFLOW is the zero-gradient copy of the
 * examples' fun_bound callbacks (examples/Jet3D/HARM_jet_breakout.jl:136-157). */
int vrn_fill_ghosts(vrn_ctx *ctx, const int bc[6]);

/* Jet nozzle used by VRN_BC_NOZZLE on the z-low face: ghost cells within
 * `radius` of the (x, y) centre hold (rho, u, |u^i| * n) with
 * n = normalize(e_z + eps (cos phi, sin phi, 0)); phi is per call/step. */
int vrn_set_nozzle(vrn_ctx *ctx, double radius, double dx, double dy, double rho, double u,
                   double umag, double eps);

/* Counters accumulated since the last reset: cells left unrecovered (NaN
 * marker) and cells that needed the con2prim cascade beyond 1DW.  Blocks. */
int vrn_get_counters(vrn_ctx *ctx, int64_t *n_unrecovered, int64_t *n_rescued);
int vrn_reset_counters(vrn_ctx *ctx);

typedef struct {
    int ref_mode;   /* 1: Verona sequencing (BC callback at step start only) */
    int rec;        /* VRN_REC_* */
    int riemann;    /* VRN_RIEMANN_* */
    int integrator; /* VRN_INT_* */
    int precision;  /* VRN_PREC_* */
    int n_iter;     /* con2prim iteration limit (kwargs[1]) */
    double tol;     /* con2prim tolerance (kwargs[2]) */
    double floor_;  /* density / internal energy floor (a number, not -0:
                       vrn_step returns VRN_E_ARG otherwise) */
    double dt;      /* fixed dt; ignored when cfl > 0 */
    double cfl;     /* > 0: dt = cfl / max_cells sum_d max|lambda_d|/dx_d */
    double dx, dy, dz;
    int bc[6];
    double nozzle_phi; /* nozzle precession angle for this step */
} vrn_step_params;

/* One full time step on the context's P (src/3dim/algos.jl:274-329):
 * per stage one fused kernel per direction (the last one also runs
 * con2prim + Limit; 3D FP64 forms: x and y in one kernel), then the
 * ghost fill.  In the 3D FP64 forms with the x-y fusion and a rule on
 * both x faces the ghost fill leaves P's x faces to be filled by the next
 * entry point that reads P's ghost cells (the fused kernel reads them
 * through the rules), so callers always see complete ghost cells.
 * *dt_used (may be NULL) receives the step's dt; in CFL mode that costs one
 * small device-to-host read. */
int vrn_step(vrn_ctx *ctx, const vrn_step_params *p, double *dt_used);

/* Per-kernel timing with HIP events on the context's stream.  enable = 1:
 * every launch of the fused path is bracketed by events; 2: only the
 * directional sweeps' launches (the x, y, x-y and z kernels), which puts
 * fewer event records into the stream; 0: off.  vrn_timing_report fills up
 * to `cap` entries with the kernel name, launch count and mean duration in
 * milliseconds.  Blocks. */
int vrn_timing_enable(vrn_ctx *ctx, int enable);
int vrn_timing_report(vrn_ctx *ctx, char *names, int name_len, int64_t *counts, double *mean_ms,
                      int cap, int *n_out);

/* Work partition of the fused sweeps (no counterpart in the reference; it
 * changes no arithmetic, every face and cell is computed once either way):
 * x sweep: rows_per_wave whole rows per wavefront, or, with xseg > 1, one of
 * xseg segments of a row per wavefront (then rows_per_wave must be 1);
 * y / z sweeps: seg_len interior cells per marching segment.  vrn_create
 * chooses a partition for the grid; vrn_get_partition reads it back. */
int vrn_set_partition(vrn_ctx *ctx, int rows_per_wave, int xseg, int seg_len);
int vrn_get_partition(vrn_ctx *ctx, int *rows_per_wave, int *xseg, int *seg_len);

/* 3D, VRN_PREC_FP64 and VRN_PREC_FP64_FAST: whether the x direction runs
 * inside the y march (on = 1, the default) or as its own sweep (0).  No
 * counterpart in the reference and no change of arithmetic: both solve the
 * same faces and apply the same updates in the same order, bit for bit; the
 * fused form moves 116 B less per cell and stage.  The FP32 form and 2D
 * always use the separate x sweep. */
int vrn_set_xy_fusion(vrn_ctx *ctx, int on);


/* ------------------------------------------------------------------------
 * Batched 1D shock tubes: many independent Verona1D.HARM_HLL problems
 * (src/1dim/algos.jl:58-315, SoA ParVector1D without ghosts, the 3x3 Newton
 * con2prim of src/1dim/structs.jl:78-105) advanced together.  The reference
 * has no batched form; each tube reproduces one HARM_HLL run step for step.
 * ------------------------------------------------------------------------ */
typedef struct vrn_tubes vrn_tubes;

/* ntubes tubes of lengths[t] >= 8 cells, adiabatic index gamma. */
int vrn_tubes_create(int ntubes, const int64_t *lengths, double gamma, vrn_tubes **out);
int vrn_tubes_destroy(vrn_tubes *tb);
const char *vrn_tubes_last_error(void);
/* Concatenated SoA primitives (tubes back to back) and per-tube dt, dx:
 * the arguments P, dt, dx of HARM_HLL; U = PtoU(P) is formed on the device
 * (src/1dim/algos.jl:82). */
int vrn_tubes_upload(vrn_tubes *tb, const double *rho, const double *u, const double *ux,
                     const double *dt, const double *dx);
int vrn_tubes_download(vrn_tubes *tb, double *rho, double *u, double *ux);
/* rec (VRN_REC_*), riemann (VRN_RIEMANN_*), n_iter and tol (kwargs of HARM_HLL). */
int vrn_tubes_set_method(vrn_tubes *tb, int rec, int riemann, int n_iter, double tol);
/* Arithmetic of the tube kernels: VRN_PREC_FP64 (default; Verona1D's FP64
 * operation order, bit-reproducible) or VRN_PREC_FP64_FAST (FMA contraction,
 * reciprocal-based division, squared-norm Newton stop).  No counterpart in
 * the reference, whose 1D solver runs one FP64 form; VRN_PREC_FP32 is
 * refused with VRN_E_ARG. */
int vrn_tubes_set_precision(vrn_tubes *tb, int precision);
/* Advance all tubes by nsteps midpoint-RK2 steps with `grid` persistent
 * workgroups of `block` threads (multiple of 64); use_graph replays a
 * captured hipGraph.  *ms receives the device time.  Blocks. */
int vrn_tubes_run(vrn_tubes *tb, int nsteps, int block, int grid, int use_graph, double *ms);
/* restore = 0 saves (P, U), restore = 1 puts them back (used by the autotuner). */
int vrn_tubes_checkpoint(vrn_tubes *tb, int restore);
int vrn_tubes_info(vrn_tubes *tb, int64_t *ncell, int *cu_count);

#ifdef __cplusplus
}
#endif
#endif /* VERONA_HIP_H */
