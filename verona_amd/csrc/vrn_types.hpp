// vrn_types.hpp -- device layout and argument blocks shared by the grid and tube kernels.
#pragma once
#include <hip/hip_runtime.h>

#ifndef VRN_MARCH_MIN_WAVES
#define VRN_MARCH_MIN_WAVES 2  // measured: 2 waves/SIMD beats the unbounded allocation
#endif

namespace vrn {

constexpr int XOFF = 13;

struct Geo {
    int NX, NY, NZ;  // with ghosts; NZ = 1 in 2D
    int px;          // row pitch in doubles
    long ps;         // plane stride in doubles
};

// Boundary rule of one face applied to one coordinate c of an axis with N
// cells (ghosts included): 0 identity, 1 periodic wrap, 2 zero-gradient copy.
// Ghost coordinates map to the interior cell whose value the ghost holds.
__host__ __device__ __forceinline__ int map_coord(int c, int N, int lo, int hi)
{
    if (c < 3) return lo == 1 ? c + (N - 6) : (lo == 2 ? 3 : c);
    if (c >= N - 3) return hi == 1 ? c - (N - 6) : (hi == 2 ? N - 4 : c);
    return c;
}

__host__ __device__ __forceinline__ long gidx(const Geo &G, int i, int j, int k)
{
    return XOFF + i + (long)G.px * (j + (long)G.NY * k);
}

struct DevState {
    double dt;       // dt of the current step (CFL mode)
    double dt_prev;  // dt of the step that just finished
    unsigned long long rate_bits;   // max signal rate (non-negative double bits)
    unsigned int rescue_count;
    unsigned int pad0;
    unsigned long long n_unrecovered;
    unsigned long long n_rescued;
    unsigned long long n_singular;
};

struct SweepArgs {
    Geo geo;
    const double *P;   // primitives of this stage (read)
    const double *U0;  // U^n (last sweep, a != 0)
    const double *Ub;  // base state (first sweep)
    double *Uacc;      // accumulator (first: write; middle: read/write; last: read)
    double *Uout;      // last sweep: combined conserved state
    double *Pout;      // last sweep: recovered primitives
    DevState *st;
    int *rescue;
    double g, floor_;
    double inv_dx[3];
    double dtfac;     // stage dt = dt * dtfac
    double dx_dim;    // spacing along this sweep's axis
    double dt_fixed;  // dt when dt_dev == 0
    int dt_dev;       // 1: dt comes from DevState (CFL mode)
    double a, b;    // stage output = a U0 + b acc (a == 0: acc)
    double tol;
    int n_iter;
    int do_rate;
    int seg_len;        // y / z sweeps: interior cells per marching segment
    int rows_per_wave;  // x sweep: rows streamed by one wave (xseg == 1)
    int xseg;           // x sweep: segments per row (> 1 only with rows_per_wave == 1)
    // 3D fast FP64 form, x and y directions in one march (0: separate x
    // sweep, the same results), and the x spacing of its flux differences
    int fuse_xy;
    double dx_x;
    // x-y march: boundary rules through which its x reads of ghost columns
    // are mapped to interior cells (0, 0: read the ghost cells; used after
    // a stage end whose ghost pass left the x faces unfilled)
    int xmap_lo, xmap_hi;
};

// Batched 1D tubes (one persistent launch, one tube per workgroup at a time).
struct TubeArgs {
    double *P[3];       // primitives P^n (in/out)
    double *Ph[3];      // stage-1 primitives
    double *U[3];       // conserved U^n (in/out)
    double *Uh[3];      // Uhalf
    const double *P0[3];  // initial primitives (stale face states)
    const long *off;    // cell offset of each tube, [ntubes + 1]
    const int *order;   // tube processing order (longest first)
    const double *dt;   // per-tube dt
    const double *dx;   // per-tube dx
    unsigned int *queue;  // work counter
    int ntubes;
    int nsteps;
    int n_iter;
    double tol;
    double g;
};

}  // namespace vrn
