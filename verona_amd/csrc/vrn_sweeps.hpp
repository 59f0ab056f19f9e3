// vrn_sweeps.hpp -- the fused directional sweep kernels of vrn_step (see
// vrn_grid.hip for the layout and the design).  Included once per arithmetic
// form: VRN_NS names the namespace of the instantiation and VRN_FAST selects
// the contracted / reciprocal-division FP64 form.
#pragma once
#include "vrn_types.hpp"
#include "vrn_physics.hpp"

#include <algorithm>
#include <type_traits>

namespace VRN_NS {
using vrn::Geo;
using vrn::DevState;
using vrn::SweepArgs;
using vrn::gidx;
using vrn::map_coord;
using vrn::XOFF;

// ------------------------------------------------------------ wave helpers
// Value of lane l-1, and `fill` in lane 0: DPP wave_shr:1 (GFX9 wave-wide
// shift, a VALU operand modifier -- no LDS round trip) with bound_ctrl off,
// so lane 0, whose source lane does not exist, keeps the old value `fill`.
// Streams that a launch reads or writes exactly once (the marching sweeps'
// look-ahead P loads, Uacc and U0 loads, and every Uacc / U / P store) use
// non-temporal accesses: same-box A/B at 384^3, z sweep 2.25 -> 2.18 ms, y
// sweep 1.54 -> 1.52, ghost pass 0.131 -> 0.108 (the stores no longer leave
// the state slots resident in L2).  The x sweep's base-state loads stay
// ordinary: as non-temporal loads they cost 1.71 -> 1.89 ms.
__device__ __forceinline__ double ld_once(const double *p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ void st_once(double *p, double v) { __builtin_nontemporal_store(v, p); }

__device__ __forceinline__ int dpp_shr1(int x, int fill)
{
    return __builtin_amdgcn_update_dpp(fill, x, 0x138, 0xf, 0xf, false);
}
__device__ __forceinline__ float shift_up1(float v, float fill)
{
    return __int_as_float(dpp_shr1(__float_as_int(v), __float_as_int(fill)));
}
__device__ __forceinline__ double shift_up1(double v, double fill)
{
    const long long b = __double_as_longlong(v), f = __double_as_longlong(fill);
    const int lo = dpp_shr1((int)(b & 0xffffffffll), (int)(f & 0xffffffffll));
    const int hi = dpp_shr1((int)(b >> 32), (int)(f >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
// Value of lane l+1, and `fill` in lane 63: DPP wave_shl:1, bound_ctrl off.
__device__ __forceinline__ int dpp_shl1(int x, int fill)
{
    return __builtin_amdgcn_update_dpp(fill, x, 0x130, 0xf, 0xf, false);
}
__device__ __forceinline__ double shift_down1(double v, double fill)
{
    const long long b = __double_as_longlong(v), f = __double_as_longlong(fill);
    const int lo = dpp_shl1((int)(b & 0xffffffffll), (int)(f & 0xffffffffll));
    const int hi = dpp_shl1((int)(b >> 32), (int)(f >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
// Value of lane 63 in every lane (v_readlane into a scalar register).
__device__ __forceinline__ float lane63(float v)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}
__device__ __forceinline__ double lane63(double v)
{
    long long b = __double_as_longlong(v);
    int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), 63);
    int hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ __forceinline__ double wave_max(double r)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) r = fmax(r, __shfl_xor(r, off));
    return r;
}

__device__ __forceinline__ void rate_reduce(DevState *st, double r)
{
    if (!(r >= 0.0)) r = 0.0;  // NaN rates are ignored, as in a `r > max` scan
    r = wave_max(r);
    if (__lane_id() == 0) atomicMax(&st->rate_bits, (unsigned long long)__double_as_longlong(r));
}

__device__ __forceinline__ void rescue_push(int *list, DevState *st, bool need, int idx)
{
    unsigned long long m = __ballot(need);
    if (m == 0) return;
    int leader = __ffsll((long long)m) - 1;
    unsigned int base = 0;
    if (__lane_id() == leader) base = atomicAdd(&st->rescue_count, (unsigned int)__popcll(m));
    base = __shfl(base, leader);
    if (need) {
        unsigned long long below = m & ((1ull << __lane_id()) - 1ull);
        list[base + __popcll(below)] = idx;
    }
}

template <int NV> __device__ __forceinline__ bool all_finite(const double *p)
{
    bool ok = true;
#pragma unroll
    for (int v = 0; v < NV; v++) ok = ok && __builtin_isfinite(p[v]);
    return ok;
}

// Combine, con2prim, Limit, rate for one updated cell of the last sweep.
// A cell counts as unrecovered when its con2prim result is not finite (the
// reference's NaN marker, or a NaN from a branch that reported convergence).
template <int NV, bool COMBINED = false>
__device__ __forceinline__ void finish_cell(const SweepArgs &a, long cu, const double *acc,
                                            const double *u0, const double *guess, bool valid,
                                            double &rate)
{
    const long ps = a.geo.ps;
    double out[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) out[v] = (COMBINED || a.a == 0.0) ? acc[v] : a.a * u0[v] + a.b * acc[v];
    bool need = false;
    if (valid) {
#pragma unroll
        for (int v = 0; v < NV; v++) st_once(&a.Uout[v * ps + cu], out[v]);
        double Pn[NV];
        int iters = 0;
        bool ok;
        if constexpr (NV == 5) {
            ok = c2p_1dw(out, guess, a.g, a.n_iter, a.tol, Pn, iters);
        } else {
#pragma unroll
            for (int v = 0; v < NV; v++) Pn[v] = guess[v];
            ok = c2p_2d(out, Pn, a.g, a.n_iter, a.tol, iters);
            if (!ok) atomicAdd(&a.st->n_singular, 1ull);
            ok = true;
        }
        if (ok) {
            if (!all_finite<NV>(Pn)) atomicAdd(&a.st->n_unrecovered, 1ull);
            Pn[0] = floor_max(a.floor_, Pn[0]);
            Pn[1] = floor_max(a.floor_, Pn[1]);
#pragma unroll
            for (int v = 0; v < NV; v++) st_once(&a.Pout[v * ps + cu], Pn[v]);
            if (a.do_rate) rate = fmax(rate, cell_rate<NV>(Pn, a.g, a.inv_dx));
        } else {
            need = true;
        }
    }
    if constexpr (NV == 5) rescue_push(a.rescue, a.st, need, (int)cu);
}

// ------------------------------------------------------------ x sweep
// Lanes walk a stream of (row, cell) pairs.  A wave owns either
// rows_per_wave whole interior rows (j, k) or, when xseg > 1, one of xseg
// segments [x0, x1) of one row.  Per row the stream holds cells
// c = x0-1 .. x1 (their reconstructions); the lane with cell c also solves
// face c-1/2 (c >= x0) and updates cell c-1 (c >= x0+1).  Whole rows are
// x0 = 3, x1 = NX-3.
// Waves per SIMD the x sweep is compiled for: the fast FP64 WENO-Z
// instantiations take 131 VGPRs at 3 (384^3 jet: 1.68 ms per launch; bounded
// to 4 they spill 12 bytes and take 1.75 ms); the others stay unbounded.
#if defined(VRN_FAST)
template <class R, int REC> constexpr int xsweep_waves() { return REC == REC_WENOZ ? 3 : 1; }
#else
template <class R, int REC> constexpr int xsweep_waves() { return 1; }
#endif

// Where the RK combination U' = a U0 + b (U_b + dt L) of a stage is formed.
// Exact FP64 and FP32 forms and 2D: by the last sweep, from the accumulated
// update (the reference's order).  3D fast FP64 form: by the first sweep of
// the stage, as a U0 + b U_b, with every sweep's update scaled by b, so the
// last sweep (the HBM-bound one) reads no U0.  Same value, different
// rounding; the x sweep + y march and the fused x-y march do it alike.
#if defined(VRN_FAST)
template <int NV> constexpr bool rk_first() { return NV == 5; }
#else
template <int NV> constexpr bool rk_first() { return false; }
#endif
// a U0 + b U_b with one explicit rounding order (left to FMA contraction,
// a x + b y may become either fma(a, x, b y) or fma(b, y, a x), and the
// x sweep and the x-y march must round alike)
__device__ __forceinline__ double rk_base(double a, double u0, double b, double ub)
{
    return __builtin_fma(a, u0, b * ub);
}

template <int NV, class R, int REC, int RIEM, bool LAST>
__global__ __launch_bounds__(256, (xsweep_waves<R, REC>())) void k_xsweep(SweepArgs a)
{
    const Geo G = a.geo;
    const int lane = threadIdx.x & 63;
    const int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const int Nyi = G.NY - 6;
    const int Nzi = NV == 5 ? G.NZ - 6 : 1;
    const int nrows = Nyi * Nzi;
    int r0, r1, x0, x1;
    if (a.xseg > 1) {
        const int Nxi = G.NX - 6, L = (Nxi + a.xseg - 1) / a.xseg;
        r0 = wave / a.xseg;
        r1 = r0 + 1;
        x0 = 3 + (wave % a.xseg) * L;
        x1 = min(x0 + L, G.NX - 3);
        if (r0 >= nrows || x0 >= x1) return;
    } else {
        r0 = wave * a.rows_per_wave;
        if (r0 >= nrows) return;
        r1 = min(r0 + a.rows_per_wave, nrows);
        x0 = 3;
        x1 = G.NX - 3;
    }
    const int E = x1 - x0 + 2;
    const int total = (r1 - r0) * E;
    const long ps = G.ps;
    const R g = (R)a.g, fl = (R)a.floor_;
    constexpr bool PRE = rk_first<NV>() && !LAST;
    const double cdt = (PRE ? a.b : 1.0) * (((a.dt_dev ? a.st->dt : a.dt_fixed) * a.dtfac) / a.dx_dim);
    const double *__restrict__ P = a.P;

    R cup[NV];
    double cF[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) {
        cup[v] = R(0);
        cF[v] = 0.0;
    }
    double rate = 0.0;
    // this lane's stream position (row offset rr, cell offset cc) and row base,
    // advanced by 64 cells per chunk without integer division
    const int nrw = r1 - r0;
    int rr = lane / E, cc = lane % E;
    int jj = (r0 + rr) % Nyi;
    long rb = gidx(G, 0, 3 + jj, NV == 5 ? 3 + (r0 + rr) / Nyi : 0);
    const long rb_last = gidx(G, 0, 3 + (r1 - 1) % Nyi, NV == 5 ? 3 + (r1 - 1) / Nyi : 0);
    for (int s0 = 0; s0 < total; s0 += 64) {
        const bool act = rr < nrw;
        const int c = act ? x0 - 1 + cc : x1;
        const long rbu = act ? rb : rb_last;
        const bool upd = act && c >= x0 + 1;
        const long cu = rbu + (c - 1);
        // base state of the cell this lane updates, issued before the flux work
        double ub[NV], u0[NV];
#pragma unroll
        for (int v = 0; v < NV; v++) {
            ub[v] = upd ? a.Ub[v * ps + cu] : 0.0;
            u0[v] = (PRE && upd && a.a != 0.0) ? a.U0[v * ps + cu] : 0.0;
        }
        R lo[NV], up[NV];
#pragma unroll
        for (int v = 0; v < NV; v++) {
            const double *q = P + v * ps + rbu + c;
            recon<REC>((R)q[-2], (R)q[-1], (R)q[0], (R)q[1], (R)q[2], lo[v], up[v]);
        }
        lo[0] = floor_max(fl, lo[0]);
        lo[1] = floor_max(fl, lo[1]);
        up[0] = floor_max(fl, up[0]);
        up[1] = floor_max(fl, up[1]);
        R upl[NV];
#pragma unroll
        for (int v = 0; v < NV; v++) upl[v] = shift_up1(up[v], cup[v]);
        R Ff[NV];
        face_flux<NV, 0, RIEM>(upl, lo, g, Ff);
        double F[NV], Fl[NV];
#pragma unroll
        for (int v = 0; v < NV; v++) {
            F[v] = (double)Ff[v];
            Fl[v] = shift_up1(F[v], cF[v]);
        }
        if constexpr (!LAST) {
            if (upd) {
                if (PRE && a.a != 0.0) {
#pragma unroll
                    for (int v = 0; v < NV; v++) ub[v] = rk_base(a.a, u0[v], a.b, ub[v]);
                }
#pragma unroll
                for (int v = 0; v < NV; v++) st_once(&a.Uacc[v * ps + cu], ub[v] - cdt * (F[v] - Fl[v]));
            }
        } else {
            // 2D: the x sweep is never last; kept for completeness of the template.
            double acc[NV], guess[NV];
#pragma unroll
            for (int v = 0; v < NV; v++) {
                acc[v] = upd ? ub[v] - cdt * (F[v] - Fl[v]) : 0.0;
                u0[v] = (upd && a.a != 0.0) ? a.U0[v * ps + cu] : 0.0;
                guess[v] = upd ? P[v * ps + cu] : 1.0;
            }
            finish_cell<NV>(a, cu, acc, u0, guess, upd, rate);
        }
#pragma unroll
        for (int v = 0; v < NV; v++) {
            cup[v] = lane63(up[v]);
            cF[v] = lane63(F[v]);
        }
        cc += 64;
        while (cc >= E) {
            cc -= E;
            rr++;
            if (++jj == Nyi) {
                jj = 0;
                rb += 7L * G.px;  // past the 6 ghost rows into the next (j, k) plane row
            } else {
                rb += G.px;
            }
        }
    }
    if constexpr (LAST)
        if (a.do_rate) rate_reduce(a.st, rate);
}

// ------------------------------------------------------------ x-y march: wave-edge faces
// The six x faces of rows r0 .. r1-1 at the ends of one wave's 64 columns
// [x0, x0 + 64) whose stencils leave those columns: faces between cells
// m - 1 and m for m = x0, x0 + 1, x0 + 2 (edge faces 0..2) and m = xr - 2,
// xr - 1, xr with xr = min(x0 + 64, NX - 3) (edge faces 3..5).  Lane L
// reconstructs cell x0 - 1 + i (i = L % 8 < 4) or xr - 3 + (i - 4) of row
// r0 + L / 8, as the x sweep does (both faces, floors), and lanes with i != 0, 4
// solve the face with their left neighbour's upper state (DPP shift):
// exactly the x sweep's faces.  Faces of rows outside [r0, r1) or whose
// stencil leaves the grid are not needed and not written.  Stencil cells in
// ghost columns are read through the x boundary rules (xmap_lo / xmap_hi)
// when the last ghost pass left the x faces unfilled, which gives the values
// that pass would have written.  Out: ebw[(row - rb) & 7 row stride
// ebs][face][var].
template <int NV, class R, int REC, int RIEM>
__device__ __forceinline__ void edge_faces(const SweepArgs &a, const Geo &G, double *ebw, int ebs, int lane, int x0,
                                           int xr, int o, int r0, int r1, int rb)
{
    const int i = lane & 7, r = r0 + (lane >> 3);
    const int cell = i < 4 ? x0 - 1 + i : xr - 7 + i;
    const bool ok = r < r1 && cell >= 2 && cell <= G.NX - 3;
    const int cc = ok ? cell : 3;
    const int rr = ok ? r : 3;
    const long rb0 = gidx(G, cc, rr, o);
    const long ps = G.ps;
    const R g = (R)a.g, fl = (R)a.floor_;
    // wave-uniform: only the waves at the x ends of the grid read ghost columns
    const bool mapx = (a.xmap_lo | a.xmap_hi) != 0 && (x0 == 3 || xr == G.NX - 3);
    R q5[5][NV];
    if (mapx) {
        const long row = rb0 - cc;
#pragma unroll
        for (int k = 0; k < 5; k++) {
            const long xs = row + map_coord(cc - 2 + k, G.NX, a.xmap_lo, a.xmap_hi);
#pragma unroll
            for (int v = 0; v < NV; v++) q5[k][v] = (R)a.P[v * ps + xs];
        }
    } else {
#pragma unroll
        for (int v = 0; v < NV; v++) {
            const double *q = a.P + v * ps + rb0;
#pragma unroll
            for (int k = 0; k < 5; k++) q5[k][v] = (R)q[k - 2];
        }
    }
    R lo[NV], up[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) recon<REC>(q5[0][v], q5[1][v], q5[2][v], q5[3][v], q5[4][v], lo[v], up[v]);
    lo[0] = floor_max(fl, lo[0]);
    lo[1] = floor_max(fl, lo[1]);
    up[0] = floor_max(fl, up[0]);
    up[1] = floor_max(fl, up[1]);
    R upl[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) upl[v] = shift_up1(up[v], up[v]);
    R Ff[NV];
    face_flux<NV, 0, RIEM>(upl, lo, g, Ff);
    // the face needs the left cell's stencil in the grid too
    const bool okl = ok && i != 0 && i != 4 && cell - 3 >= 0;
    if (okl) {
        const int e = i < 4 ? i - 1 : i - 2;
        double *w = ebw + ((r - rb) & 7) * ebs + e * NV;
#pragma unroll
        for (int v = 0; v < NV; v++) w[v] = (double)Ff[v];
    }
}

// ------------------------------------------------------------ y / z sweeps
// One lane per interior (x, o) column, marching along DIM over the segment
// [m0, m1) of interior cells.  The 5-cell reconstruction window of each lane
// lives in LDS as a ring of 5 slots laid out [slot][var][lane] (each lane
// touches only its own words: conflict-free and barrier-free), which keeps
// it out of the VGPR budget of the Riemann solve and con2prim.
template <int NV, int DIM, class R, int REC, int RIEM, bool LAST, bool XY = false>
__global__ __launch_bounds__(256, VRN_MARCH_MIN_WAVES) void k_march(SweepArgs a)
{
    // lanes per block: 256 in 3D (four o-rows of 64 columns), 64 in 2D
    constexpr int WS = NV == 5 ? 256 : 64;
    // XY (3D y march only): the march also does the x direction of each row
    // it updates, with the ring row of its 64 columns as the x stencils.  The
    // six faces at a wave's ends whose stencils leave those columns (three
    // each side) are solved every 8 rows for the next 8 by the whole wave
    // (edge_faces) into eb[row & 7][wave][face][var].
    static_assert(!XY || (DIM == 1 && NV == 5 && !LAST), "x-y fusion is the 3D y march");
    constexpr int WSR = WS;  // ring stride per (slot, variable)
    __shared__ R win[5 * NV * WSR];
    constexpr int EBR = (WS / 64) * 6 * NV;  // edge-face buffer stride per row
    __shared__ double eb[XY ? 8 * EBR : 1];
    // WENO-Z: the curvature term of each cell (wz_curv) is computed once and
    // carried: cell c needs the terms of cells c-1, c and c+1, and the term of
    // cell x lives in slot x & 1 of a two-slot ring [slot][var][lane].
    constexpr bool CARRY = REC == REC_WENOZ;
    __shared__ R curv[CARRY ? 2 * NV * WS : 1];
    const Geo G = a.geo;
    const int tid = threadIdx.x;
    const int x = 3 + blockIdx.x * 64 + (threadIdx.x & 63);
    const int o = 3 + blockIdx.y * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int Nd = DIM == 1 ? G.NY : G.NZ;
    const int No = NV == 5 ? (DIM == 1 ? G.NZ : G.NY) : 7;
    const int m0 = 3 + blockIdx.z * a.seg_len;
    // whole-wave exits only: a partially active wave keeps all lanes for the
    // cross-lane reductions and masks the inactive ones.
    if (o >= No - 3 || m0 >= Nd - 3) return;
    const bool act = x < G.NX - 3;
    // XY: every lane holds its own column (inactive lanes a ghost or the
    // last one), since the x stencils of the active lanes reach into them;
    // ghost columns through the x boundary rules when the last ghost pass
    // left the x faces unfilled
    const int xx = XY ? map_coord(min(x, G.NX - 1), G.NX, a.xmap_lo, a.xmap_hi) : (act ? x : 3);
    const int m1 = min(m0 + a.seg_len, Nd - 3);
    const long sd = DIM == 1 ? (long)G.px : (long)G.px * G.NY;
    const long cb = DIM == 1 ? gidx(G, xx, 0, NV == 5 ? o : 0) : gidx(G, xx, o, 0);
    const long ps = G.ps;
    const R g = (R)a.g, fl = (R)a.floor_;
    // 3D fast FP64 form: updates scaled by the RK weight b; the x-y march
    // forms a U0 + b U_b itself, the z sweep reads no U0
    constexpr bool PRE = rk_first<NV>();
    const double cdt = (PRE ? a.b : 1.0) * (((a.dt_dev ? a.st->dt : a.dt_fixed) * a.dtfac) / a.dx_dim);
    const double *__restrict__ P = a.P;
    // the x-y march starts from the stage's base state (there is no x sweep)
    const double *Ui = XY ? a.Ub : a.Uacc;
    const int lane = tid & 63;
    R *wl = win + tid;  // slot s, variable v at wl[(s * NV + v) * WSR]
    R *kl = curv + tid;  // curvature term of cell x, variable v at kl[((x & 1) * NV + v) * WS]

    // ring: cell m lives in slot (m - (m0 - 3)) mod 5
    double nx[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) {
        const double *q = P + v * ps + cb;
#pragma unroll
        for (int k = 0; k < 4; k++) wl[(k * NV + v) * WSR] = (R)q[(m0 - 3 + k) * sd];
        nx[v] = q[(m0 + 1) * sd];
        if constexpr (CARRY) {
            // cells m0-2 and m0-1, the neighbours of the first window's centre
            kl[(((m0 - 2) & 1) * NV + v) * WS] =
                wz_curv(wl[(0 * NV + v) * WSR], wl[(1 * NV + v) * WSR], wl[(2 * NV + v) * WSR]);
            kl[(((m0 - 1) & 1) * NV + v) * WS] =
                wz_curv(wl[(1 * NV + v) * WSR], wl[(2 * NV + v) * WSR], wl[(3 * NV + v) * WSR]);
        }
    }
    const int wv = tid >> 6;
    // XY: the edge face that is this lane's left face (-1: solved in the
    // wave): lanes 0..2 -> edge faces 0..2; lanes 62, 63 -> the right edge's
    // face at their column (x0 + lane relative to the edge xr = min(x0 + 64,
    // NX - 3)); lane 63 also takes its right face, edge face 5
    const int x0 = 3 + blockIdx.x * 64;
    const int xr = min(x0 + 64, G.NX - 3);
    const int lfe = lane < 3 ? lane : (lane >= 62 ? 3 + min(max(x0 + lane - xr + 2, 0), 2) : -1);
    const double cdtx = XY ? (PRE ? a.b : 1.0) * (((a.dt_dev ? a.st->dt : a.dt_fixed) * a.dtfac) / a.dx_x) : 0.0;
    R upp[NV];
    double Fp[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) {
        upp[v] = R(0);
        Fp[v] = 0.0;
    }
    double rate = 0.0;
    int s0 = 0;  // slot of cell c - 2
    for (int c = m0 - 1; c <= m1; c++) {
        const int s1 = s0 == 4 ? 0 : s0 + 1, s2 = s1 == 4 ? 0 : s1 + 1;
        const int s3 = s2 == 4 ? 0 : s2 + 1, s4 = s3 == 4 ? 0 : s3 + 1;
#pragma unroll
        for (int v = 0; v < NV; v++) wl[(s4 * NV + v) * WSR] = (R)nx[v];
        if constexpr (XY) {
            // the x stencils read other lanes' ring words (written by this
            // wave); keep the compiler from moving the reads above the writes
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        if (c + 3 < Nd) {
#pragma unroll
            for (int v = 0; v < NV; v++) nx[v] = ld_once(&P[v * ps + cb + (c + 3) * sd]);
        }
        // conserved state of cell c-1 (updated at the end of this iteration):
        // issued now, consumed after the reconstruction and the Riemann solve
        const bool upd = c - 1 >= m0;
        const long cu = cb + (long)(c - 1) * sd;
        double ub[NV], u0b[NV];
#pragma unroll
        for (int v = 0; v < NV; v++) {
            ub[v] = upd ? ld_once(&Ui[v * ps + cu]) : 0.0;
            u0b[v] = (((LAST && !PRE) || (XY && PRE)) && upd && a.a != 0.0) ? ld_once(&a.U0[v * ps + cu]) : 0.0;
        }
        R lo[NV], up[NV];
#pragma unroll
        for (int v = 0; v < NV; v++) {
            if constexpr (CARRY) {
                const R q2 = wl[(s2 * NV + v) * WSR], q3 = wl[(s3 * NV + v) * WSR], q4 = wl[(s4 * NV + v) * WSR];
                R *kp = kl + (((c + 1) & 1) * NV + v) * WS;  // cell c-1's term, then cell c+1's
                const R km = *kp, k0 = kl[((c & 1) * NV + v) * WS], kn = wz_curv(q2, q3, q4);
                *kp = kn;
                wenoz_k(wl[(s0 * NV + v) * WSR], wl[(s1 * NV + v) * WSR], q2, q3, q4, km, k0, kn, lo[v], up[v]);
            } else {
                recon<REC>(wl[(s0 * NV + v) * WSR], wl[(s1 * NV + v) * WSR], wl[(s2 * NV + v) * WSR],
                           wl[(s3 * NV + v) * WSR], wl[(s4 * NV + v) * WSR], lo[v], up[v]);
            }
        }
        lo[0] = floor_max(fl, lo[0]);
        lo[1] = floor_max(fl, lo[1]);
        up[0] = floor_max(fl, up[0]);
        up[1] = floor_max(fl, up[1]);
        if (c >= m0) {
            R Ff[NV];
            face_flux<NV, DIM, RIEM>(upp, lo, g, Ff);
            double F[NV];
#pragma unroll
            for (int v = 0; v < NV; v++) F[v] = (double)Ff[v];
            if (upd) {
                if constexpr (XY) {
                    if (PRE && a.a != 0.0) {
#pragma unroll
                        for (int v = 0; v < NV; v++) ub[v] = rk_base(a.a, u0b[v], a.b, ub[v]);
                    }
                    // every 8 rows: the edge faces of rows c-1 .. c+6
                    if (((c - 1 - m0) & 7) == 0) {
                        edge_faces<NV, R, REC, RIEM>(a, G, eb + wv * 6 * NV, EBR, lane, x0, xr, o, c - 1,
                                                     min(c + 7, m1), m0);
                        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                        __builtin_amdgcn_wave_barrier();
                        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                    }
                    // x direction of row c-1, as the x sweep does it: each
                    // lane reconstructs its cell (stencil from the ring row),
                    // solves its left face with its left neighbour's upper
                    // state (DPP shift) and takes its right face from its
                    // right neighbour; the faces whose stencils leave the
                    // wave's 64 columns come from edge_faces.  Uniform
                    // control flow: the shifts read every lane of the wave.
                    R lox[NV], upx[NV];
#pragma unroll
                    for (int v = 0; v < NV; v++) {
                        const R *rw = wl + (s1 * NV + v) * WSR;
                        recon<REC>(rw[-2], rw[-1], rw[0], rw[1], rw[2], lox[v], upx[v]);
                    }
                    lox[0] = floor_max(fl, lox[0]);
                    lox[1] = floor_max(fl, lox[1]);
                    upx[0] = floor_max(fl, upx[0]);
                    upx[1] = floor_max(fl, upx[1]);
                    R uplx[NV];
#pragma unroll
                    for (int v = 0; v < NV; v++) uplx[v] = shift_up1(upx[v], upx[v]);
                    R Fxf[NV];
                    face_flux<NV, 0, RIEM>(uplx, lox, g, Fxf);
                    // edge faces of row c-1: each lane reads its own (lfe)
                    // and edge face 5 (lane 63's right face)
                    const double *be = eb + ((c - 1 - m0) & 7) * EBR + wv * 6 * NV;
                    const int lf = lfe < 0 ? 0 : lfe;
#pragma unroll
                    for (int v = 0; v < NV; v++) {
                        const double fxl = lfe >= 0 ? be[lf * NV + v] : (double)Fxf[v];
                        const double fxs = shift_down1(fxl, fxl);
                        const double fxr = lane == 63 ? be[5 * NV + v] : fxs;
                        ub[v] = ub[v] - cdtx * (fxr - fxl);
                    }
                }
                double acc[NV];
#pragma unroll
                for (int v = 0; v < NV; v++) acc[v] = ub[v] - cdt * (F[v] - Fp[v]);
                if constexpr (!LAST) {
                    if (act) {
#pragma unroll
                        for (int v = 0; v < NV; v++) st_once(&a.Uacc[v * ps + cu], acc[v]);
                    }
                } else {
                    // warm-start guess: the stage's primitives of cell c-1
                    double guess[NV];
#pragma unroll
                    for (int v = 0; v < NV; v++) {
                        if constexpr (std::is_same<R, double>::value)
                            guess[v] = wl[(s1 * NV + v) * WSR];
                        else
                            guess[v] = P[v * ps + cu];
                    }
                    finish_cell<NV, PRE>(a, cu, acc, u0b, guess, act, rate);
                }
            }
#pragma unroll
            for (int v = 0; v < NV; v++) Fp[v] = F[v];
        }
#pragma unroll
        for (int v = 0; v < NV; v++) upp[v] = up[v];
        s0 = s1;
    }
    if constexpr (LAST)
        if (a.do_rate) rate_reduce(a.st, rate);
}

// ------------------------------------------------------------ 2D stage end
// Exact FP64 and FP32 forms: the y sweep leaves the stage's accumulated
// update in Uacc, like the middle sweep in 3D, and this pointwise pass forms
// the RK combination and runs the 2D Newton con2prim (2dim/structs.jl:249-330)
// from the stage's primitives, Limit and the CFL rate -- the same arithmetic
// as doing it inside the march.  With the reference's pivoted LU the Newton
// step is heavy enough that spreading the cells over every SIMD pays.  The
// fast FP64 form solves the Newton system in closed form and runs the stage
// end inside the y march instead (launch_stage), which saves the Uacc round
// trip and a read of P.
template <int NV>
__global__ __launch_bounds__(256) void k_finish2d(SweepArgs a)
{
    const Geo G = a.geo;
    const int nx = G.NX - 6, ny = G.NY - 6;
    const long ps = G.ps;
    double rate = 0.0;
    // x from the block column and lane, rows strided over blockIdx.y: no
    // integer division per cell.  The row loop bound is block-uniform, so
    // every lane of a wave takes the same trips and the wave reduction of
    // the rate below sees all 64 lanes.
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const bool xin = x < nx;
    for (int y = blockIdx.y; y < ny; y += gridDim.y) {
        const bool valid = xin;
        const long cu = gidx(G, 3 + (xin ? x : 0), 3 + y, 0);
        double acc[NV], u0[NV], guess[NV];
#pragma unroll
        for (int v = 0; v < NV; v++) {
            acc[v] = valid ? ld_once(&a.Uacc[v * ps + cu]) : 0.0;
            u0[v] = (valid && a.a != 0.0) ? ld_once(&a.U0[v * ps + cu]) : 0.0;
            guess[v] = valid ? ld_once(&a.P[v * ps + cu]) : 1.0;
        }
        finish_cell<NV>(a, cu, acc, u0, guess, valid, rate);
    }
    if (a.do_rate) rate_reduce(a.st, rate);
}

// ------------------------------------------------------------ rescue (3D)
// Cells whose 1DW Newton did not converge: rest of the cascade from the
// stage's guess, then Limit.  The list holds linear device indices.
__global__ __launch_bounds__(64) void k_rescue(SweepArgs a)
{
    const unsigned int n = a.st->rescue_count;
    const long ps = a.geo.ps;
    double rate = 0.0;
    for (unsigned int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        const long cu = a.rescue[t];
        double U[5], Pn[5];
#pragma unroll
        for (int v = 0; v < 5; v++) {
            U[v] = a.Uout[v * ps + cu];
            Pn[v] = a.P[v * ps + cu];
        }
        int iters = 0;
        c2p_cascade_tail(U, Pn, a.g, a.n_iter, a.tol, iters);
        if (!all_finite<5>(Pn)) atomicAdd(&a.st->n_unrecovered, 1ull);
        Pn[0] = floor_max(a.floor_, Pn[0]);
        Pn[1] = floor_max(a.floor_, Pn[1]);
#pragma unroll
        for (int v = 0; v < 5; v++) a.Pout[v * ps + cu] = Pn[v];
        if (a.do_rate) rate = fmax(rate, cell_rate<5>(Pn, a.g, a.inv_dx));
    }
    if (a.do_rate) rate_reduce(a.st, rate);
}


// Launch the three sweeps + rescue of one RK stage.
template <int NV, class R, int REC, int RIEM>
void launch_stage(hipStream_t s, const SweepArgs &a, const double dxs[3], void (*timer)(void *, int, int), void *tctx)
{
    const Geo &G = a.geo;
    const int nxb = (G.NX - 6 + 63) / 64;
    // 3D FP64 forms: the x direction runs inside the y march, which saves
    // the x sweep's write and the y sweep's re-read of the accumulated U and
    // one read of P per stage; the same fluxes and updates in the same
    // order as the x and y sweeps.
    const bool fxy = NV == 5 && std::is_same<R, double>::value && a.fuse_xy;
    if (fxy) {
        if constexpr (NV == 5 && std::is_same<R, double>::value) {
            SweepArgs ay = a;
            ay.dx_dim = dxs[1];
            ay.dx_x = dxs[0];
            int nseg = (G.NY - 6 + a.seg_len - 1) / a.seg_len;
            timer(tctx, 10, 0);
            hipLaunchKernelGGL((k_march<5, 1, R, REC, RIEM, false, true>), dim3(nxb, (G.NZ - 6 + 3) / 4, nseg),
                               dim3(256), 0, s, ay);
            timer(tctx, 10, 1);
        }
    } else {
        SweepArgs ax = a;
        ax.dx_dim = dxs[0];
        int nrows = (G.NY - 6) * (NV == 5 ? G.NZ - 6 : 1);
        int waves = ax.xseg > 1 ? nrows * ax.xseg : (nrows + ax.rows_per_wave - 1) / ax.rows_per_wave;
        timer(tctx, 0, 0);
        hipLaunchKernelGGL((k_xsweep<NV, R, REC, RIEM, false>), dim3((waves + 3) / 4), dim3(256), 0, s, ax);
        timer(tctx, 0, 1);
    }
    if constexpr (NV == 5) {
        if (!fxy) {
            SweepArgs ay = a;
            ay.dx_dim = dxs[1];
            int nseg = (G.NY - 6 + a.seg_len - 1) / a.seg_len;
            timer(tctx, 1, 0);
            hipLaunchKernelGGL((k_march<5, 1, R, REC, RIEM, false>), dim3(nxb, (G.NZ - 6 + 3) / 4, nseg), dim3(256), 0, s, ay);
            timer(tctx, 1, 1);
        }
        {
            SweepArgs az = a;
            az.dx_dim = dxs[2];
            int nseg = (G.NZ - 6 + a.seg_len - 1) / a.seg_len;
            timer(tctx, 2, 0);
            hipLaunchKernelGGL((k_march<5, 2, R, REC, RIEM, true>), dim3(nxb, (G.NY - 6 + 3) / 4, nseg), dim3(256), 0, s, az);
            timer(tctx, 2, 1);
        }
        {
            timer(tctx, 4, 0);
            hipLaunchKernelGGL(k_rescue, dim3(256), dim3(64), 0, s, a);
            timer(tctx, 4, 1);
        }
    } else {
        SweepArgs ay = a;
        ay.dx_dim = dxs[1];
        int nseg = (G.NY - 6 + a.seg_len - 1) / a.seg_len;
#if defined(VRN_FAST)
        // fast FP64 form: the y sweep is the last sweep (RK combination,
        // closed-form 4x4 Newton, Limit and rate inside the march); same
        // results as the separate pass, bit for bit
        timer(tctx, 1, 0);
        hipLaunchKernelGGL((k_march<4, 1, R, REC, RIEM, true>), dim3(nxb, 1, nseg), dim3(64), 0, s, ay);
        timer(tctx, 1, 1);
        return;
#endif
        timer(tctx, 1, 0);
        hipLaunchKernelGGL((k_march<4, 1, R, REC, RIEM, false>), dim3(nxb, 1, nseg), dim3(64), 0, s, ay);
        timer(tctx, 1, 1);
        // about 2048 blocks of 256 lanes: x columns of 256 cells by row groups
        const int bx = (G.NX - 6 + 255) / 256;
        const int by = std::max(1, std::min(G.NY - 6, 2048 / bx));
        timer(tctx, 3, 0);
        hipLaunchKernelGGL(k_finish2d<4>, dim3((unsigned)bx, (unsigned)by), dim3(256), 0, s, a);
        timer(tctx, 3, 1);
    }
}

template <int NV>
void dispatch_stage(hipStream_t s, const SweepArgs &a, const double dxs[3], int rec, int riem, bool fp32,
                    void (*timer)(void *, int, int), void *tctx)
{
#define VRN_ST(R_, REC_, RI_) launch_stage<NV, R_, REC_, RI_>(s, a, dxs, timer, tctx)
#define VRN_REC(R_, RI_)                                        \
    do {                                                        \
        if (rec == REC_WENOZ) VRN_ST(R_, REC_WENOZ, RI_);       \
        else if (rec == REC_PPM) VRN_ST(R_, REC_PPM, RI_);      \
        else VRN_ST(R_, REC_MINMOD, RI_);                       \
    } while (0)
    if (!fp32) {
        if (riem == RIEM_HLL) VRN_REC(double, RIEM_HLL);
        else VRN_REC(double, RIEM_HLLC);
    } else {
        if (riem == RIEM_HLL) VRN_REC(float, RIEM_HLL);
        else VRN_REC(float, RIEM_HLLC);
    }
#undef VRN_REC
#undef VRN_ST
}

}  // namespace VRN_NS
