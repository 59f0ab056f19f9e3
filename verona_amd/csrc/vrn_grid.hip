// vrn_grid.hip -- 2D/3D grid kernels of the SRHD hot path for gfx950, and
// the C ABI of include/verona_hip.h.
//
// Device layout: structure of arrays, one plane per variable.  Cell (i, j, k)
// (0-based, ghosts included) of variable v lives at
//     v * ps + XOFF + i + px * (j + NY * k)
// with a row pitch px that is a multiple of 16 doubles and XOFF = 13, so the
// first interior cell of every row (i = 3) sits on a 128-byte boundary and a
// 64-lane wave reading interior cells of a row touches exactly four lines.
//
// Two kernel sets:
//  * the reference boundary (vrn_fluxes / vrn_update / vrn_cons_to_prim /
//    vrn_limit / vrn_prim_to_cons): one thread per face or cell, same
//    arrays and regions as the Julia KA kernels they replace;
//  * the fused step (vrn_step): per RK stage, one directional sweep kernel
//    per axis.  The x sweep runs 64 lanes along a row as a continuous stream
//    of cells over many rows and passes the upper face state and the face
//    flux to the next lane with a cross-lane shift; the y and z sweeps give
//    each lane one (x, .) column and march along the axis with the 5-cell
//    reconstruction window held in an LDS ring (5 slots x NV variables per
//    lane, so advancing the window moves no data).  Either way each cell's
//    reconstruction and each face's Riemann solve run exactly once, and all
//    global loads and stores are row-contiguous across the wave.  The last
//    sweep of a stage also forms the RK combination, runs 1DW con2prim +
//    Limit, and (CFL mode) reduces the signal speed; cells 1DW leaves
//    unconverged go to a rescue list finished by the full cascade.
#include "../../include/verona_hip.h"
#include "vrn_types.hpp"
#define VRN_NS vrn
#include "vrn_physics.hpp"
#include "vrn_sweeps.hpp"

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cmath>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

namespace vrn {

// ------------------------------------------------------------ ghosts
// One thread per ghost cell.  The value written is the source cell reached
// by mapping each ghost coordinate through its face's rule (periodic wrap or
// zero-gradient copy; identity for faces without a rule), which is what
// applying whole-plane passes x, then y, then z produces.  A source cell
// that is interior along every axis comes from Pi, otherwise from Pg (with
// the floor when do_limit).  Sources are never ghosts of a mapped face, so
// the in-place form (Pd == Pi == Pg) is race free.
struct GhostArgs {
    Geo geo;
    double *Pd;
    const double *Pi;
    const double *Pg;
    int rule_lo[3], rule_hi[3];  // 0 identity, 1 periodic, 2 outflow
    int do_limit;
    double floor_;
    int nozzle;
    double noz_r2, noz_dx, noz_dy, noz_rho, noz_u, noz_ux, noz_uy, noz_uz;
    DevState *st;
    // 3D cells filled: 0 every ghost cell, 1 all but the x faces (cells with
    // an x ghost coordinate and interior y and z), 2 the x faces only
    int part;
};

template <int NV>
__global__ __launch_bounds__(256) void k_ghost(GhostArgs a)
{
    const Geo G = a.geo;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (NV == 5 && a.part == 2) t += 6L * G.NX * G.NY + 6L * G.NX * (G.NZ - 6);
    if (t == 0 && a.st != nullptr) {
        a.st->n_rescued += a.st->rescue_count;
        a.st->rescue_count = 0;
    }
    int i, j, k;
    if constexpr (NV == 5) {
        const long nz6 = 6L * G.NX * G.NY;
        const long ny6 = 6L * G.NX * (G.NZ - 6);
        const long nx6 = 6L * (G.NY - 6) * (G.NZ - 6);
        if (t < nz6) {
            long r = t;
            i = (int)(r % G.NX);
            r /= G.NX;
            j = (int)(r % G.NY);
            int q = (int)(r / G.NY);
            k = q < 3 ? q : G.NZ - 6 + q;
        } else if (t < nz6 + ny6) {
            long r = t - nz6;
            i = (int)(r % G.NX);
            r /= G.NX;
            int q = (int)(r % 6);
            j = q < 3 ? q : G.NY - 6 + q;
            k = 3 + (int)(r / 6);
        } else if (t < nz6 + ny6 + nx6) {
            long r = t - nz6 - ny6;
            int q = (int)(r % 6);
            i = q < 3 ? q : G.NX - 6 + q;
            r /= 6;
            j = 3 + (int)(r % (G.NY - 6));
            k = 3 + (int)(r / (G.NY - 6));
        } else {
            return;
        }
    } else {
        const long ny6 = 6L * G.NX;
        const long nx6 = 6L * (G.NY - 6);
        k = 0;
        if (t < ny6) {
            i = (int)(t % G.NX);
            int q = (int)(t / G.NX);
            j = q < 3 ? q : G.NY - 6 + q;
        } else if (t < ny6 + nx6) {
            long r = t - ny6;
            int q = (int)(r % 6);
            i = q < 3 ? q : G.NX - 6 + q;
            j = 3 + (int)(r / 6);
        } else {
            return;
        }
    }
    const long ps = G.ps;
    const long dst = gidx(G, i, j, k);
    if (NV == 5 && a.nozzle && k < 3 && i >= 3 && i < G.NX - 3 && j >= 3 && j < G.NY - 3) {
        double xc = ((i - 3) + 0.5) * a.noz_dx - 0.5 * (G.NX - 6) * a.noz_dx;
        double yc = ((j - 3) + 0.5) * a.noz_dy - 0.5 * (G.NY - 6) * a.noz_dy;
        if (xc * xc + yc * yc < a.noz_r2) {
            a.Pd[dst] = a.noz_rho;
            a.Pd[ps + dst] = a.noz_u;
            a.Pd[2 * ps + dst] = a.noz_ux;
            a.Pd[3 * ps + dst] = a.noz_uy;
            a.Pd[4 * ps + dst] = a.noz_uz;
            return;
        }
    }
    const int mi = map_coord(i, G.NX, a.rule_lo[0], a.rule_hi[0]);
    const int mj = map_coord(j, G.NY, a.rule_lo[1], a.rule_hi[1]);
    const int mk = NV == 5 ? map_coord(k, G.NZ, a.rule_lo[2], a.rule_hi[2]) : 0;
    const bool interior = mi >= 3 && mi < G.NX - 3 && mj >= 3 && mj < G.NY - 3 &&
                          (NV == 4 || (mk >= 3 && mk < G.NZ - 3));
    const long src = gidx(G, mi, mj, mk);
    if (interior) {
#pragma unroll
        for (int v = 0; v < NV; v++) a.Pd[v * ps + dst] = a.Pi[v * ps + src];
    } else {
        double p[NV];
#pragma unroll
        for (int v = 0; v < NV; v++) p[v] = a.Pg[v * ps + src];
        if (a.do_limit) {
            p[0] = floor_max(a.floor_, p[0]);
            p[1] = floor_max(a.floor_, p[1]);
        }
        if (src != dst || a.do_limit || a.Pd != a.Pg) {
#pragma unroll
            for (int v = 0; v < NV; v++) a.Pd[v * ps + dst] = p[v];
        }
    }
}

// ------------------------------------------------------------ CFL
// Max signal rate over interior cells: grid-stride loop, one atomic per wave.
template <int NV>
__global__ __launch_bounds__(256) void k_rate(Geo G, const double *P, double g, double ix,
                                               double iy, double iz, DevState *st)
{
    const int ni = G.NX - 6, nj = G.NY - 6, nk = NV == 5 ? G.NZ - 6 : 1;
    const long n = (long)ni * nj * nk;
    const double inv[3] = {ix, iy, iz};
    double r = 0.0;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const int i = 3 + (int)(t % ni);
        const long q = t / ni;
        const long c = gidx(G, i, 3 + (int)(q % nj), NV == 5 ? 3 + (int)(q / nj) : 0);
        double p[NV];
#pragma unroll
        for (int v = 0; v < NV; v++) p[v] = P[v * G.ps + c];
        double cr = cell_rate<NV>(p, g, inv);
        if (cr > r) r = cr;
    }
    rate_reduce(st, r);
}

__global__ void k_reset_rate(DevState *st) { st->rate_bits = 0ull; }

// dt = cfl / max(interior rate, boundary inflow rate); resets the accumulator.
__global__ void k_finalize_dt(DevState *st, double cfl, double extra_rate)
{
    double r = __longlong_as_double((long long)st->rate_bits);
    if (extra_rate > r) r = extra_rate;
    st->dt_prev = st->dt;
    st->dt = cfl / r;
    st->rate_bits = 0ull;
}

// ------------------------------------------------------------ reference-boundary kernels
template <int NV>
__global__ __launch_bounds__(256) void k_ref_ptou(Geo G, const double *P, double *U, double g)
{
    const long n = (long)G.NX * G.NY * G.NZ;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const int i = (int)(t % G.NX);
        const long r = t / G.NX;
        const int j = (int)(r % G.NY), k = (int)(r / G.NY);
        const long c = gidx(G, i, j, k);
        double p[NV], u[NV];
#pragma unroll
        for (int v = 0; v < NV; v++) p[v] = P[v * G.ps + c];
        p2u<NV>(p, u, g, lorentz<NV>(p));
#pragma unroll
        for (int v = 0; v < NV; v++) U[v * G.ps + c] = u[v];
    }
}

template <int NV>
__global__ __launch_bounds__(256) void k_ref_limit(Geo G, double *P, double fl)
{
    const long n = (long)G.NX * G.NY * G.NZ;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const int i = (int)(t % G.NX);
        const long r = t / G.NX;
        const long c = gidx(G, i, (int)(r % G.NY), (int)(r / G.NY));
        P[c] = jmax(P[c], fl);
        P[G.ps + c] = jmax(P[G.ps + c], fl);
    }
}

// Fluxes: face i+1/2 along DIM for 0-based (i, j, k) in [2, N-3)^ndim; both
// face states recomputed per face as in the reference (two reconstructions).
template <int NV, int DIM, int REC, int RIEM>
__global__ __launch_bounds__(256) void k_ref_fluxes(Geo G, const double *P, double *F, double g,
                                                     double fl)
{
    const int ni = G.NX - 5, nj = G.NY - 5, nk = NV == 5 ? G.NZ - 5 : 1;
    const long n = (long)ni * nj * nk;
    const long sd = DIM == 0 ? 1 : (DIM == 1 ? (long)G.px : (long)G.px * G.NY);
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const int i = 2 + (int)(t % ni);
        const long r = t / ni;
        const int j = 2 + (int)(r % nj);
        const int k = NV == 5 ? 2 + (int)(r / nj) : 0;
        const long c = gidx(G, i, j, k);
        double PL[NV], PR[NV];
#pragma unroll
        for (int v = 0; v < NV; v++) {
            const double *q = P + v * G.ps + c;
            double lo, hi;
            recon<REC>(q[-2 * sd], q[-sd], q[0], q[sd], q[2 * sd], lo, hi);
            PL[v] = hi;
            recon<REC>(q[-sd], q[0], q[sd], q[2 * sd], q[3 * sd], lo, hi);
            PR[v] = lo;
        }
        PL[0] = jmax(fl, PL[0]);
        PR[0] = jmax(fl, PR[0]);
        PL[1] = jmax(fl, PL[1]);
        PR[1] = jmax(fl, PR[1]);
        double Fo[NV];
        face_flux<NV, DIM, RIEM>(PL, PR, g, Fo);
#pragma unroll
        for (int v = 0; v < NV; v++) F[v * G.ps + c] = Fo[v];
    }
}

template <int NV>
__global__ __launch_bounds__(256) void k_ref_update(Geo G, const double *U, double *Ub, double dt,
                                                     double dx, double dy, double dz,
                                                     const double *Fx, const double *Fy,
                                                     const double *Fz)
{
    const int ni = G.NX - 6, nj = G.NY - 6, nk = NV == 5 ? G.NZ - 6 : 1;
    const long n = (long)ni * nj * nk;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const int i = 3 + (int)(t % ni);
        const long r = t / ni;
        const int j = 3 + (int)(r % nj);
        const int k = NV == 5 ? 3 + (int)(r / nj) : 0;
        const long c = gidx(G, i, j, k);
#pragma unroll
        for (int v = 0; v < NV; v++) {
            const long cv = v * G.ps + c;
            double val = U[cv] - dt / dx * (Fx[cv] - Fx[cv - 1]) - dt / dy * (Fy[cv] - Fy[cv - G.px]);
            if constexpr (NV == 5) val = val - dt / dz * (Fz[cv] - Fz[cv - (long)G.px * G.NY]);
            Ub[cv] = val;
        }
    }
}

template <int NV>
__global__ __launch_bounds__(64) void k_ref_c2p(Geo G, const double *U, double *P, double g,
                                                 int n_iter, double tol, DevState *st)
{
    const int ni = G.NX - 6, nj = G.NY - 6, nk = NV == 5 ? G.NZ - 6 : 1;
    const long n = (long)ni * nj * nk;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const int i = 3 + (int)(t % ni);
        const long r = t / ni;
        const int j = 3 + (int)(r % nj);
        const int k = NV == 5 ? 3 + (int)(r / nj) : 0;
        const long c = gidx(G, i, j, k);
        double u[NV], p[NV];
#pragma unroll
        for (int v = 0; v < NV; v++) {
            u[v] = U[v * G.ps + c];
            p[v] = P[v * G.ps + c];
        }
        int iters = 0;
        if constexpr (NV == 5) {
            double pn[5];
            if (c2p_1dw(u, p, g, n_iter, tol, pn, iters)) {
#pragma unroll
                for (int v = 0; v < 5; v++) p[v] = pn[v];
            } else {
                atomicAdd(&st->n_rescued, 1ull);
                c2p_cascade_tail(u, p, g, n_iter, tol, iters);
            }
        } else {
            if (!c2p_2d(u, p, g, n_iter, tol, iters)) atomicAdd(&st->n_singular, 1ull);
        }
        if (!all_finite<NV>(p)) atomicAdd(&st->n_unrecovered, 1ull);
#pragma unroll
        for (int v = 0; v < NV; v++) P[v * G.ps + c] = p[v];
    }
}

// AoS (Verona order) <-> SoA planes
template <int NV>
__global__ __launch_bounds__(256) void k_aos_to_soa(Geo G, const double *aos, double *soa)
{
    const long n = (long)G.NX * G.NY * G.NZ;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const int i = (int)(t % G.NX);
        const long r = t / G.NX;
        const long c = gidx(G, i, (int)(r % G.NY), (int)(r / G.NY));
#pragma unroll
        for (int v = 0; v < NV; v++) soa[v * G.ps + c] = aos[t * NV + v];
    }
}

template <int NV>
__global__ __launch_bounds__(256) void k_soa_to_aos(Geo G, const double *soa, double *aos)
{
    const long n = (long)G.NX * G.NY * G.NZ;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const int i = (int)(t % G.NX);
        const long r = t / G.NX;
        const long c = gidx(G, i, (int)(r % G.NY), (int)(r / G.NY));
#pragma unroll
        for (int v = 0; v < NV; v++) aos[t * NV + v] = soa[v * G.ps + c];
    }
}

}  // namespace vrn

// ====================================================================== host side
using namespace vrn;

// fast FP64 form of the sweeps, compiled in vrn_grid_fast.hip
void vrn_fast_dispatch_stage5(hipStream_t s, const vrn::SweepArgs &a, const double dxs[3], int rec,
                              int riem, void (*timer)(void *, int, int), void *tctx);
void vrn_fast_dispatch_stage4(hipStream_t s, const vrn::SweepArgs &a, const double dxs[3], int rec,
                              int riem, void (*timer)(void *, int, int), void *tctx);

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(expr)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            return fail(VRN_E_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));        \
    } while (0)

const char *kKernelNames[] = {"xsweep", "ysweep", "zsweep_c2p", "finish_2d", "rescue", "ghost",
                              "cfl",    "fluxes", "update",     "c2p",       "xysweep"};
enum { KN_X, KN_Y, KN_ZC, KN_F2D, KN_RESCUE, KN_GHOST, KN_CFL, KN_FLUX, KN_UPD, KN_C2P, KN_XY, KN_COUNT };

}  // namespace

struct vrn_ctx {
    int ndim = 3, NV = 5;
    Geo G{};
    double gamma = 4.0 / 3.0;
    double *buf = nullptr;   // P, Pn, U, Uh and the staging buffer
    double *fbuf = nullptr;  // flux slots, allocated by the first per-kernel call that uses them
    int fuse_xy = 1;  // 3D FP64 forms: x direction inside the y march
    // 3D FP64 forms with the fused x-y march and rules on both x faces:
    // the stage-end ghost passes leave the x faces unfilled (the x-y march,
    // their only reader in the step, reads ghost columns through the rules);
    // `pend` fills them (settle) before anything else reads P's ghost cells
    bool ghosts_pending = false;
    vrn::GhostArgs pend{};
    double *P = nullptr, *Pn = nullptr, *U = nullptr, *Uh = nullptr, *F[3] = {nullptr, nullptr, nullptr};
    double *stg = nullptr;
    int *rescue = nullptr;
    DevState *st = nullptr;
    DevState *st_host = nullptr;  // pinned
    hipStream_t s = nullptr;
    // nozzle
    double noz_r = 0, noz_dx = 1, noz_dy = 1, noz_rho = 0, noz_u = 0, noz_umag = 0, noz_eps = 0;
    bool have_noz = false;
    // rate_bits holds the max signal rate of the current P for spacings
    // rate_dx only while rate_ok: a CFL step leaves it there (its last stage
    // reduces the new state's rate); anything else that writes P, a fixed-dt
    // step or new spacings invalidate it, and the next CFL step recomputes it.
    bool rate_ok = false;
    double rate_dx[3] = {0, 0, 0};
    double dxs[3] = {1, 1, 1};
    // tuning
    int rows_per_wave = 8;
    int seg_len = 96;
    int xseg = 1;
    // timing
    int timing = 0;  // 0 off, 1 every kernel, 2 the directional sweeps only
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_used;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_free;
    double t_ms[KN_COUNT] = {0};
    long t_n[KN_COUNT] = {0};
};

namespace {

int harvest(vrn_ctx *c)
{
    if (c->ev_used.empty()) return 0;
    HIPCHK(hipStreamSynchronize(c->s));
    for (auto &e : c->ev_used) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, e.second.first, e.second.second));
        c->t_ms[e.first] += ms;
        c->t_n[e.first] += 1;
        c->ev_free.push_back(e.second);
    }
    c->ev_used.clear();
    return 0;
}

// Event bracket around one launch when timing is on.
struct Timed {
    vrn_ctx *c;
    int kn;
    std::pair<hipEvent_t, hipEvent_t> ev{};
    bool on;
    Timed(vrn_ctx *c_, int kn_)
        : c(c_), kn(kn_),
          on(c_->timing == 1 ||
             (c_->timing == 2 && (kn_ == KN_X || kn_ == KN_Y || kn_ == KN_ZC || kn_ == KN_XY)))
    {
        if (!on) return;
        if (c->ev_free.empty()) {
            if (c->ev_used.size() >= 8192) harvest(c);
        }
        if (c->ev_free.empty()) {
            hipEvent_t a, b;
            if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
                on = false;
                return;
            }
            c->ev_free.push_back({a, b});
        }
        ev = c->ev_free.back();
        c->ev_free.pop_back();
        (void)hipEventRecord(ev.first, c->s);
    }
    ~Timed()
    {
        if (!on) return;
        (void)hipEventRecord(ev.second, c->s);
        c->ev_used.push_back({kn, ev});
    }
};

// The flux slots (zeroed) on first use.
int ensure_flux(vrn_ctx *c)
{
    if (c->fbuf) return 0;
    const size_t arr = (size_t)c->NV * c->G.ps;
    const int nf = c->ndim == 3 ? 3 : 2;
    hipError_t e = hipMalloc(&c->fbuf, sizeof(double) * arr * nf);
    if (e != hipSuccess) {
        c->fbuf = nullptr;
        return fail(VRN_E_NOMEM, "hipMalloc of the flux slots (%zu bytes): %s", sizeof(double) * arr * nf,
                    hipGetErrorString(e));
    }
    e = hipMemsetAsync(c->fbuf, 0, sizeof(double) * arr * nf, c->s);
    if (e != hipSuccess) return fail(VRN_E_HIP, "hipMemsetAsync: %s", hipGetErrorString(e));
    for (int d = 0; d < nf; d++) c->F[d] = c->fbuf + d * arr;
    return 0;
}

double *slot_ptr(vrn_ctx *c, int slot)
{
    switch (slot) {
    case VRN_SLOT_P: return c->P;
    case VRN_SLOT_U: return c->U;
    case VRN_SLOT_UHALF: return c->Uh;
    case VRN_SLOT_FX: return c->F[0];
    case VRN_SLOT_FY: return c->F[1];
    case VRN_SLOT_FZ: return c->ndim == 3 ? c->F[2] : nullptr;
    default: return nullptr;
    }
}

int grid_blocks(long n, int block)
{
    long b = (n + block - 1) / block;
    if (b > 65536) b = 65536;
    if (b < 1) b = 1;
    return (int)b;
}

template <int NV> void launch_ghost(vrn_ctx *c, GhostArgs &ga)
{
    const Geo &G = c->G;
    long n;
    if (NV == 5)
        n = (ga.part == 2 ? 0L : 6L * G.NX * G.NY + 6L * G.NX * (G.NZ - 6)) +
            (ga.part == 1 ? 0L : 6L * (G.NY - 6) * (G.NZ - 6));
    else
        n = 6L * G.NX + 6L * (G.NY - 6);
    Timed tm(c, KN_GHOST);
    hipLaunchKernelGGL(k_ghost<NV>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->s, ga);
}

// rules: per face 0 identity, 1 periodic, 2 outflow (nozzle counts as outflow)
void make_rules(const int bc[6], bool periodic_only, bool nonperiodic_only, int *lo, int *hi,
                int &nozzle)
{
    nozzle = 0;
    for (int d = 0; d < 3; d++) {
        for (int s = 0; s < 2; s++) {
            int b = bc[2 * d + s];
            int r = 0;
            if (b == VRN_BC_PERIODIC && !nonperiodic_only) r = 1;
            if ((b == VRN_BC_OUTFLOW || b == VRN_BC_NOZZLE) && !periodic_only) r = 2;
            (s == 0 ? lo : hi)[d] = r;
        }
    }
    if (bc[4] == VRN_BC_NOZZLE && !periodic_only) nozzle = 1;
}

void fill_nozzle_args(vrn_ctx *c, GhostArgs &ga, double phi)
{
    ga.noz_r2 = c->noz_r * c->noz_r;
    ga.noz_dx = c->noz_dx;
    ga.noz_dy = c->noz_dy;
    ga.noz_rho = c->noz_rho;
    ga.noz_u = c->noz_u;
    double nx = c->noz_eps * cos(phi), ny = c->noz_eps * sin(phi), nz = 1.0;
    double nn = sqrt(nx * nx + ny * ny + nz * nz);
    ga.noz_ux = c->noz_umag * (nx / nn);
    ga.noz_uy = c->noz_umag * (ny / nn);
    ga.noz_uz = c->noz_umag * (nz / nn);
}

int check_ctx(vrn_ctx *c)
{
    if (!c) return fail(VRN_E_ARG, "null context");
    return 0;
}

}  // namespace

// ====================================================================== C ABI
// Work partition of the sweeps for this grid: enough waves to fill 256 CUs
// (x sweep: whole rows per wave while rows are plentiful, else segments of
// one row; y / z sweeps: shorter marching segments when columns are few).
// The partition changes no arithmetic: every face is solved once and every
// cell updated once from the same operands.
// Fill the x faces the last stage-end ghost pass left unfilled, if any.
static int settle(vrn_ctx *c)
{
    if (!c->ghosts_pending) return 0;
    c->ghosts_pending = false;
    c->pend.Pd = c->P;
    c->pend.Pi = c->P;
    c->pend.Pg = c->P;
    c->pend.part = 2;
    launch_ghost<5>(c, c->pend);
    HIPCHK(hipGetLastError());
    return 0;
}

static void partition_sweeps(vrn_ctx *c)
{
    const int Nxi = c->G.NX - 6, Nyi = c->G.NY - 6, Nzi = c->ndim == 3 ? c->G.NZ - 6 : 1;
    const long target = 8192;  // waves
    const long nrows = (long)Nyi * Nzi;
    long rpw = nrows / target;
    c->rows_per_wave = (int)(rpw < 1 ? 1 : (rpw > 8 ? 8 : rpw));
    c->xseg = 1;
    if (nrows < target) {
        long want = (target + nrows - 1) / nrows, most = Nxi / 32;
        c->xseg = (int)(want < most ? want : most);
        if (c->xseg < 1) c->xseg = 1;
        if (c->xseg > 1) c->rows_per_wave = 1;
    }
    const long colwaves = (long)((Nxi + 63) / 64) * (c->ndim == 3 ? (Nyi < Nzi ? Nyi : Nzi) : 1);
    const int along = c->ndim == 3 ? (Nyi > Nzi ? Nyi : Nzi) : Nyi;
    long seg = colwaves * along / 2048;
    c->seg_len = (int)(seg < 24 ? 24 : (seg > 96 ? 96 : seg));
}

extern "C" {

const char *vrn_last_error(void) { return g_err.c_str(); }

int vrn_create(int ndim, int nx, int ny, int nz, double gamma, vrn_ctx **out)
{
    if (!out) return fail(VRN_E_ARG, "out is null");
    *out = nullptr;
    if (ndim != 2 && ndim != 3) return fail(VRN_E_ARG, "ndim must be 2 or 3 (got %d)", ndim);
    if (nx < 4 || ny < 4 || (ndim == 3 && nz < 4))
        return fail(VRN_E_ARG, "interior sizes must be >= 4 (got %d %d %d)", nx, ny, nz);
    if (!(gamma > 1.0)) return fail(VRN_E_ARG, "gamma must be > 1");
    vrn_ctx *c = new vrn_ctx();
    c->ndim = ndim;
    c->NV = ndim == 3 ? 5 : 4;
    c->gamma = gamma;
    c->G.NX = nx + 6;
    c->G.NY = ny + 6;
    c->G.NZ = ndim == 3 ? nz + 6 : 1;
    c->G.px = ((XOFF + c->G.NX + 15) / 16) * 16;
    c->G.ps = (long)c->G.px * c->G.NY * c->G.NZ;
    partition_sweeps(c);
    long cells = (long)c->G.NX * c->G.NY * c->G.NZ;
    if ((long)c->G.ps >= (1L << 31)) {
        delete c;
        return fail(VRN_E_ARG, "grid too large for 32-bit rescue indices");
    }
    const int NV = c->NV;
    // P, Pn, U, Uh; the flux slots of the per-kernel interface (vrn_fluxes /
    // vrn_update) are allocated on their first use: the fused step never
    // touches them, so a step-only context holds 4 arrays + staging
    const int nplanes_arrays = 4;
    size_t arr = (size_t)NV * c->G.ps;
    size_t total = arr * nplanes_arrays + (size_t)NV * cells;
    hipError_t e = hipStreamCreate(&c->s);
    if (e != hipSuccess) {
        delete c;
        return fail(VRN_E_HIP, "hipStreamCreate: %s", hipGetErrorString(e));
    }
    e = hipMalloc(&c->buf, total * sizeof(double));
    if (e != hipSuccess) {
        (void)hipStreamDestroy(c->s);
        delete c;
        return fail(VRN_E_NOMEM, "hipMalloc(%zu bytes): %s", total * sizeof(double), hipGetErrorString(e));
    }
    double *p = c->buf;
    c->P = p; p += arr;
    c->Pn = p; p += arr;
    c->U = p; p += arr;
    c->Uh = p; p += arr;
    c->stg = p;
    e = hipMalloc(&c->rescue, sizeof(int) * (size_t)cells);
    if (e == hipSuccess) e = hipMalloc(&c->st, sizeof(DevState));
    if (e == hipSuccess) e = hipHostMalloc(&c->st_host, sizeof(DevState), hipHostMallocDefault);
    if (e == hipSuccess) e = hipMemsetAsync(c->buf, 0, total * sizeof(double), c->s);
    if (e == hipSuccess) e = hipMemsetAsync(c->st, 0, sizeof(DevState), c->s);
    if (e == hipSuccess) e = hipStreamSynchronize(c->s);
    if (e != hipSuccess) {
        vrn_destroy(c);
        return fail(VRN_E_NOMEM, "allocation failed: %s", hipGetErrorString(e));
    }
    *out = c;
    return 0;
}

int vrn_destroy(vrn_ctx *c)
{
    if (!c) return 0;
    if (c->s) (void)hipStreamSynchronize(c->s);
    for (auto &e : c->ev_used) {
        (void)hipEventDestroy(e.second.first);
        (void)hipEventDestroy(e.second.second);
    }
    for (auto &e : c->ev_free) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    if (c->buf) (void)hipFree(c->buf);
    if (c->fbuf) (void)hipFree(c->fbuf);
    if (c->rescue) (void)hipFree(c->rescue);
    if (c->st) (void)hipFree(c->st);
    if (c->st_host) (void)hipHostFree(c->st_host);
    if (c->s) (void)hipStreamDestroy(c->s);
    delete c;
    return 0;
}

int vrn_sync(vrn_ctx *c)
{
    if (int r = check_ctx(c)) return r;
    HIPCHK(hipStreamSynchronize(c->s));
    return 0;
}

int vrn_upload(vrn_ctx *c, int slot, const double *host)
{
    if (int r = check_ctx(c)) return r;
    if (slot >= VRN_SLOT_FX && slot <= VRN_SLOT_FZ)
        if (int r = ensure_flux(c)) return r;
    double *d = slot_ptr(c, slot);
    if (!d || !host) return fail(VRN_E_ARG, "bad slot %d or null host pointer", slot);
    const Geo &G = c->G;
    size_t bytes = sizeof(double) * (size_t)c->NV * G.NX * G.NY * G.NZ;
    HIPCHK(hipMemcpyAsync(c->stg, host, bytes, hipMemcpyHostToDevice, c->s));
    long n = (long)G.NX * G.NY * G.NZ;
    if (c->NV == 5)
        hipLaunchKernelGGL(k_aos_to_soa<5>, dim3(grid_blocks(n, 256)), dim3(256), 0, c->s, G, c->stg, d);
    else
        hipLaunchKernelGGL(k_aos_to_soa<4>, dim3(grid_blocks(n, 256)), dim3(256), 0, c->s, G, c->stg, d);
    HIPCHK(hipGetLastError());
    if (slot == VRN_SLOT_P) {
        c->rate_ok = false;
        c->ghosts_pending = false;
    }
    HIPCHK(hipStreamSynchronize(c->s));
    return 0;
}

int vrn_download(vrn_ctx *c, int slot, double *host)
{
    if (int r = check_ctx(c)) return r;
    if (slot >= VRN_SLOT_FX && slot <= VRN_SLOT_FZ)
        if (int r = ensure_flux(c)) return r;
    double *d = slot_ptr(c, slot);
    if (!d || !host) return fail(VRN_E_ARG, "bad slot %d or null host pointer", slot);
    if (slot == VRN_SLOT_P)
        if (int r = settle(c)) return r;
    const Geo &G = c->G;
    long n = (long)G.NX * G.NY * G.NZ;
    if (c->NV == 5)
        hipLaunchKernelGGL(k_soa_to_aos<5>, dim3(grid_blocks(n, 256)), dim3(256), 0, c->s, G, d, c->stg);
    else
        hipLaunchKernelGGL(k_soa_to_aos<4>, dim3(grid_blocks(n, 256)), dim3(256), 0, c->s, G, d, c->stg);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(host, c->stg, sizeof(double) * (size_t)c->NV * n, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    return 0;
}

int vrn_copy(vrn_ctx *c, int dst_slot, int src_slot)
{
    if (int r = check_ctx(c)) return r;
    if ((dst_slot >= VRN_SLOT_FX && dst_slot <= VRN_SLOT_FZ) || (src_slot >= VRN_SLOT_FX && src_slot <= VRN_SLOT_FZ))
        if (int r = ensure_flux(c)) return r;
    double *d = slot_ptr(c, dst_slot), *s_ = slot_ptr(c, src_slot);
    if (!d || !s_ || d == s_) return fail(VRN_E_ARG, "bad slots %d <- %d", dst_slot, src_slot);
    if (src_slot == VRN_SLOT_P)
        if (int r = settle(c)) return r;
    HIPCHK(hipMemcpyAsync(d, s_, sizeof(double) * (size_t)c->NV * c->G.ps, hipMemcpyDeviceToDevice, c->s));
    if (dst_slot == VRN_SLOT_P) {
        c->rate_ok = false;
        c->ghosts_pending = false;
    }
    return 0;
}

int vrn_slot_bytes(vrn_ctx *c, int64_t *bytes)
{
    if (int r = check_ctx(c)) return r;
    if (bytes) *bytes = (int64_t)(sizeof(double) * (size_t)c->NV * c->G.ps);
    return 0;
}

int vrn_prim_to_cons(vrn_ctx *c)
{
    if (int r = check_ctx(c)) return r;
    if (int r = settle(c)) return r;
    long n = (long)c->G.NX * c->G.NY * c->G.NZ;
    if (c->NV == 5)
        hipLaunchKernelGGL(k_ref_ptou<5>, dim3(grid_blocks(n, 256)), dim3(256), 0, c->s, c->G, c->P, c->U, c->gamma);
    else
        hipLaunchKernelGGL(k_ref_ptou<4>, dim3(grid_blocks(n, 256)), dim3(256), 0, c->s, c->G, c->P, c->U, c->gamma);
    HIPCHK(hipGetLastError());
    return 0;
}

int vrn_limit(vrn_ctx *c, double floor_)
{
    if (int r = check_ctx(c)) return r;
    if (int r = settle(c)) return r;
    long n = (long)c->G.NX * c->G.NY * c->G.NZ;
    if (c->NV == 5)
        hipLaunchKernelGGL(k_ref_limit<5>, dim3(grid_blocks(n, 256)), dim3(256), 0, c->s, c->G, c->P, floor_);
    else
        hipLaunchKernelGGL(k_ref_limit<4>, dim3(grid_blocks(n, 256)), dim3(256), 0, c->s, c->G, c->P, floor_);
    HIPCHK(hipGetLastError());
    c->rate_ok = false;
    return 0;
}

}  // extern "C"

template <int NV, int DIM>
static void launch_ref_flux(vrn_ctx *c, int rec, int riem, double fl)
{
    const Geo &G = c->G;
    long n = (long)(G.NX - 5) * (G.NY - 5) * (NV == 5 ? G.NZ - 5 : 1);
    dim3 gb(grid_blocks(n, 256)), bb(256);
    double *F = c->F[DIM];
#define VRN_FL(R_, M_) hipLaunchKernelGGL((k_ref_fluxes<NV, DIM, R_, M_>), gb, bb, 0, c->s, G, c->P, F, c->gamma, fl)
    if (riem == VRN_RIEMANN_HLL) {
        if (rec == VRN_REC_WENOZ) VRN_FL(REC_WENOZ, RIEM_HLL);
        else if (rec == VRN_REC_PPM) VRN_FL(REC_PPM, RIEM_HLL);
        else VRN_FL(REC_MINMOD, RIEM_HLL);
    } else {
        if (rec == VRN_REC_WENOZ) VRN_FL(REC_WENOZ, RIEM_HLLC);
        else if (rec == VRN_REC_PPM) VRN_FL(REC_PPM, RIEM_HLLC);
        else VRN_FL(REC_MINMOD, RIEM_HLLC);
    }
#undef VRN_FL
}

extern "C" {

int vrn_fluxes(vrn_ctx *c, int dim, double floor_, int rec, int riemann)
{
    if (int r = check_ctx(c)) return r;
    if (dim < 0 || dim >= c->ndim) return fail(VRN_E_ARG, "dim %d out of range", dim);
    if (rec < 0 || rec > 2 || riemann < 0 || riemann > 1) return fail(VRN_E_ARG, "bad rec/riemann");
    if (int r = ensure_flux(c)) return r;
    if (int r = settle(c)) return r;
    Timed tm(c, KN_FLUX);
    if (c->NV == 5) {
        if (dim == 0) launch_ref_flux<5, 0>(c, rec, riemann, floor_);
        else if (dim == 1) launch_ref_flux<5, 1>(c, rec, riemann, floor_);
        else launch_ref_flux<5, 2>(c, rec, riemann, floor_);
    } else {
        if (dim == 0) launch_ref_flux<4, 0>(c, rec, riemann, floor_);
        else launch_ref_flux<4, 1>(c, rec, riemann, floor_);
    }
    HIPCHK(hipGetLastError());
    return 0;
}

int vrn_update(vrn_ctx *c, int dst_slot, double dt, double dx, double dy, double dz)
{
    if (int r = check_ctx(c)) return r;
    if (dst_slot != VRN_SLOT_U && dst_slot != VRN_SLOT_UHALF) return fail(VRN_E_ARG, "dst must be U or UHALF");
    if (int r = ensure_flux(c)) return r;
    double *dst = slot_ptr(c, dst_slot);
    const Geo &G = c->G;
    long n = (long)(G.NX - 6) * (G.NY - 6) * (c->NV == 5 ? G.NZ - 6 : 1);
    Timed tm(c, KN_UPD);
    if (c->NV == 5)
        hipLaunchKernelGGL(k_ref_update<5>, dim3(grid_blocks(n, 256)), dim3(256), 0, c->s, G, c->U, dst, dt, dx, dy, dz, c->F[0], c->F[1], c->F[2]);
    else
        hipLaunchKernelGGL(k_ref_update<4>, dim3(grid_blocks(n, 256)), dim3(256), 0, c->s, G, c->U, dst, dt, dx, dy, dz, c->F[0], c->F[1], c->F[1]);
    HIPCHK(hipGetLastError());
    return 0;
}

int vrn_cons_to_prim(vrn_ctx *c, int src_slot, int n_iter, double tol)
{
    if (int r = check_ctx(c)) return r;
    if (src_slot != VRN_SLOT_U && src_slot != VRN_SLOT_UHALF) return fail(VRN_E_ARG, "src must be U or UHALF");
    if (n_iter < 0) return fail(VRN_E_ARG, "n_iter must be >= 0");
    if (int r = settle(c)) return r;
    const Geo &G = c->G;
    long n = (long)(G.NX - 6) * (G.NY - 6) * (c->NV == 5 ? G.NZ - 6 : 1);
    Timed tm(c, KN_C2P);
    if (c->NV == 5)
        hipLaunchKernelGGL(k_ref_c2p<5>, dim3(grid_blocks(n, 64)), dim3(64), 0, c->s, G, slot_ptr(c, src_slot), c->P, c->gamma, n_iter, tol, c->st);
    else
        hipLaunchKernelGGL(k_ref_c2p<4>, dim3(grid_blocks(n, 64)), dim3(64), 0, c->s, G, slot_ptr(c, src_slot), c->P, c->gamma, n_iter, tol, c->st);
    HIPCHK(hipGetLastError());
    c->rate_ok = false;
    return 0;
}

int vrn_set_nozzle(vrn_ctx *c, double radius, double dx, double dy, double rho, double u, double umag, double eps)
{
    if (int r = check_ctx(c)) return r;
    if (c->ndim != 3) return fail(VRN_E_ARG, "nozzle needs a 3D context");
    c->noz_r = radius;
    c->noz_dx = dx;
    c->noz_dy = dy;
    c->noz_rho = rho;
    c->noz_u = u;
    c->noz_umag = umag;
    c->noz_eps = eps;
    c->have_noz = true;
    return 0;
}

static int fill_ghosts_phi(vrn_ctx *c, const int bc[6], double phi, bool periodic_only, bool nonperiodic_only)
{
    GhostArgs ga{};
    ga.geo = c->G;
    ga.Pd = c->P;
    ga.Pi = c->P;
    ga.Pg = c->P;
    make_rules(bc, periodic_only, nonperiodic_only, ga.rule_lo, ga.rule_hi, ga.nozzle);
    if (ga.nozzle) {
        if (!c->have_noz) return fail(VRN_E_STATE, "VRN_BC_NOZZLE used before vrn_set_nozzle");
        fill_nozzle_args(c, ga, phi);
    }
    ga.do_limit = 0;
    ga.st = nullptr;
    if (c->NV == 5) launch_ghost<5>(c, ga);
    else launch_ghost<4>(c, ga);
    HIPCHK(hipGetLastError());
    return 0;
}

int vrn_fill_ghosts(vrn_ctx *c, const int bc[6])
{
    if (int r = check_ctx(c)) return r;
    if (!bc) return fail(VRN_E_ARG, "bc is null");
    for (int f = 0; f < 6; f++)
        if (bc[f] < 0 || bc[f] > 3 || (bc[f] == VRN_BC_NOZZLE && f != 4))
            return fail(VRN_E_ARG, "bad boundary kind %d on face %d", bc[f], f);
    if (int r = settle(c)) return r;
    return fill_ghosts_phi(c, bc, 0.0, false, false);
}

int vrn_get_counters(vrn_ctx *c, int64_t *n_unrecovered, int64_t *n_rescued)
{
    if (int r = check_ctx(c)) return r;
    HIPCHK(hipMemcpyAsync(c->st_host, c->st, sizeof(DevState), hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    if (n_unrecovered) *n_unrecovered = (int64_t)c->st_host->n_unrecovered;
    if (n_rescued) *n_rescued = (int64_t)(c->st_host->n_rescued + c->st_host->rescue_count);
    return 0;
}

int vrn_reset_counters(vrn_ctx *c)
{
    if (int r = check_ctx(c)) return r;
    HIPCHK(hipMemcpyAsync(c->st_host, c->st, sizeof(DevState), hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    c->st_host->n_unrecovered = 0;
    c->st_host->n_rescued = 0;
    c->st_host->n_singular = 0;
    c->st_host->rescue_count = 0;
    HIPCHK(hipMemcpyAsync(c->st, c->st_host, sizeof(DevState), hipMemcpyHostToDevice, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------- fused step
// Event bracket callback used by the sweep launchers (kn: sweep index).
static void timer_cb(void *p, int kn, int phase)
{
    vrn_ctx *c = static_cast<vrn_ctx *>(p);
    if (!c->timing) return;
    static thread_local Timed *open_tm = nullptr;
    if (phase == 0) {
        open_tm = new Timed(c, kn);
    } else if (open_tm) {
        delete open_tm;
        open_tm = nullptr;
    }
}

static void run_stage(vrn_ctx *c, const SweepArgs &a, int rec, int riem, int prec)
{
    const bool fp32 = prec == VRN_PREC_FP32;
    if (prec == VRN_PREC_FP64_FAST) {
        if (c->NV == 5)
            vrn_fast_dispatch_stage5(c->s, a, c->dxs, rec, riem, timer_cb, c);
        else
            vrn_fast_dispatch_stage4(c->s, a, c->dxs, rec, riem, timer_cb, c);
        return;
    }
    if (c->NV == 5)
        vrn::dispatch_stage<5>(c->s, a, c->dxs, rec, riem, fp32, timer_cb, c);
    else
        vrn::dispatch_stage<4>(c->s, a, c->dxs, rec, riem, fp32, timer_cb, c);
}

extern "C" {

int vrn_step(vrn_ctx *c, const vrn_step_params *p, double *dt_used)
{
    if (int r = check_ctx(c)) return r;
    if (!p) return fail(VRN_E_ARG, "params is null");
    if (p->rec < 0 || p->rec > 2 || p->riemann < 0 || p->riemann > 1 || p->integrator < 0 ||
        p->integrator > 1 || p->precision < 0 || p->precision > 2 || p->n_iter < 0)
        return fail(VRN_E_ARG, "bad step parameters");
    if (!(p->dx > 0) || !(p->dy > 0) || (c->ndim == 3 && !(p->dz > 0)))
        return fail(VRN_E_ARG, "grid spacings must be positive");
    if (!(p->cfl > 0) && !(p->dt > 0)) return fail(VRN_E_ARG, "need dt > 0 or cfl > 0");
    // floors are applied as max(floor, x) by a compare-select, which needs a
    // floor that is a number and not -0
    if (std::isnan(p->floor_) || (p->floor_ == 0.0 && std::signbit(p->floor_)))
        return fail(VRN_E_ARG, "floor must be a number other than -0");
    for (int f = 0; f < 6; f++)
        if (p->bc[f] < 0 || p->bc[f] > 3 || (p->bc[f] == VRN_BC_NOZZLE && (f != 4 || !c->have_noz)))
            return fail(VRN_E_ARG, "bad boundary kind %d on face %d (nozzle needs vrn_set_nozzle)", p->bc[f], f);
    const Geo &G = c->G;
    const bool cflm = p->cfl > 0.0;
    // stage-end ghost rules (the same every stage)
    int rlo[3], rhi[3], rnoz;
    make_rules(p->bc, p->ref_mode != 0, false, rlo, rhi, rnoz);
    // 3D FP64 forms with the fused x-y march and a rule on both x faces
    // (so each x ghost cell is a copy of an interior cell): the stage-end
    // ghost passes leave the x faces unfilled and the x-y march reads ghost
    // columns through the rules.  Anything else first fills them.
    const bool defer = (p->precision == VRN_PREC_FP64_FAST || p->precision == VRN_PREC_FP64) && c->NV == 5 &&
                       c->fuse_xy && rlo[0] && rhi[0];
    if (!defer)
        if (int r = settle(c)) return r;
    // dt: fixed, or CFL from the state at the step start
    const double dz_rate = c->ndim == 3 ? p->dz : 1.0;
    if (cflm) {
        if (!c->rate_ok || c->rate_dx[0] != p->dx || c->rate_dx[1] != p->dy || c->rate_dx[2] != dz_rate) {
            Timed tm(c, KN_CFL);
            hipLaunchKernelGGL(k_reset_rate, dim3(1), dim3(1), 0, c->s, c->st);
            if (c->NV == 5)
                hipLaunchKernelGGL(k_rate<5>, dim3(2048), dim3(256), 0, c->s, G, c->P, c->gamma, 1.0 / p->dx, 1.0 / p->dy, 1.0 / p->dz, c->st);
            else
                hipLaunchKernelGGL(k_rate<4>, dim3(2048), dim3(256), 0, c->s, G, c->P, c->gamma, 1.0 / p->dx, 1.0 / p->dy, 1.0, c->st);
        }
        // the rate of the current P is in rate_bits (from the last stage of
        // the previous CFL step, or from k_rate just now); the nozzle's inflow
        // state enters through the ghost faces and is added here.
        double extra = 0.0;
        if (p->bc[4] == VRN_BC_NOZZLE) {
            GhostArgs gn{};
            fill_nozzle_args(c, gn, p->nozzle_phi);
            const double Pz[5] = {gn.noz_rho, gn.noz_u, gn.noz_ux, gn.noz_uy, gn.noz_uz};
            const double inv[3] = {1.0 / p->dx, 1.0 / p->dy, 1.0 / p->dz};
            extra = cell_rate<5>(Pz, c->gamma, inv);
        }
        hipLaunchKernelGGL(k_finalize_dt, dim3(1), dim3(1), 0, c->s, c->st, p->cfl, extra);
    }
    c->dxs[0] = p->dx;
    c->dxs[1] = p->dy;
    c->dxs[2] = c->ndim == 3 ? p->dz : 1.0;
    // Verona sequencing: the user BC callback runs once at the step start
    if (p->ref_mode) {
        bool any = false;
        for (int f = 0; f < 6; f++) any |= (p->bc[f] == VRN_BC_OUTFLOW || p->bc[f] == VRN_BC_NOZZLE);
        if (any) {
            int r = fill_ghosts_phi(c, p->bc, p->nozzle_phi, false, true);
            if (r) return r;
        }
    }
    SweepArgs a{};
    a.geo = G;
    a.st = c->st;
    a.rescue = c->rescue;
    a.g = c->gamma;
    a.floor_ = p->floor_;
    a.inv_dx[0] = 1.0 / p->dx;
    a.inv_dx[1] = 1.0 / p->dy;
    a.inv_dx[2] = c->ndim == 3 ? 1.0 / p->dz : 0.0;
    a.tol = p->tol;
    a.n_iter = p->n_iter;
    a.dt_fixed = p->dt;
    a.dt_dev = cflm ? 1 : 0;
    a.seg_len = c->seg_len;
    a.rows_per_wave = c->rows_per_wave;
    a.xseg = c->xseg;
    a.fuse_xy = c->fuse_xy;
    a.dx_x = p->dx;
    const int nst = p->integrator == VRN_INT_RK3 ? 3 : 2;
    for (int s = 0; s < nst; s++) {
        // x ghost columns through the rules of the pass that left them unfilled
        a.xmap_lo = defer && c->ghosts_pending ? c->pend.rule_lo[0] : 0;
        a.xmap_hi = defer && c->ghosts_pending ? c->pend.rule_hi[0] : 0;
        a.P = c->P;
        a.Pout = c->Pn;
        a.do_rate = (cflm && s == nst - 1) ? 1 : 0;
        if (p->integrator == VRN_INT_RK3) {
            // Shu-Osher SSP-RK3: U1 = U + dt L(U); U2 = 3/4 U + 1/4 (U1 + dt L(U1));
            // U' = 1/3 U + 2/3 (U2 + dt L(U2))
            a.dtfac = 1.0;
            a.U0 = c->U;
            a.Ub = s == 0 ? c->U : c->Uh;
            a.Uacc = c->Uh;
            a.Uout = s == 2 ? c->U : c->Uh;
            a.a = s == 0 ? 0.0 : (s == 1 ? 0.75 : 1.0 / 3.0);
            a.b = s == 0 ? 1.0 : (s == 1 ? 0.25 : 2.0 / 3.0);
        } else {
            // midpoint RK2 (src/3dim/algos.jl:285, 311): Uhalf = U - dt/2 L(P); U = U - dt L(P*)
            a.dtfac = s == 0 ? 0.5 : 1.0;
            a.U0 = c->U;
            a.Ub = c->U;
            a.Uacc = s == 0 ? c->Uh : c->U;
            a.Uout = a.Uacc;
            a.a = 0.0;
            a.b = 1.0;
        }
        run_stage(c, a, p->rec, p->riemann, p->precision);
        HIPCHK(hipGetLastError());
        // ghosts of the new P: carried (floored) from the old one, then the
        // stage-end boundary rules (ref: periodic exchange only)
        GhostArgs ga{};
        ga.geo = G;
        ga.Pd = c->Pn;
        ga.Pi = c->Pn;
        ga.Pg = c->P;
        make_rules(p->bc, p->ref_mode != 0, false, ga.rule_lo, ga.rule_hi, ga.nozzle);
        if (ga.nozzle) fill_nozzle_args(c, ga, p->nozzle_phi);
        ga.do_limit = 1;
        ga.floor_ = p->floor_;
        ga.st = c->st;
        ga.part = defer ? 1 : 0;
        if (c->NV == 5) launch_ghost<5>(c, ga);
        else launch_ghost<4>(c, ga);
        if (defer) {
            c->pend = ga;
            c->ghosts_pending = true;
        }
        HIPCHK(hipGetLastError());
        std::swap(c->P, c->Pn);
    }
    // a CFL step's last stage left the new state's rate in rate_bits
    c->rate_ok = cflm;
    c->rate_dx[0] = p->dx;
    c->rate_dx[1] = p->dy;
    c->rate_dx[2] = dz_rate;
    if (dt_used) {
        if (cflm) {
            HIPCHK(hipMemcpyAsync(c->st_host, c->st, sizeof(DevState), hipMemcpyDeviceToHost, c->s));
            HIPCHK(hipStreamSynchronize(c->s));
            *dt_used = c->st_host->dt;
        } else {
            *dt_used = p->dt;
        }
    }
    return 0;
}

int vrn_set_partition(vrn_ctx *c, int rows_per_wave, int xseg, int seg_len)
{
    if (int r = check_ctx(c)) return r;
    const int Nxi = c->G.NX - 6;
    if (rows_per_wave < 1 || rows_per_wave > 64 || xseg < 1 || (xseg > 1 && rows_per_wave != 1) ||
        xseg > Nxi || seg_len < 8 || seg_len > 4096)
        return fail(VRN_E_ARG, "bad sweep partition (rows_per_wave %d, xseg %d, seg_len %d)", rows_per_wave,
                    xseg, seg_len);
    // a row segment streams at least 34 cells (x0-1 .. x1 with >= 32 interior
    // cells) unless the row itself is shorter
    if (xseg > 1 && (Nxi + xseg - 1) / xseg < 32)
        return fail(VRN_E_ARG, "row segments shorter than 32 cells (nx %d, xseg %d)", Nxi, xseg);
    c->rows_per_wave = rows_per_wave;
    c->xseg = xseg;
    c->seg_len = seg_len;
    return 0;
}

int vrn_set_xy_fusion(vrn_ctx *c, int on)
{
    if (int r = check_ctx(c)) return r;
    c->fuse_xy = on ? 1 : 0;
    return 0;
}

int vrn_get_partition(vrn_ctx *c, int *rows_per_wave, int *xseg, int *seg_len)
{
    if (int r = check_ctx(c)) return r;
    if (rows_per_wave) *rows_per_wave = c->rows_per_wave;
    if (xseg) *xseg = c->xseg;
    if (seg_len) *seg_len = c->seg_len;
    return 0;
}

int vrn_timing_enable(vrn_ctx *c, int enable)
{
    if (int r = check_ctx(c)) return r;
    if (int r = harvest(c)) return r;
    if (enable < 0 || enable > 2) return fail(VRN_E_ARG, "timing mode %d (0 off, 1 all kernels, 2 sweeps)", enable);
    c->timing = enable;
    for (int k = 0; k < KN_COUNT; k++) {
        c->t_ms[k] = 0;
        c->t_n[k] = 0;
    }
    return 0;
}

int vrn_timing_report(vrn_ctx *c, char *names, int name_len, int64_t *counts, double *mean_ms, int cap, int *n_out)
{
    if (int r = check_ctx(c)) return r;
    if (int r = harvest(c)) return r;
    int n = 0;
    for (int k = 0; k < KN_COUNT && n < cap; k++) {
        if (c->t_n[k] == 0) continue;
        if (names && name_len > 0) {
            strncpy(names + (size_t)n * name_len, kKernelNames[k], name_len - 1);
            names[(size_t)n * name_len + name_len - 1] = 0;
        }
        if (counts) counts[n] = c->t_n[k];
        if (mean_ms) mean_ms[n] = c->t_ms[k] / c->t_n[k];
        n++;
    }
    if (n_out) *n_out = n;
    return 0;
}

}  // extern "C"
