// vrn_tubes.hip -- thousands of independent 1D relativistic shock tubes of
// uneven length advanced by one persistent, hipGraph-launched kernel.
//
// Each tube follows Verona1D.HARM_HLL (src/1dim/algos.jl:58-315): SoA
// primitives (rho, u, u^x) without ghost cells, reconstruction into PR[i-1],
// PL[i] over the reference's index range, a face flux for every face,
// midpoint RK2 (Uhalf = U - dt/2 dx dF, then U -= dt/dx dF) on cells
// 1 .. N-3 (0-based; cells 0, N-2, N-1 are never updated), and the 3x3
// Newton con2prim of src/1dim/structs.jl:78-105 on every cell, warm-started
// from P^n in both stages.  Face states the reconstruction never writes keep
// the tube's initial primitives, as the reference's deepcopy'd PR/PL do.
//
// Execution: a persistent grid of workgroups pulls tube indices from an
// atomic queue (tubes pre-sorted longest first, LPT) and runs every step of
// a tube inside one workgroup; tubes are independent, so no grid-wide
// synchronisation is needed.  Within a tube the workgroup streams over the
// cells in chunks of blockDim cells, one cell per lane, exchanging face
// states and fluxes through LDS with a carry across chunks.
#include "../../include/verona_hip.h"
#include "vrn_physics.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

namespace vrn {

// 1dim/structs.jl:58-76  PtoU / PtoF (w = (u + p) + rho association)
__device__ __forceinline__ void p2u1(const double *P, double *U, double g)
{
    double gamma = vrn_sqrt1(P[2] * P[2] + 1.0);
    double pr = (g - 1.0) * P[1];
    U[0] = P[0] * gamma;
    U[1] = -(P[1] + pr + P[0]) * (gamma * gamma) + pr;
    U[2] = (P[1] + pr + P[0]) * gamma * P[2];
}

__device__ __forceinline__ void p2f1(const double *P, double *F, double g)
{
    double gamma = vrn_sqrt1(P[2] * P[2] + 1.0);
    double pr = (g - 1.0) * P[1];
    F[0] = P[0] * P[2];
    F[1] = -(P[1] + pr + P[0]) * gamma * P[2];
    F[2] = pr + (P[1] + pr + P[0]) * (P[2] * P[2]);
}

// 1dim/algos.jl:127-171 face flux from PL (left), PR (right)
template <int RIEM>
__device__ __forceinline__ void face1(const double *PL, const double *PR, double g, double *F)
{
    double UL[3], UR[3], FL[3], FR[3];
    p2u1(PR, UR, g);
    p2u1(PL, UL, g);
    p2f1(PR, FR, g);
    p2f1(PL, FL, g);
    double vL = PL[2] / vrn_sqrt1(PL[2] * PL[2] + 1.0);
    double vR = PR[2] / vrn_sqrt1(PR[2] * PR[2] + 1.0);
    double CL = sound_speed(PL[0], PL[1], g);
    double CR = sound_speed(PR[0], PR[1], g);
    double lpL, lmL, lpR, lmR;
    lambdas(vL, CL, PL[2] * PL[2] + 1.0, lpL, lmL);
    lambdas(vR, CR, PR[2] * PR[2] + 1.0, lpR, lmR);
    double Cmax = jmax(lpL, lpR);
    double Cmin = -jmin(lmL, lmR);
    riemann_solve<3, RIEM>(UL, UR, FL, FR, vL, vR, (g - 1.0) * PL[1], (g - 1.0) * PR[1], Cmax, Cmin,
                           2, F);
}

// 1dim/structs.jl:78-105: Newton on (rho, u, u^x) with F_ptoU / Jacobian and
// the 3x3 closed-form solve of StaticArrays.jl (adjugate / det, det =
// dot(col1, cross(col2, col3))); stops when norm(dx) < tol before applying dx.
__device__ __forceinline__ void c2p1(const double *U, double *x, double g, int n_iter, double tol)
{
    for (int it = 0; it < n_iter; it++) {
        // once all three unknowns are NaN every further iteration returns NaN:
        // stopping here gives the same result as running out the iterations
        if (__builtin_isnan(x[0]) && __builtin_isnan(x[1]) && __builtin_isnan(x[2])) break;
        double gam = vrn_sqrt1(x[2] * x[2] + 1.0);
        double w = g * x[1] + x[0];
        double f0 = gam * x[0] - U[0];
        double f1 = (g - 1.0) * x[1] - gam * gam * w - U[1];
        double f2 = x[2] * gam * w - U[2];
        double gj = vrn_sqrt1(1.0 + x[2] * x[2]);
        // column-major a[0..8]
        double a0 = gj, a1 = -(gj * gj), a2 = gj * x[2];
        double a3 = 0.0, a4 = (g - 1.0) - gj * gj * g, a5 = gj * x[2] * g;
        double a6 = x[0] * x[2] / gj, a7 = -2.0 * x[2] * w, a8 = (2.0 * (x[2] * x[2]) + 1.0) / gj * w;
        double c0 = a4 * a8 - a5 * a7;
        double c1 = a5 * a6 - a3 * a8;
        double c2 = a3 * a7 - a4 * a6;
        double d = a0 * c0 + a1 * c1 + a2 * c2;
        double d0 = ((a4 * a8 - a7 * a5) * f0 + (a6 * a5 - a3 * a8) * f1 + (a3 * a7 - a6 * a4) * f2) / d;
        double d1 = ((a7 * a2 - a1 * a8) * f0 + (a0 * a8 - a6 * a2) * f1 + (a6 * a1 - a0 * a7) * f2) / d;
        double d2 = ((a1 * a5 - a4 * a2) * f0 + (a3 * a2 - a0 * a5) * f1 + (a0 * a4 - a3 * a1) * f2) / d;
        if (vrn_sqrt(d0 * d0 + d1 * d1 + d2 * d2) < tol) break;
        x[0] = x[0] - d0;
        x[1] = x[1] - d1;
        x[2] = x[2] - d2;
    }
}

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

template <int REC>
__device__ __forceinline__ void rec_range(int N, int &lo, int &hi)
{
    if (REC == REC_MINMOD) {
        lo = 1;
        hi = N - 1;
    } else {
        lo = 2;
        hi = N - 2;
    }
}

// One RK stage of one tube by the whole workgroup.  s = 0: Pin = P, out to
// (Uh, Ph) with coefficient 0.5*dt/dx; s = 1: Pin = Ph, out to (U, P) with dt/dx.
template <int REC, int RIEM>
__device__ void tube_stage(const TubeArgs &a, int tb, int s, double *lds)
{
    const long o = a.off[tb];
    const int N = (int)(a.off[tb + 1] - o);
    const double g = a.g;
    const double coef = s == 0 ? 0.5 * a.dt[tb] / a.dx[tb] : a.dt[tb] / a.dx[tb];
    const double *Pin[3] = {s == 0 ? a.P[0] + o : a.Ph[0] + o, s == 0 ? a.P[1] + o : a.Ph[1] + o,
                            s == 0 ? a.P[2] + o : a.Ph[2] + o};
    const int T = blockDim.x, tid = threadIdx.x;
    int rlo, rhi;
    rec_range<REC>(N, rlo, rhi);
    // LDS: PL[2][T][3] and F[2][T][3], alternating halves per chunk.  A lane
    // reads its left neighbour's PL and F at position tid-1 of its own half,
    // lane 0 the previous chunk's last lane in the other half.  Races: the
    // next chunk writes the other half's PL only after this chunk's second
    // barrier (all PL reads done) and its F only after its own first barrier,
    // which no wave reaches before finishing this chunk's F reads.
    double *sPL = lds;
    double *sF = lds + 6 * T;
    int h = 0;
    for (int c0 = 0; c0 <= N; c0 += T, h ^= 1) {
        const int c = c0 + tid;  // stream element: recon of cell c, face c-1, cell c-1
        const bool in = c <= N;
        const int me = h * T + tid;
        const int prev = tid > 0 ? me - 1 : (h ^ 1) * T + T - 1;
        double PLc[3], PRm[3];  // PL[c], PR[c-1]
        if (in && c >= rlo && c < rhi) {
#pragma unroll
            for (int v = 0; v < 3; v++) {
                const double *q = Pin[v] + c;
                double lo, hi;
                if (REC == REC_MINMOD)
                    minmod(q[-1], q[0], q[1], lo, hi);
                else
                    recon<REC>(q[-2], q[-1], q[0], q[1], q[2], lo, hi);
                PLc[v] = hi;
                PRm[v] = lo;
            }
        } else {
#pragma unroll
            for (int v = 0; v < 3; v++) {
                PLc[v] = (in && c < N) ? a.P0[v][o + c] : 0.0;
                PRm[v] = (in && c >= 1) ? a.P0[v][o + c - 1] : 0.0;
            }
        }
#pragma unroll
        for (int v = 0; v < 3; v++) sPL[3 * me + v] = PLc[v];
        __syncthreads();
        double F[3] = {0.0, 0.0, 0.0};
        const bool face = in && c >= 1;  // face c-1 (between cells c-1 and c)
        if (face) {
            double PLm[3];
#pragma unroll
            for (int v = 0; v < 3; v++) PLm[v] = sPL[3 * prev + v];
            face1<RIEM>(PLm, PRm, g, F);
        }
#pragma unroll
        for (int v = 0; v < 3; v++) sF[3 * me + v] = F[v];
        __syncthreads();
        if (face) {
            const int j = c - 1;  // cell updated / recovered by this lane
            double Uo[3];
            const bool upd = j >= 1 && j < N - 2;
#pragma unroll
            for (int v = 0; v < 3; v++) {
                double ub = a.U[v][o + j];
                Uo[v] = upd ? ub - coef * (F[v] - sF[3 * prev + v]) : ub;
                if (s == 0)
                    a.Uh[v][o + j] = Uo[v];
                else if (upd)
                    a.U[v][o + j] = Uo[v];
            }
            double x[3] = {a.P[0][o + j], a.P[1][o + j], a.P[2][o + j]};
            c2p1(Uo, x, g, a.n_iter, a.tol);
#pragma unroll
            for (int v = 0; v < 3; v++) {
                if (s == 0)
                    a.Ph[v][o + j] = x[v];
                else
                    a.P[v][o + j] = x[v];
            }
        }
    }
}

template <int REC, int RIEM>
__global__ void k_tubes(TubeArgs a)
{
    extern __shared__ double lds[];
    __shared__ int cur;
    for (;;) {
        if (threadIdx.x == 0) {
            unsigned int q = atomicAdd(a.queue, 1u);
            cur = q < (unsigned int)a.ntubes ? a.order[q] : -1;
        }
        __syncthreads();
        const int tb = cur;
        __syncthreads();
        if (tb < 0) break;
        for (int st = 0; st < a.nsteps; st++) {
            tube_stage<REC, RIEM>(a, tb, 0, lds);
            __syncthreads();
            tube_stage<REC, RIEM>(a, tb, 1, lds);
            __syncthreads();
        }
    }
}

// U = PtoU(P) on every cell (1dim/structs.jl:58-66); also resets the queue.
__global__ void k_tubes_p2u(TubeArgs a, long ncell)
{
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < ncell; t += (long)gridDim.x * blockDim.x) {
        double P[3] = {a.P[0][t], a.P[1][t], a.P[2][t]}, U[3];
        p2u1(P, U, a.g);
        a.U[0][t] = U[0];
        a.U[1][t] = U[1];
        a.U[2][t] = U[2];
        a.Uh[0][t] = U[0];
        a.Uh[1][t] = U[1];
        a.Uh[2][t] = U[2];
    }
}

// Queue reset as a kernel (graph-safe: no memset node).
__global__ void k_queue_reset(unsigned int *q) { *q = 0u; }

}  // namespace vrn

using namespace vrn;

namespace {
thread_local std::string t_err;
int tfail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    t_err = buf;
    return code;
}
#define TCHK(expr)                                                                        \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess) return tfail(VRN_E_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)
}  // namespace

struct vrn_tubes {
    int ntubes = 0;
    long ncell = 0;
    double gamma = 5.0 / 3.0;
    double *buf = nullptr;
    long *off = nullptr;
    int *order = nullptr;
    double *dt = nullptr, *dx = nullptr;
    unsigned int *queue = nullptr;
    hipStream_t s = nullptr;
    TubeArgs args{};
    int rec = VRN_REC_PPM, riemann = VRN_RIEMANN_HLLC;
    bool params = false;
    // captured graph
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    int g_block = 0, g_grid = 0, g_steps = 0, g_rec = -1, g_riem = -1;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    float last_ms = 0.f;
    double *Psave[3] = {nullptr, nullptr, nullptr}, *Usave[3] = {nullptr, nullptr, nullptr};
};

extern "C" {

const char *vrn_tubes_last_error(void) { return t_err.c_str(); }

int vrn_tubes_create(int ntubes, const int64_t *lengths, double gamma, vrn_tubes **out)
{
    if (!out || !lengths || ntubes <= 0) return tfail(VRN_E_ARG, "bad arguments");
    *out = nullptr;
    long total = 0;
    for (int t = 0; t < ntubes; t++) {
        if (lengths[t] < 8) return tfail(VRN_E_ARG, "tube %d shorter than 8 cells", t);
        total += lengths[t];
    }
    vrn_tubes *c = new vrn_tubes();
    c->ntubes = ntubes;
    c->ncell = total;
    c->gamma = gamma;
    std::vector<long> off(ntubes + 1, 0);
    for (int t = 0; t < ntubes; t++) off[t + 1] = off[t] + lengths[t];
    std::vector<int> order(ntubes);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return lengths[a] > lengths[b]; });
    size_t nd = (size_t)total * 21;  // P, Ph, U, Uh, P0, Psave, Usave x 3 variables
    hipError_t e = hipStreamCreate(&c->s);
    if (e == hipSuccess) e = hipMalloc(&c->buf, nd * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&c->off, sizeof(long) * (ntubes + 1));
    if (e == hipSuccess) e = hipMalloc(&c->order, sizeof(int) * ntubes);
    if (e == hipSuccess) e = hipMalloc(&c->dt, sizeof(double) * ntubes);
    if (e == hipSuccess) e = hipMalloc(&c->dx, sizeof(double) * ntubes);
    if (e == hipSuccess) e = hipMalloc(&c->queue, sizeof(unsigned int));
    if (e == hipSuccess) e = hipMemcpy(c->off, off.data(), sizeof(long) * (ntubes + 1), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(c->order, order.data(), sizeof(int) * ntubes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipEventCreate(&c->ev0);
    if (e == hipSuccess) e = hipEventCreate(&c->ev1);
    if (e != hipSuccess) {
        vrn_tubes_destroy(c);
        return tfail(VRN_E_NOMEM, "allocation failed: %s", hipGetErrorString(e));
    }
    double *p = c->buf;
    TubeArgs &a = c->args;
    for (int v = 0; v < 3; v++) a.P[v] = p + (0 * 3 + v) * total;
    for (int v = 0; v < 3; v++) a.Ph[v] = p + (1 * 3 + v) * total;
    for (int v = 0; v < 3; v++) a.U[v] = p + (2 * 3 + v) * total;
    for (int v = 0; v < 3; v++) a.Uh[v] = p + (3 * 3 + v) * total;
    for (int v = 0; v < 3; v++) a.P0[v] = p + (4 * 3 + v) * total;
    for (int v = 0; v < 3; v++) c->Psave[v] = p + (5 * 3 + v) * total;
    for (int v = 0; v < 3; v++) c->Usave[v] = p + (6 * 3 + v) * total;
    a.off = c->off;
    a.order = c->order;
    a.dt = c->dt;
    a.dx = c->dx;
    a.queue = c->queue;
    a.ntubes = ntubes;
    a.g = gamma;
    *out = c;
    return 0;
}

int vrn_tubes_destroy(vrn_tubes *c)
{
    if (!c) return 0;
    if (c->s) (void)hipStreamSynchronize(c->s);
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    (void)hipFree(c->buf);
    (void)hipFree(c->off);
    (void)hipFree(c->order);
    (void)hipFree(c->dt);
    (void)hipFree(c->dx);
    (void)hipFree(c->queue);
    if (c->s) (void)hipStreamDestroy(c->s);
    delete c;
    return 0;
}

// Upload concatenated SoA primitives (rho, u, u^x; each ncell doubles, tubes
// back to back) with per-tube dt and dx; computes U = PtoU(P) and keeps the
// initial primitives for the face states the reconstruction never writes.
int vrn_tubes_upload(vrn_tubes *c, const double *rho, const double *u, const double *ux,
                     const double *dt, const double *dx)
{
    if (!c || !rho || !u || !ux || !dt || !dx) return tfail(VRN_E_ARG, "null argument");
    const size_t b = sizeof(double) * c->ncell;
    const double *src[3] = {rho, u, ux};
    for (int v = 0; v < 3; v++) {
        TCHK(hipMemcpyAsync(c->args.P[v], src[v], b, hipMemcpyHostToDevice, c->s));
        TCHK(hipMemcpyAsync((void *)c->args.P0[v], src[v], b, hipMemcpyHostToDevice, c->s));
    }
    TCHK(hipMemcpyAsync(c->dt, dt, sizeof(double) * c->ntubes, hipMemcpyHostToDevice, c->s));
    TCHK(hipMemcpyAsync(c->dx, dx, sizeof(double) * c->ntubes, hipMemcpyHostToDevice, c->s));
    hipLaunchKernelGGL(k_tubes_p2u, dim3(1024), dim3(256), 0, c->s, c->args, c->ncell);
    TCHK(hipGetLastError());
    TCHK(hipStreamSynchronize(c->s));
    return 0;
}

int vrn_tubes_download(vrn_tubes *c, double *rho, double *u, double *ux)
{
    if (!c || !rho || !u || !ux) return tfail(VRN_E_ARG, "null argument");
    const size_t b = sizeof(double) * c->ncell;
    double *dst[3] = {rho, u, ux};
    for (int v = 0; v < 3; v++) TCHK(hipMemcpyAsync(dst[v], c->args.P[v], b, hipMemcpyDeviceToHost, c->s));
    TCHK(hipStreamSynchronize(c->s));
    return 0;
}

int vrn_tubes_set_method(vrn_tubes *c, int rec, int riemann, int n_iter, double tol)
{
    if (!c || rec < 0 || rec > 2 || riemann < 0 || riemann > 1 || n_iter < 0)
        return tfail(VRN_E_ARG, "bad method");
    c->rec = rec;
    c->riemann = riemann;
    c->args.n_iter = n_iter;
    c->args.tol = tol;
    c->params = true;
    return 0;
}

static void launch_tubes(vrn_tubes *c, int block, int grid)
{
    size_t lds = sizeof(double) * 12 * block;
    hipLaunchKernelGGL(k_queue_reset, dim3(1), dim3(1), 0, c->s, c->queue);
#define VRN_TL(R_, M_) hipLaunchKernelGGL((k_tubes<R_, M_>), dim3(grid), dim3(block), lds, c->s, c->args)
    if (c->riemann == VRN_RIEMANN_HLL) {
        if (c->rec == VRN_REC_PPM) VRN_TL(REC_PPM, RIEM_HLL);
        else if (c->rec == VRN_REC_WENOZ) VRN_TL(REC_WENOZ, RIEM_HLL);
        else VRN_TL(REC_MINMOD, RIEM_HLL);
    } else {
        if (c->rec == VRN_REC_PPM) VRN_TL(REC_PPM, RIEM_HLLC);
        else if (c->rec == VRN_REC_WENOZ) VRN_TL(REC_WENOZ, RIEM_HLLC);
        else VRN_TL(REC_MINMOD, RIEM_HLLC);
    }
#undef VRN_TL
}

// Advance every tube by nsteps steps with `grid` persistent workgroups of
// `block` threads.  use_graph != 0 captures (queue reset + persistent kernel)
// into a hipGraph once per (block, grid, nsteps, method) and replays it.
// *ms (may be NULL) receives the device time of the launch.  Blocks.
int vrn_tubes_run(vrn_tubes *c, int nsteps, int block, int grid, int use_graph, double *ms)
{
    if (!c || !c->params) return tfail(VRN_E_STATE, "vrn_tubes_set_method first");
    if (nsteps < 0 || block < 64 || block > 1024 || (block & 63) || grid < 1)
        return tfail(VRN_E_ARG, "bad launch shape (block %d, grid %d)", block, grid);
    c->args.nsteps = nsteps;
    if (use_graph) {
        bool stale = !c->gexec || c->g_block != block || c->g_grid != grid || c->g_steps != nsteps ||
                     c->g_rec != c->rec || c->g_riem != c->riemann;
        if (stale) {
            if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
            if (c->graph) (void)hipGraphDestroy(c->graph);
            c->gexec = nullptr;
            c->graph = nullptr;
            TCHK(hipStreamBeginCapture(c->s, hipStreamCaptureModeThreadLocal));
            launch_tubes(c, block, grid);
            TCHK(hipStreamEndCapture(c->s, &c->graph));
            TCHK(hipGraphInstantiate(&c->gexec, c->graph, nullptr, nullptr, 0));
            c->g_block = block;
            c->g_grid = grid;
            c->g_steps = nsteps;
            c->g_rec = c->rec;
            c->g_riem = c->riemann;
        }
        TCHK(hipEventRecord(c->ev0, c->s));
        TCHK(hipGraphLaunch(c->gexec, c->s));
        TCHK(hipEventRecord(c->ev1, c->s));
    } else {
        TCHK(hipEventRecord(c->ev0, c->s));
        launch_tubes(c, block, grid);
        TCHK(hipGetLastError());
        TCHK(hipEventRecord(c->ev1, c->s));
    }
    TCHK(hipStreamSynchronize(c->s));
    float t = 0.f;
    TCHK(hipEventElapsedTime(&t, c->ev0, c->ev1));
    c->last_ms = t;
    if (ms) *ms = t;
    return 0;
}

// Save / restore the state (P, U) so the autotuner can try launch shapes on
// the real data without advancing it.
int vrn_tubes_checkpoint(vrn_tubes *c, int restore)
{
    if (!c) return tfail(VRN_E_ARG, "null");
    const size_t b = sizeof(double) * c->ncell;
    for (int v = 0; v < 3; v++) {
        if (restore) {
            TCHK(hipMemcpyAsync(c->args.P[v], c->Psave[v], b, hipMemcpyDeviceToDevice, c->s));
            TCHK(hipMemcpyAsync(c->args.U[v], c->Usave[v], b, hipMemcpyDeviceToDevice, c->s));
        } else {
            TCHK(hipMemcpyAsync(c->Psave[v], c->args.P[v], b, hipMemcpyDeviceToDevice, c->s));
            TCHK(hipMemcpyAsync(c->Usave[v], c->args.U[v], b, hipMemcpyDeviceToDevice, c->s));
        }
    }
    TCHK(hipStreamSynchronize(c->s));
    return 0;
}

int vrn_tubes_info(vrn_tubes *c, int64_t *ncell, int *cu_count)
{
    if (!c) return tfail(VRN_E_ARG, "null");
    if (ncell) *ncell = c->ncell;
    if (cu_count) {
        int dev = 0;
        TCHK(hipGetDevice(&dev));
        TCHK(hipDeviceGetAttribute(cu_count, hipDeviceAttributeMultiprocessorCount, dev));
    }
    return 0;
}

}  // extern "C"
