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
//
// This file holds the C ABI and the exact-form kernels; the kernels
// themselves are in vrn_tubes_kern.hpp, compiled again as the fast FP64 form
// in vrn_tubes_fast.hip (vrn_tubes_set_precision selects one).
#include "../../include/verona_hip.h"
#include "vrn_tubes_kern.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

// fast FP64 form of the tube kernels (vrn_tubes_fast.hip)
void vrn_fast_launch_tubes(hipStream_t s, const vrn::TubeArgs &a, int rec, int riem, int block, int grid);

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
    int precision = VRN_PREC_FP64;
    bool params = false;
    // captured graph
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    int g_block = 0, g_grid = 0, g_steps = 0, g_rec = -1, g_riem = -1, g_prec = -1;
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

// Arithmetic of the tube kernels: VRN_PREC_FP64 (default) reproduces the
// oracle bit for bit; VRN_PREC_FP64_FAST is the same algorithm with FMA
// contraction and reciprocal-based division (vrn_tubes_fast.hip).
int vrn_tubes_set_precision(vrn_tubes *c, int precision)
{
    if (!c) return tfail(VRN_E_ARG, "null");
    if (precision != VRN_PREC_FP64 && precision != VRN_PREC_FP64_FAST)
        return tfail(VRN_E_ARG, "tubes support VRN_PREC_FP64 and VRN_PREC_FP64_FAST (got %d)", precision);
    c->precision = precision;
    return 0;
}

static void launch_tubes(vrn_tubes *c, int block, int grid)
{
    // VRN_REC_* / VRN_RIEMANN_* equal the kernels' REC_* / RIEM_* values
    if (c->precision == VRN_PREC_FP64_FAST)
        vrn_fast_launch_tubes(c->s, c->args, c->rec, c->riemann, block, grid);
    else
        vrn::launch_tubes_kernels(c->s, c->args, c->rec, c->riemann, block, grid);
}

// Advance every tube by nsteps steps with `grid` persistent workgroups of
// `block` threads.  use_graph != 0 captures (queue reset + persistent kernel)
// into a hipGraph once per (block, grid, nsteps, method, precision) and replays it.
// *ms (may be NULL) receives the device time of the launch.  Blocks.
int vrn_tubes_run(vrn_tubes *c, int nsteps, int block, int grid, int use_graph, double *ms)
{
    if (!c || !c->params) return tfail(VRN_E_STATE, "vrn_tubes_set_method first");
    if (nsteps < 0 || block < 64 || block > 1024 || (block & 63) || grid < 1)
        return tfail(VRN_E_ARG, "bad launch shape (block %d, grid %d)", block, grid);
    c->args.nsteps = nsteps;
    if (use_graph) {
        bool stale = !c->gexec || c->g_block != block || c->g_grid != grid || c->g_steps != nsteps ||
                     c->g_rec != c->rec || c->g_riem != c->riemann || c->g_prec != c->precision;
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
            c->g_prec = c->precision;
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
