// vrn_tubes_kern.hpp -- device code of the batched 1D tubes, compiled once
// per arithmetic form: VRN_NS names the namespace of the instantiation and
// VRN_FAST selects the fast FP64 form (see vrn_physics.hpp); the exact form
// reproduces the oracle bit for bit.
#pragma once
#include "vrn_physics.hpp"
#include "vrn_types.hpp"

namespace VRN_NS {

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
    double vL = fdiv(PL[2], vrn_sqrt1(PL[2] * PL[2] + 1.0));
    double vR = fdiv(PR[2], vrn_sqrt1(PR[2] * PR[2] + 1.0));
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
// dot(col1, cross(col2, col3))); stops when norm(dx) < tol before applying dx
// (fast form: |dx|^2 < tol^2, one reciprocal each of gamma and the determinant).
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
#if defined(VRN_FAST)
        // fast form: one reciprocal of gamma and one of the determinant
        const double ig = recip(gj);
        double a6 = x[0] * x[2] * ig, a7 = -2.0 * x[2] * w, a8 = (2.0 * (x[2] * x[2]) + 1.0) * ig * w;
#else
        double a6 = x[0] * x[2] / gj, a7 = -2.0 * x[2] * w, a8 = (2.0 * (x[2] * x[2]) + 1.0) / gj * w;
#endif
        double c0 = a4 * a8 - a5 * a7;
        double c1 = a5 * a6 - a3 * a8;
        double c2 = a3 * a7 - a4 * a6;
        double d = a0 * c0 + a1 * c1 + a2 * c2;
#if defined(VRN_FAST)
        const double id = recip(d);
        double d0 = ((a4 * a8 - a7 * a5) * f0 + (a6 * a5 - a3 * a8) * f1 + (a3 * a7 - a6 * a4) * f2) * id;
        double d1 = ((a7 * a2 - a1 * a8) * f0 + (a0 * a8 - a6 * a2) * f1 + (a6 * a1 - a0 * a7) * f2) * id;
        double d2 = ((a1 * a5 - a4 * a2) * f0 + (a3 * a2 - a0 * a5) * f1 + (a0 * a4 - a3 * a1) * f2) * id;
        if (d0 * d0 + d1 * d1 + d2 * d2 < tol * tol) break;
#else
        double d0 = ((a4 * a8 - a7 * a5) * f0 + (a6 * a5 - a3 * a8) * f1 + (a3 * a7 - a6 * a4) * f2) / d;
        double d1 = ((a7 * a2 - a1 * a8) * f0 + (a0 * a8 - a6 * a2) * f1 + (a6 * a1 - a0 * a7) * f2) / d;
        double d2 = ((a1 * a5 - a4 * a2) * f0 + (a3 * a2 - a0 * a5) * f1 + (a0 * a4 - a3 * a1) * f2) / d;
        if (vrn_sqrt(d0 * d0 + d1 * d1 + d2 * d2) < tol) break;
#endif
        x[0] = x[0] - d0;
        x[1] = x[1] - d1;
        x[2] = x[2] - d2;
    }
}


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
__device__ void tube_stage(const vrn::TubeArgs &a, int tb, int s, double *lds)
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
            // the stage's outputs are written once per step: non-temporal
            // stores (same-box A/B: +0.9 % exact form, +1-2 % fast form)
            double Uo[3];
            const bool upd = j >= 1 && j < N - 2;
#pragma unroll
            for (int v = 0; v < 3; v++) {
                double ub = a.U[v][o + j];
                Uo[v] = upd ? ub - coef * (F[v] - sF[3 * prev + v]) : ub;
                if (s == 0)
                    __builtin_nontemporal_store(Uo[v], &a.Uh[v][o + j]);
                else if (upd)
                    __builtin_nontemporal_store(Uo[v], &a.U[v][o + j]);
            }
            double x[3] = {a.P[0][o + j], a.P[1][o + j], a.P[2][o + j]};
            c2p1(Uo, x, g, a.n_iter, a.tol);
#pragma unroll
            for (int v = 0; v < 3; v++) {
                if (s == 0)
                    __builtin_nontemporal_store(x[v], &a.Ph[v][o + j]);
                else
                    __builtin_nontemporal_store(x[v], &a.P[v][o + j]);
            }
        }
    }
}

template <int REC, int RIEM>
__global__ void k_tubes(vrn::TubeArgs a)
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

#if !defined(VRN_FAST)
// U = PtoU(P) on every cell (1dim/structs.jl:58-66), Uhalf = U.  Part of the
// upload, so both arithmetic forms start from the exact-form U.
__global__ void k_tubes_p2u(vrn::TubeArgs a, long ncell)
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
#endif

// Queue reset as a kernel (graph-safe: no memset node).
__global__ void k_queue_reset(unsigned int *q) { *q = 0u; }


// Queue reset + the persistent kernel for (rec, riem) on stream s.
inline void launch_tubes_kernels(hipStream_t s, const vrn::TubeArgs &a, int rec, int riem, int block,
                                 int grid)
{
    const size_t lds = sizeof(double) * 12 * block;
    hipLaunchKernelGGL(k_queue_reset, dim3(1), dim3(1), 0, s, a.queue);
#define VRN_TL(R_, M_) hipLaunchKernelGGL((k_tubes<R_, M_>), dim3(grid), dim3(block), lds, s, a)
    if (riem == RIEM_HLL) {
        if (rec == REC_PPM) VRN_TL(REC_PPM, RIEM_HLL);
        else if (rec == REC_WENOZ) VRN_TL(REC_WENOZ, RIEM_HLL);
        else VRN_TL(REC_MINMOD, RIEM_HLL);
    } else {
        if (rec == REC_PPM) VRN_TL(REC_PPM, RIEM_HLLC);
        else if (rec == REC_WENOZ) VRN_TL(REC_WENOZ, RIEM_HLLC);
        else VRN_TL(REC_MINMOD, RIEM_HLLC);
    }
#undef VRN_TL
}

}  // namespace VRN_NS
