/*
 * orc_3d.c -- CPU oracle, 3D solver (TEST INFRASTRUCTURE ONLY; see orc_core.h).
 *
 * Arrays use Verona's AoS byte order (3dim/structs.jl:25-37): arr[v, i, j, k]
 * column-major with the variable fastest, 3 ghost cells per side, so the flat
 * 0-based index is v + 5*(i + NX*(j + NY*k)) with NX = Nx + 6 etc.
 */
#include "verona_oracle.h"
#include "orc_core.h"

#include <stdlib.h>
#include <string.h>

#define IX(v, i, j, k) ((size_t)(v) + 5 * ((size_t)(i) + (size_t)NX * ((size_t)(j) + (size_t)NY * (size_t)(k))))

/* 3dim/structs.jl:66-81  kernel_PtoU over the whole array (ghosts included). */
void orc3_ptou(const double *P, double *U, int NX, int NY, int NZ, double g)
{
    size_t n = (size_t)NX * NY * NZ;
#pragma omp parallel for schedule(static)
    for (size_t c = 0; c < n; c++) orc3_p2u(P + 5 * c, U + 5 * c, g);
}

/* 3dim/algos.jl:5-12  function_Limit over the whole array. */
void orc3_limit(double *P, double floor_, int NX, int NY, int NZ)
{
    size_t n = (size_t)NX * NY * NZ;
#pragma omp parallel for schedule(static)
    for (size_t c = 0; c < n; c++) {
        P[5 * c + 0] = jl_max(P[5 * c + 0], floor_);
        P[5 * c + 1] = jl_max(P[5 * c + 1], floor_);
    }
}

/* Face states + Riemann flux at the face between cell c and c + e_dim.
 * 3dim/algos.jl:57-174: PL = upper WENOZ face of cell c, PR = lower WENOZ
 * face of cell c+e_dim, floors on rho and u, PtoU/PtoF, wave speeds, HLL
 * (riemann == 0) or HLLC (riemann == 1, no reference counterpart). */
static void face_flux3(const double *P, int NX, int NY, int i, int j, int k, int dim, double g,
                       double floor_, int rec, int riemann, double *F)
{
    int di = dim == 0, dj = dim == 1, dk = dim == 2;
    double PL[5], PR[5], UL[5], UR[5], FL[5], FR[5];
    for (int v = 0; v < 5; v++) {
        double q[6];
        for (int s = 0; s < 6; s++) {
            int o = s - 2;
            q[s] = P[IX(v, i + o * di, j + o * dj, k + o * dk)];
        }
        double lo, hi;
        orc_recon(rec, q[0], q[1], q[2], q[3], q[4], &lo, &hi);
        PL[v] = hi;
        orc_recon(rec, q[1], q[2], q[3], q[4], q[5], &lo, &hi);
        PR[v] = lo;
    }
    for (int v = 0; v < 2; v++) {
        PL[v] = jl_max(floor_, PL[v]);
        PR[v] = jl_max(floor_, PR[v]);
    }
    orc3_p2u(PR, UR, g);
    orc3_p2u(PL, UL, g);
    orc3_p2f(PR, FR, g, dim);
    orc3_p2f(PL, FL, g, dim);
    double lorL = sqrt(PL[2] * PL[2] + PL[3] * PL[3] + PL[4] * PL[4] + 1.0);
    double lorR = sqrt(PR[2] * PR[2] + PR[3] * PR[3] + PR[4] * PR[4] + 1.0);
    double vL = PL[2 + dim] / lorL;
    double vR = PR[2 + dim] / lorR;
    double CL = orc_sound_speed(PL[0], PL[1], g);
    double CR = orc_sound_speed(PR[0], PR[1], g);
    double lpL, lmL, lpR, lmR;
    orc_lambdas(vL, CL, lorL * lorL, &lpL, &lmL);
    orc_lambdas(vR, CR, lorR * lorR, &lpR, &lmR);
    double Cmax = jl_max(lpL, lpR);
    double Cmin = -jl_min(lmL, lmR);
    if (riemann == 0)
        orc_hll(5, UL, UR, FL, FR, Cmax, Cmin, F);
    else
        orc_hllc(5, 2 + dim, UL, UR, FL, FR, vL, vR, (g - 1.0) * PL[1], (g - 1.0) * PR[1], Cmax,
                 Cmin, F);
}

/* 3dim/algos.jl:14-176  function_Fluxes: writes F[:, i, j, k] (face i+1/2
 * along dim) for 0-based i, j, k in [2, N-3) in every axis; other entries of F
 * are left untouched. */
void orc3_fluxes(const double *P, double g, double floor_, double *F, int dim, int NX, int NY,
                 int NZ, int rec, int riemann)
{
#pragma omp parallel for schedule(static)
    for (int k = 2; k < NZ - 3; k++)
        for (int j = 2; j < NY - 3; j++)
            for (int i = 2; i < NX - 3; i++)
                face_flux3(P, NX, NY, i, j, k, dim, g, floor_, rec, riemann, F + IX(0, i, j, k));
}

/* 3dim/algos.jl:178-200  function_Update on interior cells (0-based [3, N-3)):
 * Ub = U - dt/dx (Fx_i - Fx_{i-1}) - dt/dy (...) - dt/dz (...).  Ub may alias U. */
void orc3_update(const double *U, double *Ub, double dt, double dx, double dy, double dz,
                 const double *Fx, const double *Fy, const double *Fz, int NX, int NY, int NZ)
{
#pragma omp parallel for schedule(static)
    for (int k = 3; k < NZ - 3; k++)
        for (int j = 3; j < NY - 3; j++)
            for (int i = 3; i < NX - 3; i++)
                for (int v = 0; v < 5; v++) {
                    size_t c = IX(v, i, j, k);
                    Ub[c] = U[c] - dt / dx * (Fx[c] - Fx[IX(v, i - 1, j, k)]) -
                            dt / dy * (Fy[c] - Fy[IX(v, i, j - 1, k)]) -
                            dt / dz * (Fz[c] - Fz[IX(v, i, j, k - 1)]);
                }
}

/* ------------------------------------------------------------------------ */
/* con2prim cascade, 3dim/structs.jl:218-628 (one cell).                     */
/* ------------------------------------------------------------------------ */

/* 3dim/structs.jl:144-179  LU_dec_5D! (no pivoting), column-major flat 5x5. */
static void lu5(double *A, const double *t, double *x)
{
#define A_(i, j) A[((j) - 1) * 5 + (i) - 1]
    for (int k = 1; k <= 5; k++)
        for (int i = k + 1; i <= 5; i++) {
            A_(i, k) /= A_(k, k);
            for (int j = k + 1; j <= 5; j++) A_(i, j) -= A_(i, k) * A_(k, j);
        }
    for (int i = 1; i <= 5; i++) {
        x[i - 1] = t[i - 1];
        for (int j = 1; j <= i - 1; j++) x[i - 1] -= A_(i, j) * x[j - 1];
    }
    for (int i = 5; i >= 1; i--) {
        for (int j = i + 1; j <= 5; j++) x[i - 1] -= A_(i, j) * x[j - 1];
        x[i - 1] /= A_(i, i);
    }
#undef A_
}

/* 3dim/structs.jl:181-214  LU_dec_2D! */
static void lu2(double *A, const double *t, double *x)
{
    A[1] /= A[0];
    A[3] -= A[1] * A[2];
    x[0] = t[0];
    x[1] = t[1];
    x[1] -= A[1] * x[0];
    x[1] /= A[3];
    x[0] -= A[2] * x[1];
    x[0] /= A[0];
}

/* The 1DW residual f(W) of 3dim/structs.jl:315-317 with v^2 given. */
static inline double res_1dw(double Qt, double W, double v2, double D, double c)
{
    return Qt + W - c * (W * (1.0 - v2) - D * sqrt(1.0 - v2));
}

/* Its derivative f'(W) with v^2 = S^2/W^2, 3dim/structs.jl:318-321. */
static inline double jac_1dw(double W, double v2, double S2, double D, double c)
{
    double W3 = W * W * W;
    if (ORC_MUT == ORC_MUT_F1DW) return 1.0 - c * (1.0 - v2) - c * (W * (2.0 * S2 / W3));
    return 1.0 - c * (1.0 - v2) - c * (W * (2.0 * S2 / W3) - D * (S2 / (W3 * sqrt(1.0 - v2))));
}

/* f(W) and f'(W) of the 1DW iteration for the conserved state Uc at W,
 * with v^2 = S^2/W^2 clamped to [0, 1 - 1e-15] as the iteration does;
 * exported for the pins. */
void orc3_f1dw(const double *Uc, double W, double g, double *f, double *fp)
{
    const double c = (g - 1.0) / g;
    double S2 = Uc[2] * Uc[2] + Uc[3] * Uc[3] + Uc[4] * Uc[4];
    double v2 = S2 / (W * W);
    if (v2 < 0.0)
        v2 = 0.0;
    else if (v2 > 1.0 - 1e-15)
        v2 = 1.0 - 1e-15;
    *f = res_1dw(Uc[1], W, v2, Uc[0], c);
    *fp = jac_1dw(W, v2, S2, Uc[0], c);
}

/* Bisection residual, 3dim/structs.jl:546-550. */
static inline double res_bis(double Qt, double W, double S2, double D, double c)
{
    return Qt + W - c * (W * (1.0 - S2 / (W * W)) - D * sqrt(jl_max(0.0, 1.0 - S2 / (W * W))));
}

int orc3_c2p_cell(const double *Uc, double *Pc, double g, int n_iter, double tol, int *iters)
{
    const double D = Uc[0], Q_t = Uc[1], Q_x = Uc[2], Q_y = Uc[3], Q_z = Uc[4];
    const double c = (g - 1.0) / g;
    int conv1 = 0, conv2 = 0, conv5 = 0, convb = 0;
    int it_total = 0;
    double W = 0.0, v2, S2;

    /* 1DW (Noble et al. 2006), 3dim/structs.jl:274-343 */
    {
        S2 = Q_x * Q_x + Q_y * Q_y + Q_z * Q_z;
        double gm = sqrt(Pc[2] * Pc[2] + Pc[3] * Pc[3] + Pc[4] * Pc[4] + 1.0);
        double w_small = Pc[0] + g * Pc[1];
        W = w_small * (gm * gm);
        const double W_max = 1e30;
        double W_min = sqrt(S2) * (1.0 + 1e-15);
        for (int t10 = 0; t10 < 1100 && S2 / (W * W) >= 1.0 && W < W_max; t10++) W *= 10.0; /* unbounded in the reference; W = 0 would never exit */
        v2 = jl_min(S2 / (W * W), 1.0 - 1e-15);
        for (int it = 0; it < n_iter; it++) {
            it_total++;
            if (W < W_min) W = W_min;
            v2 = S2 / (W * W);
            if (v2 < 0.0)
                v2 = 0.0;
            else if (v2 > 1.0 - 1e-15)
                v2 = 1.0 - 1e-15;
            double fun = res_1dw(Q_t, W, v2, D, c);
            double jac = jac_1dw(W, v2, S2, D, c);
            double dW = fun / jac;
            double Wp = W - dW;
            if (Wp < W_min)
                W = 0.5 * (W + W_min);
            else
                W = Wp;
            if (dW < 0.0) dW = -dW;
            if (dW * dW < tol * tol) {
                conv1 = 1;
                break;
            }
        }
    }

    /* 2D Newton with backtracking, 3dim/structs.jl:346-447 */
    if (!conv1) {
        S2 = Q_x * Q_x + Q_y * Q_y + Q_z * Q_z;
        double gm = sqrt(Pc[2] * Pc[2] + Pc[3] * Pc[3] + Pc[4] * Pc[4] + 1.0);
        double w_small = Pc[0] + g * Pc[1];
        W = w_small * (gm * gm);
        const double W_max = 1e30;
        for (int t10 = 0; t10 < 1100 && S2 / (W * W) >= 1.0 && W < W_max; t10++) W *= 10.0; /* unbounded in the reference; W = 0 would never exit */
        v2 = jl_min(S2 / (W * W), 1.0 - 1e-15);
        for (int it = 0; it < n_iter; it++) {
            it_total++;
            double W_old = W, v2_old = v2;
            double f[2], J[4], o[2];
            f[0] = v2 - S2 / (W * W);
            f[1] = res_1dw(Q_t, W, v2, D, c);
            J[0] = 2.0 * S2 / (W * W * W);
            J[1] = 1.0 - c * (1.0 - v2);
            J[2] = 1.0;
            J[3] = c * (W - D / (2.0 * sqrt(1.0 - v2)));
            lu2(J, f, o);
            double dW = o[0], dv2 = o[1], alpha = 1.0;
            int converged = 0;
            for (int b = 0; b < 10; b++) {
                double sq = sqrt(W - alpha * dW);
                double Wc = sq * sq;
                double v2c = v2 - alpha * dv2;
                if (Wc <= 0.0 || Wc > W_max) Wc = W_old;
                if (v2c < 0.0)
                    v2c = 0.0;
                else if (v2c >= 1.0)
                    v2c = 1.0 - 1e-15;
                double r1 = v2c - S2 / (Wc * Wc);
                double r2 = res_1dw(Q_t, Wc, v2c, D, c);
                double errc = r1 * r1 + r2 * r2;
                double r1o = v2 - S2 / (W * W);
                double r2o = res_1dw(Q_t, W, v2, D, c);
                double erro = r1o * r1o + r2o * r2o;
                if (errc <= erro) {
                    W = Wc;
                    v2 = v2c;
                    converged = 1;
                    break;
                } else {
                    alpha *= 0.5;
                }
            }
            if (!converged) {
                W = W_old;
                v2 = v2_old;
            }
            double relW = sqrt((W - W_old) * (W - W_old)) / jl_max(fabs(W_old), tol * tol);
            double relv = sqrt((v2 - v2_old) * (v2 - v2_old)) / jl_max(fabs(v2_old), tol * tol);
            if (relW + relv < tol) {
                conv2 = 1;
                break;
            }
        }
    }

    /* 5D Newton on primitives, 3dim/structs.jl:450-539 (modifies Pc in place) */
    if (!conv1 && !conv2) {
        for (int it = 0; it < n_iter; it++) {
            it_total++;
            double gam = sqrt(Pc[2] * Pc[2] + Pc[3] * Pc[3] + Pc[4] * Pc[4] + 1.0);
            double w = g * Pc[1] + Pc[0];
            double f[5], J[25], o[5];
            f[0] = gam * Pc[0] - Uc[0];
            f[1] = (g - 1.0) * Pc[1] - gam * gam * w - Uc[1];
            f[2] = Pc[2] * gam * w - Uc[2];
            f[3] = Pc[3] * gam * w - Uc[3];
            f[4] = Pc[4] * gam * w - Uc[4];
            J[0] = gam;
            J[5] = 0.0;
            J[10] = Pc[2] * Pc[0] / gam;
            J[15] = Pc[3] * Pc[0] / gam;
            J[20] = Pc[4] * Pc[0] / gam;
            J[1] = -(gam * gam);
            J[6] = g * (-(gam * gam)) + g - 1.0;
            J[11] = -2.0 * Pc[2] * w;
            J[16] = -2.0 * Pc[3] * w;
            J[21] = -2.0 * Pc[4] * w;
            J[2] = Pc[2] * gam;
            J[7] = g * Pc[2] * gam;
            J[12] = Pc[2] * Pc[2] * w / gam + w * gam;
            J[17] = Pc[2] * Pc[3] * w / gam;
            J[22] = Pc[2] * Pc[4] * w / gam;
            J[3] = Pc[3] * gam;
            J[8] = g * Pc[3] * gam;
            J[13] = Pc[2] * Pc[3] * w / gam;
            J[18] = Pc[3] * Pc[3] * w / gam + w * gam;
            J[23] = Pc[3] * Pc[4] * w / gam;
            J[4] = Pc[4] * gam;
            J[9] = g * Pc[4] * gam;
            J[14] = Pc[2] * Pc[4] * w / gam;
            J[19] = Pc[3] * Pc[4] * w / gam;
            J[24] = Pc[4] * Pc[4] * w / gam + w * gam;
            lu5(J, f, o);
            if (o[0] * o[0] + o[1] * o[1] + o[2] * o[2] + o[3] * o[3] + o[4] * o[4] < tol * tol) {
                if (isfinite(Pc[0]) && isfinite(Pc[1]) && isfinite(Pc[2]) && isfinite(Pc[3]) &&
                    isfinite(Pc[4])) {
                    conv5 = 1;
                    Pc[0] = jl_max(1e-8, Pc[0]);
                    Pc[1] = jl_max(1e-8, Pc[1]);
                    double gm = sqrt(Pc[2] * Pc[2] + Pc[3] * Pc[3] + Pc[4] * Pc[4] + 1.0);
                    Pc[2] = Pc[2] / gm;
                    Pc[3] = Pc[3] / gm;
                    Pc[4] = Pc[4] / gm;
                    gm = jl_min(gm, 50.0);
                    Pc[2] = Pc[2] * gm;
                    Pc[3] = Pc[3] * gm;
                    Pc[4] = Pc[4] * gm;
                }
                break;
            }
            for (int v = 0; v < 5; v++) Pc[v] = Pc[v] - o[v];
        }
    }

    /* bisection on W, 3dim/structs.jl:542-596 */
    if (!conv1 && !conv2 && !conv5) {
        S2 = Q_x * Q_x + Q_y * Q_y + Q_z * Q_z;
        double W_min = jl_max(1e-15, sqrt(S2));
        double W_max = W_min * 10.0;
        double fmin = res_bis(Q_t, W_min, S2, D, c);
        double fmax = res_bis(Q_t, W_max, S2, D, c);
        int count = 0;
        while (fmin * fmax > 0.0 && count < 100) {
            W_max *= 10.0;
            fmax = res_bis(Q_t, W_max, S2, D, c);
            count++;
        }
        if (fmin * fmax <= 0.0) {
            const double thr = 1e-6 * 1e-6;
            for (int it = 0; it < n_iter * 10; it++) {
                it_total++;
                double W_mid = 0.5 * (W_min + W_max);
                double fmid = res_bis(Q_t, W_mid, S2, D, c);
                if (fmid * fmid < thr) {
                    convb = 1;
                    W = W_mid;
                    break;
                }
                if (fmin * fmid < 0.0) {
                    W_max = W_mid;
                    fmax = fmid;
                } else {
                    W_min = W_mid;
                    fmin = fmid;
                }
                W = W_mid;
            }
        }
    }

    /* recovery, 3dim/structs.jl:598-609 */
    if (conv1 || conv2 || convb) {
        v2 = S2 / (W * W);
        double gm = 1.0 / sqrt(1.0 - v2);
        gm = jl_min(gm, 50.0);
        double rho = jl_max(1e-8, D / gm);
        double UU = jl_max(1e-8, (W / (gm * gm) - D / gm) / g);
        Pc[0] = rho;
        Pc[1] = UU;
        Pc[2] = gm * Q_x / W;
        Pc[3] = gm * Q_y / W;
        Pc[4] = gm * Q_z / W;
    }
    /* failure marker, 3dim/structs.jl:613-619 */
    if (!conv1 && !conv2 && !conv5 && !convb) {
        for (int v = 0; v < 5; v++) Pc[v] = NAN;
    }
    if (iters) *iters = it_total;
    if (conv1) return 1;
    if (conv2) return 2;
    if (conv5) return 3;
    if (convb) return 4;
    return 0;
}

/* 3dim/structs.jl:218-628  function_UtoP on interior cells; P holds the
 * warm-start guess on entry.  method_counts[0..4] receives the number of
 * cells finished by: failure, 1DW, 2D, 5D, bisection.  Returns the number of
 * unrecovered cells: those whose result is not finite (the NaN marker, or a
 * NaN from a branch that reported convergence). */
long orc3_c2p(const double *U, double *P, int NX, int NY, int NZ, double g, int n_iter,
              double tol, long *method_counts)
{
    long cnt[5] = {0, 0, 0, 0, 0};
    long bad = 0;
#pragma omp parallel
    {
        long lc[5] = {0, 0, 0, 0, 0};
        long bad_l = 0;
#pragma omp for schedule(static)
        for (int k = 3; k < NZ - 3; k++)
            for (int j = 3; j < NY - 3; j++)
                for (int i = 3; i < NX - 3; i++) {
                    size_t c = IX(0, i, j, k);
                    int m = orc3_c2p_cell(U + c, P + c, g, n_iter, tol, NULL);
                    lc[m]++;
                    if (!(isfinite(P[c]) && isfinite(P[c + 1]) && isfinite(P[c + 2]) &&
                          isfinite(P[c + 3]) && isfinite(P[c + 4])))
                        bad_l++;
                }
#pragma omp critical
        {
            for (int m = 0; m < 5; m++) cnt[m] += lc[m];
            bad += bad_l;
        }
    }
    if (method_counts)
        for (int m = 0; m < 5; m++) method_counts[m] = cnt[m];
    return bad;
}

/* Batched con2prim over n independent cells (U, P are [n][5]); used for the
 * con2prim workload.  iters (optional) receives the per-cell iteration count. */
long orc3_c2p_batch(const double *U, double *P, long n, double g, int n_iter, double tol,
                    int *method, int *iters)
{
    long bad = 0;
#pragma omp parallel for schedule(dynamic, 4096) reduction(+ : bad)
    for (long c = 0; c < n; c++) {
        int it = 0;
        int m = orc3_c2p_cell(U + 5 * c, P + 5 * c, g, n_iter, tol, &it);
        if (method) method[c] = m;
        if (iters) iters[c] = it;
        const double *p = P + 5 * c;
        bad += !(isfinite(p[0]) && isfinite(p[1]) && isfinite(p[2]) && isfinite(p[3]) && isfinite(p[4]));
        (void)m;
    }
    return bad;
}

/* ------------------------------------------------------------------------ */
/* Ghost fills.                                                               */
/* ------------------------------------------------------------------------ */

/* Periodic wrap of the 3-cell slabs along the axes in `mask` (bit d = axis d),
 * axis by axis (x, then y, then z), each copying whole planes.  This is the
 * single-rank form of 3dim/sync_CUDA.jl:3-129 with a periodic communicator. */
void orc3_fill_periodic(double *P, int NX, int NY, int NZ, int mask)
{
    if (mask & 1)
        for (int k = 0; k < NZ; k++)
            for (int j = 0; j < NY; j++)
                for (int g3 = 0; g3 < 3; g3++)
                    for (int v = 0; v < 5; v++) {
                        P[IX(v, g3, j, k)] = P[IX(v, NX - 6 + g3, j, k)];
                        P[IX(v, NX - 3 + g3, j, k)] = P[IX(v, 3 + g3, j, k)];
                    }
    if (mask & 2)
        for (int k = 0; k < NZ; k++)
            for (int g3 = 0; g3 < 3; g3++)
                for (int i = 0; i < NX; i++)
                    for (int v = 0; v < 5; v++) {
                        P[IX(v, i, g3, k)] = P[IX(v, i, NY - 6 + g3, k)];
                        P[IX(v, i, NY - 3 + g3, k)] = P[IX(v, i, 3 + g3, k)];
                    }
    if (mask & 4)
        for (int g3 = 0; g3 < 3; g3++)
            for (int j = 0; j < NY; j++)
                for (int i = 0; i < NX; i++)
                    for (int v = 0; v < 5; v++) {
                        P[IX(v, i, j, g3)] = P[IX(v, i, j, NZ - 6 + g3)];
                        P[IX(v, i, j, NZ - 3 + g3)] = P[IX(v, i, j, 3 + g3)];
                    }
}

/* Zero-gradient (outflow) copy of the first/last interior plane into the 3
 * ghost planes, for faces flagged in lo_mask / hi_mask (bit d = axis d), axis
 * by axis x, y, z.  Same copies as the TurnOff callback of
 * examples/Jet3D/HARM_jet_breakout.jl:136-157. */
void orc3_fill_outflow(double *P, int NX, int NY, int NZ, int lo_mask, int hi_mask)
{
    for (int k = 0; k < NZ; k++)
        for (int j = 0; j < NY; j++)
            for (int g3 = 0; g3 < 3; g3++)
                for (int v = 0; v < 5; v++) {
                    if (lo_mask & 1) P[IX(v, g3, j, k)] = P[IX(v, 3, j, k)];
                    if (hi_mask & 1) P[IX(v, NX - 3 + g3, j, k)] = P[IX(v, NX - 4, j, k)];
                }
    for (int k = 0; k < NZ; k++)
        for (int g3 = 0; g3 < 3; g3++)
            for (int i = 0; i < NX; i++)
                for (int v = 0; v < 5; v++) {
                    if (lo_mask & 2) P[IX(v, i, g3, k)] = P[IX(v, i, 3, k)];
                    if (hi_mask & 2) P[IX(v, i, NY - 3 + g3, k)] = P[IX(v, i, NY - 4, k)];
                }
    for (int g3 = 0; g3 < 3; g3++)
        for (int j = 0; j < NY; j++)
            for (int i = 0; i < NX; i++)
                for (int v = 0; v < 5; v++) {
                    if (lo_mask & 4) P[IX(v, i, j, g3)] = P[IX(v, i, j, 3)];
                    if (hi_mask & 4) P[IX(v, i, j, NZ - 3 + g3)] = P[IX(v, i, j, NZ - 4)];
                }
}

/* Jet nozzle on the z-low face: ghost cells k = 0..2 whose cell centre lies
 * within `radius` of the (x, y) box centre hold the injected state
 * (rho, u, u^i = |u| * n) with n = normalize(e_z + eps (cos phi, sin phi, 0)).
 * Applied after the outflow copy.  No reference counterpart (the reference's
 * jet is an initial condition plus a user callback). */
void orc3_fill_nozzle(double *P, int NX, int NY, int NZ, double dx, double dy, double radius,
                      double rho, double u, double umag, double eps, double phi)
{
    (void)NZ;
    double nx = eps * cos(phi), ny = eps * sin(phi), nz = 1.0;
    double nn = sqrt(nx * nx + ny * ny + nz * nz);
    nx /= nn;
    ny /= nn;
    nz /= nn;
    double cx = 0.5 * (NX - 6) * dx, cy = 0.5 * (NY - 6) * dy;
    for (int g3 = 0; g3 < 3; g3++)
        for (int j = 3; j < NY - 3; j++)
            for (int i = 3; i < NX - 3; i++) {
                double x = ((i - 3) + 0.5) * dx - cx;
                double y = ((j - 3) + 0.5) * dy - cy;
                if (x * x + y * y < radius * radius) {
                    P[IX(0, i, j, g3)] = rho;
                    P[IX(1, i, j, g3)] = u;
                    P[IX(2, i, j, g3)] = umag * nx;
                    P[IX(3, i, j, g3)] = umag * ny;
                    P[IX(4, i, j, g3)] = umag * nz;
                }
            }
}

/* ------------------------------------------------------------------------ */
/* Drivers.                                                                   */
/* ------------------------------------------------------------------------ */

/* Apply the boundary spec bc[6] = {xlo, xhi, ylo, yhi, zlo, zhi}. */
static void apply_bc3(double *P, int NX, int NY, int NZ, const orc_bc3 *bc, int nonperiodic)
{
    int pmask = 0, lo = 0, hi = 0;
    for (int d = 0; d < 3; d++) {
        if (bc->type[2 * d] == ORC_BC_PERIODIC) pmask |= 1 << d;
        if (bc->type[2 * d] == ORC_BC_OUTFLOW || bc->type[2 * d] == ORC_BC_NOZZLE) lo |= 1 << d;
        if (bc->type[2 * d + 1] == ORC_BC_OUTFLOW) hi |= 1 << d;
    }
    if (nonperiodic) {
        if (lo | hi) orc3_fill_outflow(P, NX, NY, NZ, lo, hi);
        if (bc->type[4] == ORC_BC_NOZZLE)
            orc3_fill_nozzle(P, NX, NY, NZ, bc->dx, bc->dy, bc->noz_radius, bc->noz_rho,
                             bc->noz_u, bc->noz_umag, bc->noz_eps, bc->noz_phi);
    } else if (pmask) {
        orc3_fill_periodic(P, NX, NY, NZ, pmask);
    }
}

/* Cell-centred CFL: max over interior cells of sum_d max(|lam+_d|, |lam-_d|)/dx_d
 * (wave speeds as in 3dim/algos.jl:130-157, evaluated at the cell state). */
double orc3_max_rate(const double *P, int NX, int NY, int NZ, double g, double dx, double dy,
                     double dz)
{
    double mx = 0.0;
    double inv[3] = {1.0 / dx, 1.0 / dy, 1.0 / dz};
#pragma omp parallel for schedule(static) reduction(max : mx)
    for (int k = 3; k < NZ - 3; k++)
        for (int j = 3; j < NY - 3; j++)
            for (int i = 3; i < NX - 3; i++) {
                const double *p = P + IX(0, i, j, k);
                double lor = sqrt(p[2] * p[2] + p[3] * p[3] + p[4] * p[4] + 1.0);
                double cs = orc_sound_speed(p[0], p[1], g);
                double r = 0.0;
                for (int d = 0; d < 3; d++) {
                    double lp, lm;
                    orc_lambdas(p[2 + d] / lor, cs, lor * lor, &lp, &lm);
                    double a = jl_max(fabs(lp), fabs(lm));
                    r = r + a * inv[d];
                }
                if (r > mx) mx = r;
            }
    return mx;
}

static double cell_rate3(const double *p, double g, const double *inv)
{
    double lor = sqrt(p[2] * p[2] + p[3] * p[3] + p[4] * p[4] + 1.0);
    double cs = orc_sound_speed(p[0], p[1], g);
    double r = 0.0;
    for (int d = 0; d < 3; d++) {
        double lp, lm;
        orc_lambdas(p[2 + d] / lor, cs, lor * lor, &lp, &lm);
        r = r + jl_max(fabs(lp), fabs(lm)) * inv[d];
    }
    return r;
}

/* Signal rate of the nozzle's inflow state at the current angle. */
static double nozzle_rate(const orc_bc3 *bc, double g, double dx, double dy, double dz)
{
    double nx = bc->noz_eps * cos(bc->noz_phi), ny = bc->noz_eps * sin(bc->noz_phi), nz = 1.0;
    double nn = sqrt(nx * nx + ny * ny + nz * nz);
    double p[5] = {bc->noz_rho, bc->noz_u, bc->noz_umag * (nx / nn), bc->noz_umag * (ny / nn),
                   bc->noz_umag * (nz / nn)};
    double inv[3] = {1.0 / dx, 1.0 / dy, 1.0 / dz};
    return cell_rate3(p, g, inv);
}

/* One stage of the fused form: Uout = a*U0 + b*acc with
 * acc = ((Ubase - dt/dx dFx) - dt/dy dFy) - dt/dz dFz  (a == 0: Uout = acc). */
static void stage_update3(const double *U0, const double *Ub, double *Uout, double a, double b,
                          double dt, double dx, double dy, double dz, const double *Fx,
                          const double *Fy, const double *Fz, int NX, int NY, int NZ)
{
#pragma omp parallel for schedule(static)
    for (int k = 3; k < NZ - 3; k++)
        for (int j = 3; j < NY - 3; j++)
            for (int i = 3; i < NX - 3; i++)
                for (int v = 0; v < 5; v++) {
                    size_t c = IX(v, i, j, k);
                    double acc = Ub[c] - dt / dx * (Fx[c] - Fx[IX(v, i - 1, j, k)]);
                    acc = acc - dt / dy * (Fy[c] - Fy[IX(v, i, j - 1, k)]);
                    acc = acc - dt / dz * (Fz[c] - Fz[IX(v, i, j, k - 1)]);
                    Uout[c] = a == 0.0 ? acc : a * U0[c] + b * acc;
                }
}

/*
 * Run nsteps steps of the 3D solver on P (in/out, AoS with ghosts).
 * mode->ref != 0 reproduces Verona3D.HARM_HLL (3dim/algos.jl:203-362):
 *   ghost exchange + PtoU at start; per step: outflow/nozzle callback at step
 *   start only, Fluxes x,y,z, Update(U -> Uhalf, dt/2), UtoP, Limit, exchange,
 *   Fluxes, Update(U -> U, dt), UtoP (guess P*), Limit, exchange, t += dt.
 * mode->ref == 0 is the extended form: boundaries after every stage, midpoint
 *   RK2 (integrator 0) or SSP-RK3 (integrator 1), fixed dt or CFL dt.
 * Returns the number of unrecovered cells summed over all stages;
 * *t_out receives the final time.
 */
long orc3_run(double *P, int NX, int NY, int NZ, const orc_mode *mode, const orc_bc3 *bc,
              int nsteps, double *t_out, double *dt_last)
{
    size_t n = (size_t)NX * NY * NZ * 5;
    double *U = calloc(n, sizeof(double));
    double *Us = calloc(n, sizeof(double));
    double *Fx = calloc(n, sizeof(double));
    double *Fy = calloc(n, sizeof(double));
    double *Fz = calloc(n, sizeof(double));
    double g = mode->gamma, fl = mode->floor_;
    double dx = mode->dx, dy = mode->dy, dz = mode->dz;
    long bad = 0;
    double t = 0.0, dt = mode->dt;
    orc_bc3 bcs = *bc;

    /* start: exchange (periodic) + boundaries, then PtoU */
    apply_bc3(P, NX, NY, NZ, &bcs, 0);
    if (!mode->ref) apply_bc3(P, NX, NY, NZ, &bcs, 1);
    orc3_ptou(P, U, NX, NY, NZ, g);

    for (int s = 0; s < nsteps; s++) {
        if (bcs.type[4] == ORC_BC_NOZZLE && bc->phi_seq) bcs.noz_phi = bc->phi_seq[s];
        if (mode->cfl > 0.0) {
            double r = orc3_max_rate(P, NX, NY, NZ, g, dx, dy, dz);
            if (bcs.type[4] == ORC_BC_NOZZLE) {
                /* the nozzle inflow state enters through the ghost faces */
                double ex = nozzle_rate(&bcs, g, dx, dy, dz);
                if (ex > r) r = ex;
            }
            dt = mode->cfl / r;
        }
        if (mode->ref) apply_bc3(P, NX, NY, NZ, &bcs, 1);
        int nst = mode->integrator == 1 ? 3 : 2;
        for (int st = 0; st < nst; st++) {
            orc3_fluxes(P, g, fl, Fx, 0, NX, NY, NZ, mode->rec, mode->riemann);
            orc3_fluxes(P, g, fl, Fy, 1, NX, NY, NZ, mode->rec, mode->riemann);
            orc3_fluxes(P, g, fl, Fz, 2, NX, NY, NZ, mode->rec, mode->riemann);
            if (mode->integrator == 1) {
                /* SSP-RK3 (Shu & Osher): (a, b, base) = (0,1,U), (3/4,1/4,Us), (1/3,2/3,Us) */
                if (st == 0)
                    stage_update3(U, U, Us, 0.0, 1.0, dt, dx, dy, dz, Fx, Fy, Fz, NX, NY, NZ);
                else if (st == 1)
                    stage_update3(U, Us, Us, 0.75, 0.25, dt, dx, dy, dz, Fx, Fy, Fz, NX, NY, NZ);
                else
                    stage_update3(U, Us, U, ORC_MUT == ORC_MUT_RK3 ? 0.5 : 1.0 / 3.0,
                                  ORC_MUT == ORC_MUT_RK3 ? 0.5 : 2.0 / 3.0, dt, dx, dy, dz, Fx, Fy, Fz, NX, NY, NZ);
            } else {
                if (st == 0)
                    orc3_update(U, Us, dt / 2, dx, dy, dz, Fx, Fy, Fz, NX, NY, NZ);
                else
                    orc3_update(U, U, dt, dx, dy, dz, Fx, Fy, Fz, NX, NY, NZ);
            }
            const double *Uc = (mode->integrator == 1) ? (st == 2 ? U : Us) : (st == 0 ? Us : U);
            bad += orc3_c2p(Uc, P, NX, NY, NZ, g, mode->n_iter, mode->tol, NULL);
            orc3_limit(P, fl, NX, NY, NZ);
            apply_bc3(P, NX, NY, NZ, &bcs, 0);
            if (!mode->ref) apply_bc3(P, NX, NY, NZ, &bcs, 1);
        }
        t += dt;
    }
    if (t_out) *t_out = t;
    if (dt_last) *dt_last = dt;
    free(U);
    free(Us);
    free(Fx);
    free(Fy);
    free(Fz);
    return bad;
}
