/*
 * orc_2d.c -- CPU oracle, 2D solver (TEST INFRASTRUCTURE ONLY; see orc_core.h).
 *
 * Arrays use Verona's 2D AoS order (2dim/structs.jl:26-34): arr[v, i, j],
 * 4 variables, 3 ghost cells per side; flat index v + 4*(i + NX*j).
 */
#include "verona_oracle.h"
#include "orc_core.h"

#include <stdlib.h>

#define IX2(v, i, j) ((size_t)(v) + 4 * ((size_t)(i) + (size_t)NX * (size_t)(j)))

/* 2dim/structs.jl:62-76 */
void orc2_ptou(const double *P, double *U, int NX, int NY, double g)
{
    size_t n = (size_t)NX * NY;
#pragma omp parallel for schedule(static)
    for (size_t c = 0; c < n; c++) orc2_p2u(P + 4 * c, U + 4 * c, g);
}

/* 2dim/algos.jl:5-9 */
void orc2_limit(double *P, double floor_, int NX, int NY)
{
    size_t n = (size_t)NX * NY;
#pragma omp parallel for schedule(static)
    for (size_t c = 0; c < n; c++) {
        P[4 * c + 0] = jl_max(P[4 * c + 0], floor_);
        P[4 * c + 1] = jl_max(P[4 * c + 1], floor_);
    }
}

/* 2dim/algos.jl:11-168 function_Fluxes at face (i+1/2 or j+1/2), for 0-based
 * i, j in [2, N-3).  The reference's PR guard tests `j < Nx-2` where `i` is
 * meant (2dim/algos.jl:82); on square grids the flux region is unaffected and
 * this restatement uses the intended guard everywhere. */
void orc2_fluxes(const double *P, double g, double floor_, double *F, int dim, int NX, int NY,
                 int rec, int riemann)
{
    int di = dim == 0, dj = dim == 1;
#pragma omp parallel for schedule(static)
    for (int j = 2; j < NY - 3; j++)
        for (int i = 2; i < NX - 3; i++) {
            double PL[4], PR[4], UL[4], UR[4], FL[4], FR[4];
            for (int v = 0; v < 4; v++) {
                double q[6];
                for (int s = 0; s < 6; s++) q[s] = P[IX2(v, i + (s - 2) * di, j + (s - 2) * dj)];
                double lo, hi;
                orc_recon(rec, q[0], q[1], q[2], q[3], q[4], &lo, &hi);
                PL[v] = hi;
                orc_recon(rec, q[1], q[2], q[3], q[4], q[5], &lo, &hi);
                PR[v] = lo;
            }
            /* floors applied twice in the reference (:103-106, :111-114) */
            for (int r = 0; r < 2; r++)
                for (int v = 0; v < 2; v++) {
                    PL[v] = jl_max(floor_, PL[v]);
                    PR[v] = jl_max(floor_, PR[v]);
                }
            orc2_p2u(PR, UR, g);
            orc2_p2u(PL, UL, g);
            orc2_p2f(PR, FR, g, dim);
            orc2_p2f(PL, FL, g, dim);
            double lorL = sqrt(PL[2] * PL[2] + PL[3] * PL[3] + 1.0);
            double lorR = sqrt(PR[2] * PR[2] + PR[3] * PR[3] + 1.0);
            double vL = PL[2 + dim] / lorL;
            double vR = PR[2 + dim] / lorR;
            double CL = orc_sound_speed(PL[0], PL[1], g);
            double CR = orc_sound_speed(PR[0], PR[1], g);
            double lpL, lmL, lpR, lmR;
            orc_lambdas(vL, CL, lorL * lorL, &lpL, &lmL);
            orc_lambdas(vR, CR, lorR * lorR, &lpR, &lmR);
            double Cmax = jl_max(lpL, lpR);
            double Cmin = -jl_min(lmL, lmR);
            double *Fo = F + IX2(0, i, j);
            if (riemann == 0)
                orc_hll(4, UL, UR, FL, FR, Cmax, Cmin, Fo);
            else
                orc_hllc(4, 2 + dim, UL, UR, FL, FR, vL, vR, (g - 1.0) * PL[1],
                         (g - 1.0) * PR[1], Cmax, Cmin, Fo);
        }
}

/* 2dim/algos.jl:170-189 */
void orc2_update(const double *U, double *Ub, double dt, double dx, double dy, const double *Fx,
                 const double *Fy, int NX, int NY)
{
#pragma omp parallel for schedule(static)
    for (int j = 3; j < NY - 3; j++)
        for (int i = 3; i < NX - 3; i++)
            for (int v = 0; v < 4; v++) {
                size_t c = IX2(v, i, j);
                Ub[c] = U[c] - dt / dx * (Fx[c] - Fx[IX2(v, i - 1, j)]) -
                        dt / dy * (Fy[c] - Fy[IX2(v, i, j - 1)]);
            }
}

/* 2dim/structs.jl:121-208  LU_dec_pivot! (row-scaled partial pivoting, explicit
 * fma).  Returns 1 when a row is identically zero (the reference raises). */
static int lu4_pivot(double *A, double *t, double *x)
{
#define A_(i, j) A[((j) - 1) * 4 + (i) - 1]
    double rn[4];
    for (int i = 1; i <= 4; i++) {
        double absmax = 0.0;
        for (int j = 1; j <= 4; j++) {
            double tmp = fabs(A_(i, j));
            if (tmp > absmax) absmax = tmp;
        }
        if (absmax == 0.0) return 1;
        rn[i - 1] = 1.0 / absmax;
    }
    for (int k = 1; k <= 4; k++) {
        int pivot = k;
        double pv = fabs(A_(k, k)) * rn[k - 1];
        for (int i = k + 1; i <= 4; i++) {
            double tmp = fabs(A_(i, k)) * rn[i - 1];
            if (tmp > pv) {
                pv = tmp;
                pivot = i;
            }
        }
        if (pivot != k) {
            for (int j = 1; j <= 4; j++) {
                double tmp = A_(k, j);
                A_(k, j) = A_(pivot, j);
                A_(pivot, j) = tmp;
            }
            double tmp = t[k - 1];
            t[k - 1] = t[pivot - 1];
            t[pivot - 1] = tmp;
            tmp = rn[k - 1];
            rn[k - 1] = rn[pivot - 1];
            rn[pivot - 1] = tmp;
        }
        for (int i = k + 1; i <= 4; i++) {
            A_(i, k) /= A_(k, k);
            for (int j = k + 1; j <= 4; j++) A_(i, j) = fma(-A_(i, k), A_(k, j), A_(i, j));
        }
    }
    for (int i = 1; i <= 4; i++) {
        x[i - 1] = t[i - 1];
        for (int j = 1; j <= i - 1; j++) x[i - 1] = fma(-A_(i, j), x[j - 1], x[i - 1]);
    }
    for (int i = 4; i >= 1; i--) {
        for (int j = i + 1; j <= 4; j++) x[i - 1] = fma(-A_(i, j), x[j - 1], x[i - 1]);
        x[i - 1] /= A_(i, i);
    }
#undef A_
    return 0;
}

/* 2dim/structs.jl:249-330  function_UtoP: 4x4 Newton on the interior, warm
 * start = P, break-before-update when ||dx||_2 < tol.  Returns the number of
 * singular-row events (cell left at its current iterate). */
long orc2_c2p(const double *U, double *P, int NX, int NY, double g, int n_iter, double tol,
              long *n_singular)
{
    long sing = 0;
#pragma omp parallel for schedule(static) reduction(+ : sing)
    for (int j = 3; j < NY - 3; j++)
        for (int i = 3; i < NX - 3; i++) {
            double *p = P + IX2(0, i, j);
            const double *u = U + IX2(0, i, j);
            for (int it = 0; it < n_iter; it++) {
                double gam = sqrt(p[2] * p[2] + p[3] * p[3] + 1.0);
                double w = g * p[1] + p[0];
                double f[4], J[16], o[4];
                f[0] = gam * p[0] - u[0];
                f[1] = (g - 1.0) * p[1] - gam * gam * w - u[1];
                f[2] = p[2] * gam * w - u[2];
                f[3] = p[3] * gam * w - u[3];
                J[0] = gam;
                J[4] = 0.0;
                J[8] = p[0] * p[2] / gam;
                J[12] = p[0] * p[3] / gam;
                J[1] = -(gam * gam);
                J[5] = (g - 1.0) - gam * gam * g;
                J[9] = -2.0 * p[2] * w;
                J[13] = -2.0 * p[3] * w;
                J[2] = gam * p[2];
                J[6] = gam * p[2] * g;
                J[10] = (2.0 * (p[2] * p[2]) + p[3] * p[3] + 1.0) / gam * w;
                J[14] = p[2] * p[3] / gam * w;
                J[3] = gam * p[3];
                J[7] = gam * p[3] * g;
                J[11] = p[2] * p[3] / gam * w;
                J[15] = (2.0 * (p[3] * p[3]) + p[2] * p[2] + 1.0) / gam * w;
                if (lu4_pivot(J, f, o)) {
                    sing++;
                    break;
                }
                if (sqrt(o[0] * o[0] + o[1] * o[1] + o[2] * o[2] + o[3] * o[3]) < tol) break;
                for (int v = 0; v < 4; v++) p[v] = p[v] - o[v];
            }
        }
    if (n_singular) *n_singular = sing;
    return sing;
}

/* Single-rank periodic exchange, 2dim/sync_CUDA.jl:3-72 (x then y, whole planes). */
void orc2_fill_periodic(double *P, int NX, int NY, int mask)
{
    if (mask & 1)
        for (int j = 0; j < NY; j++)
            for (int g3 = 0; g3 < 3; g3++)
                for (int v = 0; v < 4; v++) {
                    P[IX2(v, g3, j)] = P[IX2(v, NX - 6 + g3, j)];
                    P[IX2(v, NX - 3 + g3, j)] = P[IX2(v, 3 + g3, j)];
                }
    if (mask & 2)
        for (int g3 = 0; g3 < 3; g3++)
            for (int i = 0; i < NX; i++)
                for (int v = 0; v < 4; v++) {
                    P[IX2(v, i, g3)] = P[IX2(v, i, NY - 6 + g3)];
                    P[IX2(v, i, NY - 3 + g3)] = P[IX2(v, i, 3 + g3)];
                }
}

/* Zero-gradient copies (examples/Jet2D/HARM_jet_breakout.jl:88-98), x then y. */
void orc2_fill_outflow(double *P, int NX, int NY, int lo_mask, int hi_mask)
{
    for (int j = 0; j < NY; j++)
        for (int g3 = 0; g3 < 3; g3++)
            for (int v = 0; v < 4; v++) {
                if (lo_mask & 1) P[IX2(v, g3, j)] = P[IX2(v, 3, j)];
                if (hi_mask & 1) P[IX2(v, NX - 3 + g3, j)] = P[IX2(v, NX - 4, j)];
            }
    for (int g3 = 0; g3 < 3; g3++)
        for (int i = 0; i < NX; i++)
            for (int v = 0; v < 4; v++) {
                if (lo_mask & 2) P[IX2(v, i, g3)] = P[IX2(v, i, 3)];
                if (hi_mask & 2) P[IX2(v, i, NY - 3 + g3)] = P[IX2(v, i, NY - 4)];
            }
}

double orc2_max_rate(const double *P, int NX, int NY, double g, double dx, double dy)
{
    double mx = 0.0;
    double inv[2] = {1.0 / dx, 1.0 / dy};
#pragma omp parallel for schedule(static) reduction(max : mx)
    for (int j = 3; j < NY - 3; j++)
        for (int i = 3; i < NX - 3; i++) {
            const double *p = P + IX2(0, i, j);
            double lor = sqrt(p[2] * p[2] + p[3] * p[3] + 1.0);
            double cs = orc_sound_speed(p[0], p[1], g);
            double r = 0.0;
            for (int d = 0; d < 2; d++) {
                double lp, lm;
                orc_lambdas(p[2 + d] / lor, cs, lor * lor, &lp, &lm);
                r = r + jl_max(fabs(lp), fabs(lm)) * inv[d];
            }
            if (r > mx) mx = r;
        }
    return mx;
}

static void apply_bc2(double *P, int NX, int NY, const int *bc, int nonperiodic)
{
    int pmask = 0, lo = 0, hi = 0;
    for (int d = 0; d < 2; d++) {
        if (bc[2 * d] == ORC_BC_PERIODIC) pmask |= 1 << d;
        if (bc[2 * d] == ORC_BC_OUTFLOW) lo |= 1 << d;
        if (bc[2 * d + 1] == ORC_BC_OUTFLOW) hi |= 1 << d;
    }
    if (nonperiodic) {
        if (lo | hi) orc2_fill_outflow(P, NX, NY, lo, hi);
    } else if (pmask) {
        orc2_fill_periodic(P, NX, NY, pmask);
    }
}

static void stage_update2(const double *U0, const double *Ub, double *Uout, double a, double b,
                          double dt, double dx, double dy, const double *Fx, const double *Fy,
                          int NX, int NY)
{
#pragma omp parallel for schedule(static)
    for (int j = 3; j < NY - 3; j++)
        for (int i = 3; i < NX - 3; i++)
            for (int v = 0; v < 4; v++) {
                size_t c = IX2(v, i, j);
                double acc = Ub[c] - dt / dx * (Fx[c] - Fx[IX2(v, i - 1, j)]);
                acc = acc - dt / dy * (Fy[c] - Fy[IX2(v, i, j - 1)]);
                Uout[c] = a == 0.0 ? acc : a * U0[c] + b * acc;
            }
}

/* Verona2D.HARM_HLL step sequence (2dim/algos.jl:192-324) in ref mode; the
 * extended form as in orc3_run.  bc_type[4] = {xlo, xhi, ylo, yhi}. */
long orc2_run(double *P, int NX, int NY, const orc_mode *mode, const int *bc_type, int nsteps,
              double *t_out, double *dt_last)
{
    size_t n = (size_t)NX * NY * 4;
    double *U = calloc(n, sizeof(double));
    double *Us = calloc(n, sizeof(double));
    double *Fx = calloc(n, sizeof(double));
    double *Fy = calloc(n, sizeof(double));
    double g = mode->gamma, fl = mode->floor_, dx = mode->dx, dy = mode->dy;
    double t = 0.0, dt = mode->dt;
    long bad = 0;
    apply_bc2(P, NX, NY, bc_type, 0);
    if (!mode->ref) apply_bc2(P, NX, NY, bc_type, 1);
    orc2_ptou(P, U, NX, NY, g);
    for (int s = 0; s < nsteps; s++) {
        if (mode->cfl > 0.0) dt = mode->cfl / orc2_max_rate(P, NX, NY, g, dx, dy);
        if (mode->ref) apply_bc2(P, NX, NY, bc_type, 1);
        int nst = mode->integrator == 1 ? 3 : 2;
        for (int st = 0; st < nst; st++) {
            orc2_fluxes(P, g, fl, Fx, 0, NX, NY, mode->rec, mode->riemann);
            orc2_fluxes(P, g, fl, Fy, 1, NX, NY, mode->rec, mode->riemann);
            if (mode->integrator == 1) {
                if (st == 0)
                    stage_update2(U, U, Us, 0.0, 1.0, dt, dx, dy, Fx, Fy, NX, NY);
                else if (st == 1)
                    stage_update2(U, Us, Us, 0.75, 0.25, dt, dx, dy, Fx, Fy, NX, NY);
                else
                    stage_update2(U, Us, U, ORC_MUT == ORC_MUT_RK3 ? 0.5 : 1.0 / 3.0,
                                  ORC_MUT == ORC_MUT_RK3 ? 0.5 : 2.0 / 3.0, dt, dx, dy, Fx, Fy, NX, NY);
            } else {
                if (st == 0)
                    orc2_update(U, Us, dt / 2, dx, dy, Fx, Fy, NX, NY);
                else
                    orc2_update(U, U, dt, dx, dy, Fx, Fy, NX, NY);
            }
            const double *Uc = (mode->integrator == 1) ? (st == 2 ? U : Us) : (st == 0 ? Us : U);
            bad += orc2_c2p(Uc, P, NX, NY, g, mode->n_iter, mode->tol, NULL);
            orc2_limit(P, fl, NX, NY);
            apply_bc2(P, NX, NY, bc_type, 0);
            if (!mode->ref) apply_bc2(P, NX, NY, bc_type, 1);
        }
        t += dt;
    }
    if (t_out) *t_out = t;
    if (dt_last) *dt_last = dt;
    free(U);
    free(Us);
    free(Fx);
    free(Fy);
    return bad;
}
