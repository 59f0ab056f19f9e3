/*
 * orc_1d.c -- CPU oracle, 1D solver (TEST INFRASTRUCTURE ONLY; see orc_core.h).
 *
 * ParVector1D is SoA with no ghost cells (1dim/structs.jl:16-28): three
 * arrays rho[N], u[N], ux[N] (1-based cells 1..N in the reference).
 */
#include "verona_oracle.h"
#include "orc_core.h"

#include <stdlib.h>
#include <string.h>

/* 1dim/structs.jl:48-54  F_ptoU */
static inline void f_ptou(const double *x, double *b, double g)
{
    double gam = sqrt(x[2] * x[2] + 1.0);
    double w = g * x[1] + x[0];
    b[0] = gam * x[0];
    b[1] = (g - 1.0) * x[1] - gam * gam * w;
    b[2] = x[2] * gam * w;
}

/* 1dim/structs.jl:32-46  Jacobian, column-major 3x3 */
static inline void jacobian1(const double *x, double *J, double g)
{
    double gam = sqrt(1.0 + x[2] * x[2]);
    double w = g * x[1] + x[0];
#define J_(i, j) J[((j) - 1) * 3 + (i) - 1]
    J_(1, 1) = gam;
    J_(1, 2) = 0.0;
    J_(1, 3) = x[0] * x[2] / gam;
    J_(2, 1) = -(gam * gam);
    J_(2, 2) = (g - 1.0) - gam * gam * g;
    J_(2, 3) = ORC_MUT == ORC_MUT_JAC ? -(x[2] * w) : -2.0 * x[2] * w;
    J_(3, 1) = gam * x[2];
    J_(3, 2) = gam * x[2] * g;
    J_(3, 3) = (2.0 * (x[2] * x[2]) + 1.0) / gam * w;
#undef J_
}

/*
 * `J \ b` for a 3x3 static matrix as done by StaticArrays.jl 1.9.11 (a
 * third-party dependency of the reference, pinned in Manifest.toml:597-601
 * and not vendored): closed-form adjugate solve divided by
 * det = dot(col1, cross(col2, col3)).  Restated from StaticArrays' published
 * solve/det source; its last-ulp rounding is not pinned by any reference test.
 */
static inline void sa_solve3(const double *a, const double *b, double *x)
{
#define a_(i, j) a[((j) - 1) * 3 + (i) - 1]
    double c0 = a[4] * a[8] - a[5] * a[7];
    double c1 = a[5] * a[6] - a[3] * a[8];
    double c2 = a[3] * a[7] - a[4] * a[6];
    double d = a[0] * c0 + a[1] * c1 + a[2] * c2;
    x[0] = ((a_(2, 2) * a_(3, 3) - a_(2, 3) * a_(3, 2)) * b[0] +
            (a_(1, 3) * a_(3, 2) - a_(1, 2) * a_(3, 3)) * b[1] +
            (a_(1, 2) * a_(2, 3) - a_(1, 3) * a_(2, 2)) * b[2]) / d;
    x[1] = ((a_(2, 3) * a_(3, 1) - a_(2, 1) * a_(3, 3)) * b[0] +
            (a_(1, 1) * a_(3, 3) - a_(1, 3) * a_(3, 1)) * b[1] +
            (a_(1, 3) * a_(2, 1) - a_(1, 1) * a_(2, 3)) * b[2]) / d;
    x[2] = ((ORC_MUT == ORC_MUT_SOLVE3 ? a_(1, 2) * a_(2, 3) - a_(2, 2) * a_(3, 1)
                                       : a_(2, 1) * a_(3, 2) - a_(2, 2) * a_(3, 1)) * b[0] +
            (a_(1, 2) * a_(3, 1) - a_(1, 1) * a_(3, 2)) * b[1] +
            (a_(1, 1) * a_(2, 2) - a_(1, 2) * a_(2, 1)) * b[2]) / d;
#undef a_
}

/* Point functions of the 1D Newton, exported for their pins: F_ptoU(x) - U
 * is the residual, J its analytic Jacobian (column-major, J[(j-1)*3+i-1] =
 * J[i,j]), sa_solve3 the Newton step J^-1 f. */
void orc1_f_ptou(const double *x, double g, double *f) { f_ptou(x, f, g); }
void orc1_jacobian(const double *x, double g, double *J) { jacobian1(x, J, g); }
void orc1_solve3(const double *A, const double *b, double *x) { sa_solve3(A, b, x); }

/* 1dim/structs.jl:78-105  UtoP: 3x3 Newton per cell, warm start = P,
 * break when norm(dx) < tol before applying dx. */
void orc1_c2p(const double *U1, const double *U2, const double *U3, double *P1, double *P2,
              double *P3, long N, double g, int n_iter, double tol)
{
#pragma omp parallel for schedule(static)
    for (long i = 0; i < N; i++) {
        double x[3] = {P1[i], P2[i], P3[i]};
        for (int it = 0; it < n_iter; it++) {
            double f[3], J[9], d[3];
            f_ptou(x, f, g);
            jacobian1(x, J, g);
            f[0] = f[0] - U1[i];
            f[1] = f[1] - U2[i];
            f[2] = f[2] - U3[i];
            sa_solve3(J, f, d);
            if (ORC_MUT == ORC_MUT_BREAK) {
                x[0] = x[0] - d[0];
                x[1] = x[1] - d[1];
                x[2] = x[2] - d[2];
            }
            if (sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) < tol) break;
            if (ORC_MUT != ORC_MUT_BREAK) {
                x[0] = x[0] - d[0];
                x[1] = x[1] - d[1];
                x[2] = x[2] - d[2];
            }
        }
        P1[i] = x[0];
        P2[i] = x[1];
        P3[i] = x[2];
    }
}

/* HLL (or HLLC) flux at one face from PL = left state, PR = right state,
 * 1dim/algos.jl:127-171 (the reference loop is serial over faces). */
static inline void face1(const double *PL, const double *PR, double g, int riemann, double *F)
{
    double UL[3], UR[3], FL[3], FR[3];
    orc1_p2u(PR, UR, g);
    orc1_p2u(PL, UL, g);
    orc1_p2f(PR, FR, g);
    orc1_p2f(PL, FL, g);
    double vL = PL[2] / sqrt(PL[2] * PL[2] + 1.0);
    double vR = PR[2] / sqrt(PR[2] * PR[2] + 1.0);
    double CL = orc_sound_speed(PL[0], PL[1], g);
    double CR = orc_sound_speed(PR[0], PR[1], g);
    double lpL, lmL, lpR, lmR;
    orc_lambdas(vL, CL, PL[2] * PL[2] + 1.0, &lpL, &lmL);
    orc_lambdas(vR, CR, PR[2] * PR[2] + 1.0, &lpR, &lmR);
    double Cmax = jl_max(lpL, lpR);
    double Cmin = -jl_min(lmL, lmR);
    if (riemann == 0)
        orc_hll(3, UL, UR, FL, FR, Cmax, Cmin, F);
    else
        orc_hllc(3, 2, UL, UR, FL, FR, vL, vR, (g - 1.0) * PL[1], (g - 1.0) * PR[1], Cmax, Cmin, F);
}

/* Reconstruction pass, 1dim/algos.jl:91-117: writes PR[i-1], PL[i]
 * (0-based cells) for i in [1, N-1) (minmod) or [2, N-2) (PPM/WENOZ). */
static void recon1(int rec, const double *const *P, double *const *PL, double *const *PR, long N)
{
    long lo = rec == ORC_REC_MINMOD ? 1 : 2;
    long hi = rec == ORC_REC_MINMOD ? N - 1 : N - 2;
    for (long i = lo; i < hi; i++)
        for (int v = 0; v < 3; v++) {
            const double *q = P[v];
            double l, h;
            if (rec == ORC_REC_MINMOD)
                orc_minmod(q[i - 1], q[i], q[i + 1], &l, &h);
            else
                orc_recon(rec, q[i - 2], q[i - 1], q[i], q[i + 1], q[i + 2], &l, &h);
            PR[v][i - 1] = l;
            PL[v][i] = h;
        }
}

/*
 * Verona1D.HARM_HLL (1dim/algos.jl:58-315) on P (in/out).  Runs `while t < T`
 * with t += dt, stopping early after max_steps steps if max_steps > 0.
 * PR/PL start as copies of the initial P and keep stale entries where the
 * reconstruction does not write (the reference's frozen-edge behaviour);
 * cells 0, N-2, N-1 (0-based) are never updated.  riemann == 1 swaps the HLL
 * face solve for HLLC.  Returns the number of steps taken.
 */
int orc1_run(double *P1, double *P2, double *P3, long N, double dt, double dx, double T,
             int max_steps, double g, int rec, int riemann, int n_iter, double tol, double *t_out)
{
    size_t sz = (size_t)N * sizeof(double);
    double *buf = malloc(sz * 3 * 7);
    double *PLa[3], *PRa[3], *Ua[3], *Uh[3], *Ph[3], *Fa[3];
    double *Pa[3] = {P1, P2, P3};
    for (int v = 0; v < 3; v++) {
        PLa[v] = buf + (size_t)N * (0 * 3 + v);
        PRa[v] = buf + (size_t)N * (1 * 3 + v);
        Ua[v] = buf + (size_t)N * (2 * 3 + v);
        Uh[v] = buf + (size_t)N * (3 * 3 + v);
        Ph[v] = buf + (size_t)N * (4 * 3 + v);
        Fa[v] = buf + (size_t)N * (5 * 3 + v);
        memcpy(PLa[v], Pa[v], sz);
        memcpy(PRa[v], Pa[v], sz);
        memset(Fa[v], 0, sz);
    }
    for (long i = 0; i < N; i++) {
        double p[3] = {P1[i], P2[i], P3[i]}, u[3];
        orc1_p2u(p, u, g);
        Ua[0][i] = u[0];
        Ua[1][i] = u[1];
        Ua[2][i] = u[2];
    }
    for (int v = 0; v < 3; v++) memcpy(Uh[v], Ua[v], sz);

    double t = 0.0;
    int steps = 0;
    while (t < T) {
        if (max_steps > 0 && steps >= max_steps) break;
        for (int stage = 0; stage < 2; stage++) {
            double *const *Pin = stage == 0 ? Pa : Ph;
            recon1(rec, (const double *const *)Pin, PLa, PRa, N);
            for (long i = 0; i < N; i++) {
                double pl[3] = {PLa[0][i], PLa[1][i], PLa[2][i]};
                double pr[3] = {PRa[0][i], PRa[1][i], PRa[2][i]};
                double f[3];
                face1(pl, pr, g, riemann, f);
                Fa[0][i] = f[0];
                Fa[1][i] = f[1];
                Fa[2][i] = f[2];
            }
            if (stage == 0) {
                for (long i = 1; i < N - 2; i++)
                    for (int v = 0; v < 3; v++)
                        Uh[v][i] = Ua[v][i] - 0.5 * dt / dx * (Fa[v][i] - Fa[v][i - 1]);
                for (int v = 0; v < 3; v++) memcpy(Ph[v], Pa[v], sz);
                orc1_c2p(Uh[0], Uh[1], Uh[2], Ph[0], Ph[1], Ph[2], N, g, n_iter, tol);
            } else {
                for (long i = 1; i < N - 2; i++)
                    for (int v = 0; v < 3; v++)
                        Ua[v][i] = Ua[v][i] - dt / dx * (Fa[v][i] - Fa[v][i - 1]);
                orc1_c2p(Ua[0], Ua[1], Ua[2], P1, P2, P3, N, g, n_iter, tol);
            }
        }
        t += dt;
        steps++;
    }
    if (t_out) *t_out = t;
    free(buf);
    return steps;
}

/* ---------------- vectorised point functions for tests ---------------- */

void orc_wenoz_v(const double *q5, long n, double *lo, double *hi)
{
    for (long i = 0; i < n; i++)
        orc_wenoz(q5[5 * i], q5[5 * i + 1], q5[5 * i + 2], q5[5 * i + 3], q5[5 * i + 4], lo + i, hi + i);
}

void orc_ppm_v(const double *q5, long n, double *lo, double *hi)
{
    for (long i = 0; i < n; i++)
        orc_ppm(q5[5 * i], q5[5 * i + 1], q5[5 * i + 2], q5[5 * i + 3], q5[5 * i + 4], lo + i, hi + i);
}

void orc_minmod_v(const double *q3, long n, double *lo, double *hi)
{
    for (long i = 0; i < n; i++) orc_minmod(q3[3 * i], q3[3 * i + 1], q3[3 * i + 2], lo + i, hi + i);
}

void orc3_p2u_v(const double *P, long n, double g, double *U)
{
    for (long i = 0; i < n; i++) orc3_p2u(P + 5 * i, U + 5 * i, g);
}

void orc3_p2f_v(const double *P, long n, double g, int dim, double *F)
{
    for (long i = 0; i < n; i++) orc3_p2f(P + 5 * i, F + 5 * i, g, dim);
}

/* Riemann flux from already-reconstructed, already-floored face states. */
void orc3_face_flux_v(const double *PLv, const double *PRv, long n, double g, int dim, int riemann,
                      double *F, double *Cmaxo, double *Cmino)
{
    for (long c = 0; c < n; c++) {
        const double *PL = PLv + 5 * c, *PR = PRv + 5 * c;
        double UL[5], UR[5], FL[5], FR[5];
        orc3_p2u(PR, UR, g);
        orc3_p2u(PL, UL, g);
        orc3_p2f(PR, FR, g, dim);
        orc3_p2f(PL, FL, g, dim);
        double lorL = sqrt(PL[2] * PL[2] + PL[3] * PL[3] + PL[4] * PL[4] + 1.0);
        double lorR = sqrt(PR[2] * PR[2] + PR[3] * PR[3] + PR[4] * PR[4] + 1.0);
        double vL = PL[2 + dim] / lorL, vR = PR[2 + dim] / lorR;
        double CL = orc_sound_speed(PL[0], PL[1], g), CR = orc_sound_speed(PR[0], PR[1], g);
        double lpL, lmL, lpR, lmR;
        orc_lambdas(vL, CL, lorL * lorL, &lpL, &lmL);
        orc_lambdas(vR, CR, lorR * lorR, &lpR, &lmR);
        double Cmax = jl_max(lpL, lpR);
        double Cmin = -jl_min(lmL, lmR);
        if (Cmaxo) Cmaxo[c] = Cmax;
        if (Cmino) Cmino[c] = Cmin;
        if (riemann == 0)
            orc_hll(5, UL, UR, FL, FR, Cmax, Cmin, F + 5 * c);
        else
            orc_hllc(5, 2 + dim, UL, UR, FL, FR, vL, vR, (g - 1.0) * PL[1], (g - 1.0) * PR[1],
                     Cmax, Cmin, F + 5 * c);
    }
}

void orc1_face_flux_v(const double *PL, const double *PR, long n, double g, int riemann, double *F)
{
    for (long c = 0; c < n; c++) face1(PL + 3 * c, PR + 3 * c, g, riemann, F + 3 * c);
}
