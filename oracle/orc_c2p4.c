/*
 * orc_c2p4.c -- CPU oracle (test infrastructure) of the batched con2prim
 * with a choice of initial guess.
 *
 * The iteration is the 1DW Newton of Verona's UtoP,
 * src/3dim/structs.jl:292-342 (W_min clamp, v^2 clamp to [0, 1 - 1e-15],
 * halving towards W_min), with the stop |dW| < rtol * W instead of the
 * reference's absolute |dW| < tol, and the recovery of :598-609 without the
 * Gamma <= 50 cap.  The initial W of a primitive guess is formed as
 * :276-287 forms it.  Same operation order as the HIP kernel, so for the same
 * initial W the results agree bit for bit.
 */
#include <math.h>
#include <stdint.h>

#include "orc_core.h"

/* W = (rho + gamma u) Gamma^2 of a primitive guess (src/3dim/structs.jl:276-287). */
void orc4_w0_from_prims(const double *Pg, long n, double g, double *W0)
{
    for (long c = 0; c < n; c++) {
        const double *p = Pg + 5 * c;
        double gm = sqrt(p[2] * p[2] + p[3] * p[3] + p[4] * p[4] + 1.0);
        W0[c] = (p[0] + g * p[1]) * (gm * gm);
    }
}

/* One cell: returns 1 when the 1DW iteration converged (P written), else 0. */
int orc4_newton_cell(const double *U, double W, double g, int n_iter, double rtol, double *P, int *iters)
{
    const double D = U[0], Qt = U[1], Qx = U[2], Qy = U[3], Qz = U[4];
    const double c = (g - 1.0) / g;
    const double S2 = Qx * Qx + Qy * Qy + Qz * Qz;
    const double W_max = 1e30;
    const double W_min = sqrt(S2) * (1.0 + 1e-15);
    for (int t10 = 0; t10 < 1100 && S2 / (W * W) >= 1.0 && W < W_max; t10++) W *= 10.0;
    int conv = 0;
    for (int it = 0; it < n_iter; it++) {
        (*iters)++;
        if (W < W_min) W = W_min;
        double v2 = S2 / (W * W);
        if (v2 < 0.0)
            v2 = 0.0;
        else if (v2 > 1.0 - 1e-15)
            v2 = 1.0 - 1e-15;
        double sv = sqrt(1.0 - v2);
        double fun = Qt + W - c * (W * (1.0 - v2) - D * sv);
        double W3 = W * W * W;
        double jac = 1.0 - c * (1.0 - v2) - c * (W * (2.0 * S2 / W3) - D * (S2 / (W3 * sv)));
        double dW = fun / jac;
        double Wp = W - dW;
        W = (Wp < W_min) ? 0.5 * (W + W_min) : Wp;
        if (dW < 0.0) dW = -dW;
        if (dW < rtol * W) {
            conv = 1;
            break;
        }
    }
    if (!conv) return 0;
    double v2 = S2 / (W * W);
    double gam = 1.0 / sqrt(1.0 - v2);
    P[0] = jl_max(1e-8, D / gam);
    P[1] = jl_max(1e-8, (W / (gam * gam) - D / gam) / g);
    P[2] = gam * Qx / W;
    P[3] = gam * Qy / W;
    P[4] = gam * Qz / W;
    return 1;
}

/* Batch: U, P are [n][5]; returns the number of cells that did not converge. */
long orc4_solve(const double *U, const double *W0, long n, double g, int n_iter, double rtol, double *P,
                int *iters, unsigned char *conv)
{
    long bad = 0;
#pragma omp parallel for schedule(static) reduction(+ : bad)
    for (long c = 0; c < n; c++) {
        int it = 0;
        int ok = orc4_newton_cell(U + 5 * c, W0[c], g, n_iter, rtol, P + 5 * c, &it);
        if (iters) iters[c] = it;
        if (conv) conv[c] = (unsigned char)ok;
        bad += !ok;
    }
    return bad;
}

/* Features of the learned guess (FP32), E = -T^t_t and z = (E^2 - S^2 - D^2)/E^2
 * as the HIP kernel forms them; ok[c] = 0 outside the model's domain. */
void orc4_features(const double *U, long n, float *X, double *E, double *z, unsigned char *ok)
{
    for (long c = 0; c < n; c++) {
        const double *u = U + 5 * c;
        const double D = u[0];
        const double S2 = u[2] * u[2] + u[3] * u[3] + u[4] * u[4];
        const double e = -u[1];
        const double tau = e - D;
        const double rE = 1.0 / e;
        const double zz = (tau * (e + D) - S2) * (rE * rE);
        const int good = D > 0.0 && tau > 0.0 && zz > 0.0 && zz < 1.0 && isfinite(e) && isfinite(S2);
        float *x = X + 4 * c;
        if (good) {
            const float zf = (float)zz;
            x[0] = (float)(D * rE);
            x[1] = sqrtf((float)S2) * (float)rE;
            x[2] = sqrtf(zf);
            x[3] = 0.1f * logf(zf);
        } else {
            x[0] = x[1] = x[2] = x[3] = 0.0f;
        }
        E[c] = e;
        z[c] = zz;
        ok[c] = (unsigned char)good;
    }
}
