/*
 * orc_core.h -- CPU oracle: point functions of Verona's per-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference's
 * algorithm used as the checker for the HIP library.  Only tests/, the smoke
 * check and the CPU-baseline leg of bench.py may load it; the product
 * (verona_amd/) never includes, links or calls it.
 *
 * Semantics follow the Julia reference operation-for-operation (compile with
 * -ffp-contract=off): literal powers x^2 = x*x and x^3 = (x*x)*x, left-to-right
 * association, NaN-propagating max/min (Julia Base 1.11 IEEE-float max/min),
 * Julia sign().  Citations are to the reference tree (Wesenheit/Verona 0.1.0).
 *
 * Parity pinning of each function is listed in oracle/README.md and is
 * exercised by tests/test_oracle_*.py.
 */
#ifndef ORC_CORE_H
#define ORC_CORE_H

#include <math.h>
#include <stdint.h>

#define ORC_INLINE static inline

/*
 * Planted mistakes.  The real oracle is built with ORC_MUT undefined (0).
 * tests/test_oracle_mutants.py builds one extra copy of the oracle per value
 * below, each carrying exactly one plausible restatement mistake, and checks
 * that the pin written for that function fails on it: a pin that a wrong
 * oracle also passes would pin nothing.
 */
#ifndef ORC_MUT
#define ORC_MUT 0
#endif
enum {
    ORC_MUT_RK3 = 1,   /* SSP-RK3 last stage (1/2, 1/2) instead of (1/3, 2/3) */
    ORC_MUT_PPM_C2,    /* PPM extremum limiter C2 = 1.0 instead of 1.25 */
    ORC_MUT_PPM_EXT,   /* PPM step 4: extremum test with && instead of || */
    ORC_MUT_JAC,       /* 1D Jacobian J[2,3] = -x3 w instead of -2 x3 w */
    ORC_MUT_SOLVE3,    /* 3x3 solve: one adjugate entry transposed */
    ORC_MUT_BREAK,     /* 1D Newton applies the step before the norm test */
    ORC_MUT_F1DW,      /* 1DW f'(W): the D S^2 / (W^3 sqrt(1 - v^2)) term dropped */
    ORC_MUT_COUNT
};

/* Julia Base: max(x::T, y::T) for IEEE floats (NaN-propagating, -0 < +0). */
ORC_INLINE double jl_max(double x, double y)
{
    double diff = x - y;
    double am = signbit(diff) ? y : x;
    return (isnan(x) || isnan(y)) ? diff : am;
}

/* Julia Base: min(x::T, y::T) for IEEE floats. */
ORC_INLINE double jl_min(double x, double y)
{
    double diff = x - y;
    double am = signbit(diff) ? x : y;
    return (isnan(x) || isnan(y)) ? diff : am;
}

/* Julia Base: sign(x) -> -1, +1, or x itself for +-0 and NaN. */
ORC_INLINE double jl_sign(double x)
{
    return x < 0.0 ? -1.0 : (x > 0.0 ? 1.0 : x);
}

ORC_INLINE double sqr(double x) { return x * x; }

/* eos.jl:15-17  SoundSpeed(rho, u, eos) */
ORC_INLINE double orc_sound_speed(double rho, double u, double g)
{
    return sqrt((g * (g - 1.0) * u) / (rho + g * u));
}

/* fluxlimiter.jl:1-13  MINMOD -> (q_i - dU, q_i + dU) = (lower face, upper face) */
ORC_INLINE void orc_minmod(double qm1, double q, double qp1, double *lo, double *hi)
{
    double sp = qp1 - q;
    double sm = q - qm1;
    double ssp = jl_sign(sp);
    double ssm = jl_sign(sm);
    double asp = fabs(sp);
    double asm_ = fabs(sm);
    double dU = 0.25 * (ssp + ssm) * jl_min(asp, asm_);
    *lo = q - dU;
    *hi = q + dU;
}

/* fluxlimiter.jl:19-71  WENOZ -> (qr_i, ql_ip1) = (lower face i-1/2, upper face i+1/2) */
ORC_INLINE void orc_wenoz(double qm2, double qm1, double q, double qp1, double qp2,
                          double *lo, double *hi)
{
    const double b1c = 13.0 / 12.0, b2c = 0.25;
    double beta0 = b1c * sqr(qm2 + q - 2.0 * qm1) + b2c * sqr(qm2 + 3.0 * q - 4.0 * qm1);
    double beta1 = b1c * sqr(qm1 + qp1 - 2.0 * q) + b2c * sqr(qm1 - qp1);
    double beta2 = b1c * sqr(qp2 + q - 2.0 * qp1) + b2c * sqr(qp2 + 3.0 * q - 4.0 * qp1);
    const double epsL = 1.0e-42;
    double tau5 = fabs(beta0 - beta2);
    double ind0 = tau5 / (beta0 + epsL);
    double ind1 = tau5 / (beta1 + epsL);
    double ind2 = tau5 / (beta2 + epsL);

    double f0 = 2.0 * qm2 - 7.0 * qm1 + 11.0 * q;
    double f1 = -qm1 + 5.0 * q + 2.0 * qp1;
    double f2 = 2.0 * q + 5.0 * qp1 - qp2;
    double a0 = 0.1 * (1.0 + sqr(ind0));
    double a1 = 0.6 * (1.0 + sqr(ind1));
    double a2 = 0.3 * (1.0 + sqr(ind2));
    double asum = 6.0 * (a0 + a1 + a2);
    *hi = (f0 * a0 + f1 * a1 + f2 * a2) / asum;

    f0 = 2.0 * qp2 - 7.0 * qp1 + 11.0 * q;
    f1 = -qp1 + 5.0 * q + 2.0 * qm1;
    f2 = 2.0 * q + 5.0 * qm1 - qm2;
    a0 = 0.1 * (1.0 + sqr(ind2));
    a1 = 0.6 * (1.0 + sqr(ind1));
    a2 = 0.3 * (1.0 + sqr(ind0));
    asum = 6.0 * (a0 + a1 + a2);
    *lo = (f0 * a0 + f1 * a1 + f2 * a2) / asum;
}

/* fluxlimiter.jl:74-180  PPM (AthenaPK form) -> (lower face, upper face) */
ORC_INLINE void orc_ppm(double q_im2, double q_im1, double q_i, double q_ip1, double q_ip2,
                        double *lo, double *hi)
{
    const double C2 = ORC_MUT == ORC_MUT_PPM_C2 ? 1.0 : 1.25;
    double qa = q_i - q_im1;
    double qb = q_ip1 - q_i;
    double dd_im1 = 0.5 * qa + 0.5 * (q_im1 - q_im2);
    double dd = 0.5 * qb + 0.5 * qa;
    double dd_ip1 = 0.5 * (q_ip2 - q_ip1) + 0.5 * qb;

    double dph = 0.5 * (q_im1 + q_i) + (dd_im1 - dd) / 6.0;
    double dph_ip1 = 0.5 * (q_i + q_ip1) + (dd - dd_ip1) / 6.0;

    double d2qc_im1 = q_im2 + q_i - 2.0 * q_im1;
    double d2qc = q_im1 + q_ip1 - 2.0 * q_i;
    double d2qc_ip1 = q_i + q_ip2 - 2.0 * q_ip1;

    /* i-1/2 */
    double qa_tmp = dph - q_im1;
    double qb_tmp = q_i - dph;
    qa = 3.0 * (q_im1 + q_i - 2.0 * dph);
    qb = d2qc_im1;
    double qc = d2qc;
    double qd = 0.0;
    if (jl_sign(qa) == jl_sign(qb) && jl_sign(qa) == jl_sign(qc))
        qd = jl_sign(qa) * jl_min(C2 * fabs(qb), jl_min(C2 * fabs(qc), fabs(qa)));
    double dph_tmp = 0.5 * (q_im1 + q_i) - qd / 6.0;
    if (qa_tmp * qb_tmp < 0.0)
        dph = dph_tmp;

    /* i+1/2 */
    qa_tmp = dph_ip1 - q_i;
    qb_tmp = q_ip1 - dph_ip1;
    qa = 3.0 * (q_i + q_ip1 - 2.0 * dph_ip1);
    qb = d2qc;
    qc = d2qc_ip1;
    qd = 0.0;
    if (jl_sign(qa) == jl_sign(qb) && jl_sign(qa) == jl_sign(qc))
        qd = jl_sign(qa) * jl_min(C2 * fabs(qb), jl_min(C2 * fabs(qc), fabs(qa)));
    double dphip1_tmp = 0.5 * (q_i + q_ip1) - qd / 6.0;
    if (qa_tmp * qb_tmp < 0.0)
        dph_ip1 = dphip1_tmp;

    double d2qf = 6.0 * (dph + dph_ip1 - 2.0 * q_i);
    double qr_i = dph;
    double ql_ip1 = dph_ip1;

    double dqf_minus = q_i - qr_i;
    double dqf_plus = ql_ip1 - q_i;

    qa_tmp = dqf_minus * dqf_plus;
    qb_tmp = (q_ip1 - q_i) * (q_i - q_im1);

    qa = d2qc_im1;
    qb = d2qc;
    qc = d2qc_ip1;
    qd = d2qf;
    double qe = 0.0;
    if (jl_sign(qa) == jl_sign(qb) && jl_sign(qa) == jl_sign(qc) && jl_sign(qa) == jl_sign(qd))
        qe = jl_sign(qd) * jl_min(jl_min(C2 * fabs(qa), C2 * fabs(qb)), jl_min(C2 * fabs(qc), fabs(qd)));

    qa = jl_max(fabs(q_im1), fabs(q_im2));
    qb = jl_max(jl_max(fabs(q_i), fabs(q_ip1)), fabs(q_ip2));

    double rho = 0.0;
    if (fabs(qd) > 1.0e-12 * jl_max(qa, qb))
        rho = qe / qd;

    double tmp_m = q_i - rho * dqf_minus;
    double tmp_p = q_i + rho * dqf_plus;
    double tmp2_m = q_i - 2.0 * dqf_plus;
    double tmp2_p = q_i + 2.0 * dqf_minus;

    if (ORC_MUT == ORC_MUT_PPM_EXT ? (qa_tmp <= 0.0 && qb_tmp <= 0.0) : (qa_tmp <= 0.0 || qb_tmp <= 0.0)) {
        if (rho <= 1.0 - 1.0e-12) {
            qr_i = tmp_m;
            ql_ip1 = tmp_p;
        }
    } else {
        if (fabs(dqf_minus) >= 2.0 * fabs(dqf_plus))
            qr_i = tmp2_m;
        if (fabs(dqf_plus) >= 2.0 * fabs(dqf_minus))
            ql_ip1 = tmp2_p;
    }
    *lo = qr_i;
    *hi = ql_ip1;
}

/* Reconstruction selector: 0 = MINMOD, 1 = PPM, 2 = WENOZ
 * (fluxlimiter.jl:183-200 dispatches the same three by name). */
enum { ORC_REC_MINMOD = 0, ORC_REC_PPM = 1, ORC_REC_WENOZ = 2 };

ORC_INLINE void orc_recon(int rec, double qm2, double qm1, double q, double qp1, double qp2,
                          double *lo, double *hi)
{
    if (rec == ORC_REC_MINMOD)
        orc_minmod(qm1, q, qp1, lo, hi);
    else if (rec == ORC_REC_PPM)
        orc_ppm(qm2, qm1, q, qp1, qp2, lo, hi);
    else
        orc_wenoz(qm2, qm1, q, qp1, qp2, lo, hi);
}

/* ---------------- 3D point conversions (3dim/structs.jl:84-142) ---------------- */

ORC_INLINE void orc3_p2u(const double *P, double *U, double g)
{
    double gam = sqrt(P[2] * P[2] + P[3] * P[3] + P[4] * P[4] + 1.0);
    double w = g * P[1] + P[0];
    U[0] = gam * P[0];
    U[1] = (g - 1.0) * P[1] - gam * gam * w;
    U[2] = P[2] * gam * w;
    U[3] = P[3] * gam * w;
    U[4] = P[4] * gam * w;
}

/* function_PtoFx / Fy / Fz: dim 0,1,2 */
ORC_INLINE void orc3_p2f(const double *P, double *F, double g, int dim)
{
    double w = g * P[1] + P[0];
    double gam = sqrt(P[2] * P[2] + P[3] * P[3] + P[4] * P[4] + 1.0);
    if (dim == 0) {
        F[0] = P[0] * P[2];
        F[1] = -w * P[2] * gam;
        F[2] = P[2] * P[2] * w + (g - 1.0) * P[1];
        F[3] = P[2] * P[3] * w;
        F[4] = P[2] * P[4] * w;
    } else if (dim == 1) {
        F[0] = P[0] * P[3];
        F[1] = -w * P[3] * gam;
        F[2] = P[2] * P[3] * w;
        F[3] = P[3] * P[3] * w + (g - 1.0) * P[1];
        F[4] = P[4] * P[3] * w;
    } else {
        F[0] = P[0] * P[4];
        F[1] = -w * P[4] * gam;
        F[2] = P[2] * P[4] * w;
        F[3] = P[3] * P[4] * w;
        F[4] = w * (P[4] * P[4]) + (g - 1.0) * P[1];
    }
}

/* ---------------- 2D point conversions (2dim/structs.jl:79-119) ---------------- */

ORC_INLINE void orc2_p2u(const double *P, double *U, double g)
{
    double gam = sqrt(P[2] * P[2] + P[3] * P[3] + 1.0);
    double w = g * P[1] + P[0];
    U[0] = gam * P[0];
    U[1] = (g - 1.0) * P[1] - gam * gam * w;
    U[2] = P[2] * gam * w;
    U[3] = P[3] * gam * w;
}

ORC_INLINE void orc2_p2f(const double *P, double *F, double g, int dim)
{
    double gam = sqrt(P[2] * P[2] + P[3] * P[3] + 1.0);
    double w = g * P[1] + P[0];
    if (dim == 0) {
        F[0] = P[0] * P[2];
        F[1] = -w * P[2] * gam;
        F[2] = P[2] * P[2] * w + (g - 1.0) * P[1];
        F[3] = P[2] * P[3] * w;
    } else {
        F[0] = P[0] * P[3];
        F[1] = -w * P[3] * gam;
        F[2] = P[2] * P[3] * w;
        F[3] = P[3] * P[3] * w + (g - 1.0) * P[1];
    }
}

/* ---------------- 1D point conversions (1dim/structs.jl:58-76) ---------------- */

ORC_INLINE void orc1_p2u(const double *P, double *U, double g)
{
    double gamma = sqrt(P[2] * P[2] + 1.0);
    double pressure = (g - 1.0) * P[1];
    U[0] = P[0] * gamma;
    U[1] = -(P[1] + pressure + P[0]) * (gamma * gamma) + pressure;
    U[2] = (P[1] + pressure + P[0]) * gamma * P[2];
}

ORC_INLINE void orc1_p2f(const double *P, double *F, double g)
{
    double gamma = sqrt(P[2] * P[2] + 1.0);
    double pressure = (g - 1.0) * P[1];
    F[0] = P[0] * P[2];
    F[1] = -(P[1] + pressure + P[0]) * gamma * P[2];
    F[2] = pressure + (P[1] + pressure + P[0]) * (P[2] * P[2]);
}

/* ---------------- wave speeds ----------------
 * 3dim/algos.jl:130-157 (2D: 2dim/algos.jl:125-149):
 *   sigma = c^2/(lor^2 (1-c^2)),  lam+- = (v +- sqrt(sigma (1 - v^2 + sigma)))/(1 + sigma)
 * Returns lam+ in *lp and lam- in *lm.  lor2 is the caller's Lorentz-factor
 * square expression (3D/2D: lor*lor; 1D: ux^2 + 1, 1dim/algos.jl:133). */
ORC_INLINE void orc_lambdas(double v, double cs, double lor2, double *lp, double *lm)
{
    double sig = cs * cs / (lor2 * (1.0 - cs * cs));
    double rt = sqrt(sig * (1.0 - v * v + sig));
    *lp = (v + rt) / (1.0 + sig);
    *lm = (v - rt) / (1.0 + sig);
}

/* HLL flux (3dim/algos.jl:158-174): branch on Cmax<0 / Cmin<0, otherwise
 * (FR*Cmin + FL*Cmax - Cmax*Cmin*(UR-UL))/(Cmax+Cmin). */
ORC_INLINE void orc_hll(int nv, const double *UL, const double *UR, const double *FL,
                        const double *FR, double Cmax, double Cmin, double *F)
{
    if (Cmax < 0.0) {
        for (int v = 0; v < nv; v++) F[v] = FR[v];
    } else if (Cmin < 0.0) {
        for (int v = 0; v < nv; v++) F[v] = FL[v];
    } else {
        for (int v = 0; v < nv; v++)
            F[v] = (FR[v] * Cmin + FL[v] * Cmax - Cmax * Cmin * (UR[v] - UL[v])) / (Cmax + Cmin);
    }
}

/*
 * HLLC flux with HLL fallback -- no reference counterpart (Verona ships HLL
 * only).  Restated from Mignone & Bodo (2005, MNRAS 364, 126), written in
 * Verona's variables, where U[1] = T^t_t = -E and F[1] = -m_d:
 *   HLL state/flux with S_L = -Cmin, S_R = Cmax,
 *   contact speed: F_E lam*^2 - (E + F_m) lam* + m = 0 (hll quantities),
 *     taken as lam* = 2m / ((E+F_m) + sqrt((E+F_m)^2 - 4 F_E m)),
 *   p* = F_m - lam* F_E,
 *   star state on side K: D* = D(S-v)/(S-lam*), m_d* = (m_d(S-v) + p* - p)/(S-lam*),
 *     m_t* = m_t(S-v)/(S-lam*), E* = (E(S-v) + p* lam* - p v)/(S-lam*),
 *   F*_K = F_K + S_K (U*_K - U_K).
 * Falls back to the HLL formula when the discriminant is negative, lam* is
 * not strictly inside (S_L, S_R), p* <= 0, or the chosen star state has
 * D* <= 0 or E* <= 0.  `nd` is the index of the normal momentum (2 + dim).
 */
ORC_INLINE void orc_hllc(int nv, int nd, const double *UL, const double *UR, const double *FL,
                         const double *FR, double vL, double vR, double pL, double pR,
                         double Cmax, double Cmin, double *F)
{
    if (Cmax < 0.0) {
        for (int v = 0; v < nv; v++) F[v] = FR[v];
        return;
    }
    if (Cmin < 0.0) {
        for (int v = 0; v < nv; v++) F[v] = FL[v];
        return;
    }
    double SL = -Cmin, SR = Cmax;
    double den = Cmax + Cmin;
    double inv = 1.0 / den;
    /* HLL quantities needed for the contact speed */
    double E_h = -(Cmax * UR[1] + Cmin * UL[1] + FL[1] - FR[1]) * inv;
    double m_h = (Cmax * UR[nd] + Cmin * UL[nd] + FL[nd] - FR[nd]) * inv;
    double FE_h = -(FR[1] * Cmin + FL[1] * Cmax - Cmax * Cmin * (UR[1] - UL[1])) * inv;
    double Fm_h = (FR[nd] * Cmin + FL[nd] * Cmax - Cmax * Cmin * (UR[nd] - UL[nd])) * inv;
    double b = E_h + Fm_h;
    double disc = b * b - 4.0 * FE_h * m_h;
    int ok = disc >= 0.0;
    double lam = 0.0, ps = 0.0;
    if (ok) {
        double q = b + sqrt(disc);
        lam = 2.0 * m_h / q;
        ps = Fm_h - lam * FE_h;
        ok = isfinite(lam) && lam > SL && lam < SR && ps > 0.0;
    }
    if (ok) {
        const double *U_K = lam >= 0.0 ? UL : UR;
        const double *F_K = lam >= 0.0 ? FL : FR;
        double S = lam >= 0.0 ? SL : SR;
        double vK = lam >= 0.0 ? vL : vR;
        double pK = lam >= 0.0 ? pL : pR;
        double fac = 1.0 / (S - lam);
        double Us[5];
        double sv = S - vK;
        Us[0] = U_K[0] * sv * fac;
        double E_K = -U_K[1];
        double Es = (E_K * sv + ps * lam - pK * vK) * fac;
        Us[1] = -Es;
        for (int v = 2; v < nv; v++) {
            if (v == nd)
                Us[v] = (U_K[v] * sv + ps - pK) * fac;
            else
                Us[v] = U_K[v] * sv * fac;
        }
        if (Us[0] > 0.0 && Es > 0.0) {
            for (int v = 0; v < nv; v++) F[v] = F_K[v] + S * (Us[v] - U_K[v]);
            return;
        }
    }
    orc_hll(nv, UL, UR, FL, FR, Cmax, Cmin, F);
}

#endif /* ORC_CORE_H */
