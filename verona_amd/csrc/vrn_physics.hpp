// vrn_physics.hpp -- device point functions of the SRHD hot path (gfx950).
//
// Every function follows the operation order of the Julia reference so that
// an FP64 build with -ffp-contract=off reproduces it to the last bit where
// the reference itself is IEEE-exact (no contraction, correctly rounded
// division and square root).  Citations are to Wesenheit/Verona 0.1.0.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "vrn_sqrt.hpp"

// VRN_NS: namespace of this instantiation.  With VRN_FAST the same functions
// are built as the fast FP64 form: a*b+c contracted to FMA (the including
// file sets the contraction pragma) and fdiv() = reciprocal refined by one
// Newton step times the numerator with one residual correction (within an
// ulp of the correctly rounded quotient), used only where the divisor is
// bounded away from zero and from the subnormal range.
#ifndef VRN_NS
#define VRN_NS vrn
#endif

namespace VRN_NS {

// Julia Base max/min for IEEE floats: NaN-propagating, -0 < +0.
template <class T> __host__ __device__ __forceinline__ T jmax(T x, T y)
{
    T d = x - y;
    T am = __builtin_signbit(d) ? y : x;
    return (__builtin_isnan(x) || __builtin_isnan(y)) ? d : am;
}
template <class T> __host__ __device__ __forceinline__ T jmin(T x, T y)
{
    T d = x - y;
    T am = __builtin_signbit(d) ? x : y;
    return (__builtin_isnan(x) || __builtin_isnan(y)) ? d : am;
}
// Julia max(c, x) / min(x, c) against a bound c that is neither NaN nor -0
// (floors, caps): one compare and a select.  Same value as jmax/jmin for
// every x; a NaN x comes back as itself (jmax returns c - x, also NaN).
template <class T> __host__ __device__ __forceinline__ T floor_max(T c, T x) { return x <= c ? c : x; }
template <class T> __host__ __device__ __forceinline__ T cap_min(T x, T c) { return x >= c ? c : x; }
// Julia sign(): -1, +1, or x itself for +-0 and NaN.
template <class T> __device__ __forceinline__ T jsign(T x)
{
    return x < T(0) ? T(-1) : (x > T(0) ? T(1) : x);
}
template <class T> __device__ __forceinline__ T sq(T x) { return x * x; }

// PPM's limiter tests, written with plain compares.  same_sign3/4 equal
// Julia's sign(a) == sign(b) == ...: true when all arguments are < 0, all
// > 0 or all == 0 (+-0), false as soon as one is NaN.  nn_min / nn_max are
// Julia's min / max for arguments that are not NaN and not -0 (the callers
// pass absolute values, scaled or not), where a compare-select is exact.
template <class T> __device__ __forceinline__ bool same_sign3(T a, T b, T c)
{
    return ((a < T(0)) & (b < T(0)) & (c < T(0))) | ((a > T(0)) & (b > T(0)) & (c > T(0))) |
           ((a == T(0)) & (b == T(0)) & (c == T(0)));
}
template <class T> __device__ __forceinline__ bool same_sign4(T a, T b, T c, T d)
{
    return ((a < T(0)) & (b < T(0)) & (c < T(0)) & (d < T(0))) |
           ((a > T(0)) & (b > T(0)) & (c > T(0)) & (d > T(0))) |
           ((a == T(0)) & (b == T(0)) & (c == T(0)) & (d == T(0)));
}
template <class T> __device__ __forceinline__ T nn_min(T x, T y) { return y < x ? y : x; }
template <class T> __device__ __forceinline__ T nn_max(T x, T y) { return y > x ? y : x; }

__host__ __device__ __forceinline__ double dsqrt(double x) { return __builtin_sqrt(x); }
__host__ __device__ __forceinline__ float dsqrt(float x) { return __builtin_sqrtf(x); }
__host__ __device__ __forceinline__ double dabs(double x) { return __builtin_fabs(x); }
__host__ __device__ __forceinline__ float dabs(float x) { return __builtin_fabsf(x); }

#if defined(VRN_FAST)
__host__ __device__ __forceinline__ double fdiv(double a, double b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    // v_rcp_f64 is good to ~2^-24 (2^-24.4 at most over 2^22 random b on
    // gfx950); one Newton step leaves |1 - b r| ~ 2^-48,
    // and the residual correction of the quotient squares that again, so
    // the result is within an ulp (a second step on r would change nothing)
    double r = __builtin_amdgcn_rcp(b);
    r = __builtin_fma(r, __builtin_fma(-b, r, 1.0), r);
    double q = a * r;
    double rem = __builtin_fma(-b, q, a);
    return __builtin_fma(rem, r, q);
#else
    return a / b;
#endif
}
// 1/b: reciprocal refined by two Newton steps (about an ulp), for b bounded
// away from zero and from the subnormal range.
__host__ __device__ __forceinline__ double frcp(double b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rcp(b);
    r = __builtin_fma(r, __builtin_fma(-b, r, 1.0), r);
    return __builtin_fma(r, __builtin_fma(-b, r, 1.0), r);
#else
    return 1.0 / b;
#endif
}
// 1/b with one Newton step (max 11 ulp measured over 2^22 random b on
// gfx950; 59 % correctly rounded), for b bounded away from zero and from
// the subnormal range.
__host__ __device__ __forceinline__ double frcp1(double b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const double r = __builtin_amdgcn_rcp(b);
    return __builtin_fma(r, __builtin_fma(-b, r, 1.0), r);
#else
    return 1.0 / b;
#endif
}
#else
__host__ __device__ __forceinline__ double fdiv(double a, double b) { return a / b; }
#endif
// FP32 form: reciprocal instruction times the numerator plus one residual
// correction, so quotients that are representable come out exact (WENO-Z
// returns constant data unchanged) at 4 instructions instead of IEEE
// division's 10.
__host__ __device__ __forceinline__ float fdiv(float a, float b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const float r = __builtin_amdgcn_rcpf(b);
    const float q = a * r;
    return __builtin_fmaf(__builtin_fmaf(-b, q, a), r, q);
#else
    return a / b;
#endif
}

// 1/x: refined reciprocal in the fast FP64 form, IEEE division otherwise
// (FP32: the reciprocal-plus-correction of fdiv).
#if defined(VRN_FAST)
__host__ __device__ __forceinline__ double recip(double x) { return frcp(x); }
#else
__host__ __device__ __forceinline__ double recip(double x) { return 1.0 / x; }
#endif
__host__ __device__ __forceinline__ float recip(float x) { return fdiv(1.0f, x); }
// Reciprocals of the face states' wave-speed terms and of the HLL/HLLC
// denominators: one Newton step in the fast FP64 form.
#if defined(VRN_FAST)
__host__ __device__ __forceinline__ double recip1(double x) { return frcp1(x); }
#else
__host__ __device__ __forceinline__ double recip1(double x) { return recip(x); }
#endif
__host__ __device__ __forceinline__ float recip1(float x) { return recip(x); }

enum { REC_MINMOD = 0, REC_PPM = 1, REC_WENOZ = 2 };
enum { RIEM_HLL = 0, RIEM_HLLC = 1 };

// WENO-Z epsilon: 1e-42 in FP64 (fluxlimiter.jl:56).  1e-42 is subnormal in
// FP32 and flushes to zero there, so the FP32 form uses 1e-30f instead.
template <class T> struct WenoEps;
template <> struct WenoEps<double> { static constexpr double v = 1.0e-42; };
template <> struct WenoEps<float> { static constexpr float v = 1.0e-30f; };

// WENO-Z curvature term 13/12 (a - 2 b + c)^2 of three consecutive cells,
// evaluated as (a + c) - 2 b: the smoothness indicators of the stencils
// centred on cells i-1, i, i+1 (fluxlimiter.jl:45-50) use it with the same
// three cells in the same order up to the commutative first sum, so one
// value serves all three positions bit for bit (the marching sweeps compute
// it once per cell and carry it).
#if defined(VRN_FAST)
// Fast FP64 form: every smoothness indicator is carried scaled by 12/13 with
// the equally scaled epsilon added, e = (12/13)(beta + 1e-42) =
// c^2 + (3/13) l^2 + WZ_EPS12 (c the curvature second difference, l the
// first-difference combination of the stencil), so the curvature term is
// c^2 + WZ_EPS12 and the ratios tau5 / (beta + eps) are unchanged.
constexpr double WZ_EPS12 = 12.0 / 13.0 * 1.0e-42;
#endif
template <class T> __device__ __forceinline__ T wz_curv(T a, T b, T c)
{
#if defined(VRN_FAST)
    if constexpr (std::is_same<T, double>::value) {
        const double d = a + c - 2.0 * b;
        return d * d + WZ_EPS12;
    }
#endif
    return T(13.0) / T(12.0) * sq(a + c - T(2) * b);
}

#if defined(VRN_FAST)
// Fast FP64 WENO-Z from the scaled indicators above (fluxlimiter.jl:19-71
// rearranged; equal to it up to rounding):
//  * e_k = (3/13) l_k^2 + k_k, tau = |e_0 - e_2|; one reciprocal of
//    e_0 e_1 e_2 (>= WZ_EPS12^3 ~ 8e-127, normal) gives the three ratios
//    tau / e_k, and s_k = 1 + (tau / e_k)^2;
//  * the six stencil polynomials share their terms: with A = q_{i-1} + 5 q_i
//    and B = q_{i+1} + 5 q_i, f0 = 2 l_0 + A, f2 = B - l_2 (face i+1/2) and
//    f0' = 2 l_2 + B, f2' = A - l_0, f1' = f1 + 3 l_1 (face i-1/2);
//  * the linear weights (1, 6, 3)/10 enter both numerator and denominator,
//    and both faces are normalised with one reciprocal of the product of
//    their denominators (each >= 10, finite unless a stencil jump exceeds
//    ~1e17).
__device__ __forceinline__ void wenoz_fast(double qm2, double qm1, double q, double qp1, double qp2, double k0,
                                           double k1, double k2, double &lo, double &hi)
{
    constexpr double c313 = 3.0 / 13.0;
    const double l0 = (qm2 + 3.0 * q) - 4.0 * qm1;
    const double l1 = qm1 - qp1;
    const double l2 = (qp2 + 3.0 * q) - 4.0 * qp1;
    const double e0 = c313 * (l0 * l0) + k0;
    const double e1 = c313 * (l1 * l1) + k1;
    const double e2 = c313 * (l2 * l2) + k2;
    const double tau = dabs(e0 - e2);
    const double e01 = e0 * e1;
    const double r = tau * frcp1(e01 * e2);
    const double i0 = r * (e1 * e2), i1 = r * (e0 * e2), i2 = r * e01;
    const double s0 = i0 * i0 + 1.0, s1 = i1 * i1 + 1.0, s2 = i2 * i2 + 1.0;
    const double q5 = 5.0 * q;
    const double A = q5 + qm1, B = q5 + qp1;
    const double f0h = 2.0 * l0 + A;
    const double f1h = 2.0 * qp1 + (q5 - qm1);
    const double f2h = B - l2;
    const double f0l = 2.0 * l2 + B;
    const double f1l = 3.0 * l1 + f1h;
    const double f2l = A - l0;
    const double z1 = 6.0 * s1, z2 = 3.0 * s2, z0 = 3.0 * s0;
    const double nh = s0 * f0h + z1 * f1h + z2 * f2h, dh = s0 + z1 + z2;
    const double nl = s2 * f0l + z1 * f1l + z0 * f2l, dl = s2 + z1 + z0;
    const double r6 = frcp1(dh * dl) * (1.0 / 6.0);
    hi = nh * (dl * r6);
    lo = nl * (dh * r6);
}
#endif

// fluxlimiter.jl:19-71  WENOZ(q_{i-2..i+2}) -> lo = q at i-1/2, hi = q at i+1/2,
// from the five values and the curvature terms k0 = wz_curv(q_{i-2}, q_{i-1},
// q_i), k1 = wz_curv(q_{i-1}, q_i, q_{i+1}), k2 = wz_curv(q_i, q_{i+1}, q_{i+2}).
template <class T>
__device__ __forceinline__ void wenoz_k(T qm2, T qm1, T q, T qp1, T qp2, T k0, T k1, T k2, T &lo, T &hi)
{
#if defined(VRN_FAST)
    if constexpr (std::is_same<T, double>::value) {
        wenoz_fast(qm2, qm1, q, qp1, qp2, k0, k1, k2, lo, hi);
        return;
    }
#endif
    const T b2c = T(0.25);
    T beta0 = k0 + b2c * sq(qm2 + T(3) * q - T(4) * qm1);
    T beta1 = k1 + b2c * sq(qm1 - qp1);
    T beta2 = k2 + b2c * sq(qp2 + T(3) * q - T(4) * qp1);
    const T eps = WenoEps<T>::v;
    T tau5 = dabs(beta0 - beta2);
    T ind0 = fdiv(tau5, beta0 + eps);
    T ind1 = fdiv(tau5, beta1 + eps);
    T ind2 = fdiv(tau5, beta2 + eps);
    T s0, s1, s2;
    if constexpr (std::is_same<T, float>::value) {
        // FP32 form: ind^2 overflows FP32 once ind > ~1.8e19 (a flat side next
        // to a small jump gives tau5 / eps that large), and inf/inf would give
        // NaN.  The weights enter only as ratios, so all three are scaled by a
        // common factor 1/ind_max^2 when ind_max > 1e15.
        T im = ind0 > ind1 ? ind0 : ind1;
        im = im > ind2 ? im : ind2;
        T sc = im > T(1e15) ? T(1) / im : T(1);
        T sc2 = sc * sc;
        s0 = sc2 + sq(ind0 * sc);
        s1 = sc2 + sq(ind1 * sc);
        s2 = sc2 + sq(ind2 * sc);
    } else {
        s0 = T(1) + sq(ind0);
        s1 = T(1) + sq(ind1);
        s2 = T(1) + sq(ind2);
    }

    T a0 = T(0.1) * s0, a1 = T(0.6) * s1, a2 = T(0.3) * s2;
    T f0 = T(2) * qm2 - T(7) * qm1 + T(11) * q;
    T f1 = -qm1 + T(5) * q + T(2) * qp1;
    T f2 = T(2) * q + T(5) * qp1 - qp2;
    const T nh = f0 * a0 + f1 * a1 + f2 * a2, dh = T(6) * (a0 + a1 + a2);
    hi = fdiv(nh, dh);

    a0 = T(0.1) * s2;
    a2 = T(0.3) * s0;
    f0 = T(2) * qp2 - T(7) * qp1 + T(11) * q;
    f1 = -qp1 + T(5) * q + T(2) * qm1;
    f2 = T(2) * q + T(5) * qm1 - qm2;
    const T nl = f0 * a0 + f1 * a1 + f2 * a2, dl = T(6) * (a0 + a1 + a2);
    lo = fdiv(nl, dl);
}

template <class T>
__device__ __forceinline__ void wenoz(T qm2, T qm1, T q, T qp1, T qp2, T &lo, T &hi)
{
    wenoz_k(qm2, qm1, q, qp1, qp2, wz_curv(qm2, qm1, q), wz_curv(qm1, q, qp1), wz_curv(q, qp1, qp2), lo, hi);
}

// fluxlimiter.jl:1-13  MINMOD(q_{i-1}, q_i, q_{i+1})
template <class T> __device__ __forceinline__ void minmod(T qm1, T q, T qp1, T &lo, T &hi)
{
    T sp = qp1 - q, sm = q - qm1;
    T dU = T(0.25) * (jsign(sp) + jsign(sm)) * jmin(dabs(sp), dabs(sm));
    lo = q - dU;
    hi = q + dU;
}

// x / 6.  Fast FP64 form: x (1/6) plus one remainder correction (Markstein:
// with 1/6 rounded to nearest and a first quotient within an ulp, the
// corrected quotient is the correctly rounded x / 6 away from the subnormal
// range); otherwise the IEEE division.
template <class T> __device__ __forceinline__ T div6(T x)
{
#if defined(VRN_FAST)
    if constexpr (std::is_same<T, double>::value) {
        constexpr double c = 1.0 / 6.0;
        const double q = x * c;
        return __builtin_fma(__builtin_fma(-q, 6.0, x), c, q);
    }
#endif
    return x / T(6);
}

// fluxlimiter.jl:74-180  PPM (AthenaPK form)
template <class T>
__device__ __forceinline__ void ppm(T q_im2, T q_im1, T q_i, T q_ip1, T q_ip2, T &lo, T &hi)
{
    const T C2 = T(1.25);
    T qa = q_i - q_im1;
    T qb = q_ip1 - q_i;
    T dd_im1 = T(0.5) * qa + T(0.5) * (q_im1 - q_im2);
    T dd = T(0.5) * qb + T(0.5) * qa;
    T dd_ip1 = T(0.5) * (q_ip2 - q_ip1) + T(0.5) * qb;
    T dph = T(0.5) * (q_im1 + q_i) + div6(dd_im1 - dd);
    T dph_ip1 = T(0.5) * (q_i + q_ip1) + div6(dd - dd_ip1);
    T d2qc_im1 = q_im2 + q_i - T(2) * q_im1;
    T d2qc = q_im1 + q_ip1 - T(2) * q_i;
    T d2qc_ip1 = q_i + q_ip2 - T(2) * q_ip1;

    T qa_tmp = dph - q_im1;
    T qb_tmp = q_i - dph;
    qa = T(3) * (q_im1 + q_i - T(2) * dph);
    qb = d2qc_im1;
    T qc = d2qc;
    T qd = T(0);
    // sign(qa) * min(...) with equal signs: the min is +0 exactly when qa is
    // +-0, so the product is copysign(min, qa)
    if (same_sign3(qa, qb, qc))
        qd = __builtin_copysign(nn_min(C2 * dabs(qb), nn_min(C2 * dabs(qc), dabs(qa))), qa);
    T dph_tmp = T(0.5) * (q_im1 + q_i) - div6(qd);
    if (qa_tmp * qb_tmp < T(0)) dph = dph_tmp;

    qa_tmp = dph_ip1 - q_i;
    qb_tmp = q_ip1 - dph_ip1;
    qa = T(3) * (q_i + q_ip1 - T(2) * dph_ip1);
    qb = d2qc;
    qc = d2qc_ip1;
    qd = T(0);
    if (same_sign3(qa, qb, qc))
        qd = __builtin_copysign(nn_min(C2 * dabs(qb), nn_min(C2 * dabs(qc), dabs(qa))), qa);
    T dphip1_tmp = T(0.5) * (q_i + q_ip1) - div6(qd);
    if (qa_tmp * qb_tmp < T(0)) dph_ip1 = dphip1_tmp;

    T d2qf = T(6) * (dph + dph_ip1 - T(2) * q_i);
    T qr_i = dph, ql_ip1 = dph_ip1;
    T dqf_minus = q_i - qr_i;
    T dqf_plus = ql_ip1 - q_i;
    qa_tmp = dqf_minus * dqf_plus;
    qb_tmp = (q_ip1 - q_i) * (q_i - q_im1);
    qa = d2qc_im1;
    qb = d2qc;
    qc = d2qc_ip1;
    qd = d2qf;
    T qe = T(0);
    if (same_sign4(qa, qb, qc, qd))
        qe = __builtin_copysign(nn_min(nn_min(C2 * dabs(qa), C2 * dabs(qb)), nn_min(C2 * dabs(qc), dabs(qd))), qd);
    // a NaN among the five inputs makes qd = d2qf NaN, so the test below is
    // false whatever the maximum: Julia's NaN-propagating max is not needed
    qa = nn_max(dabs(q_im1), dabs(q_im2));
    qb = nn_max(nn_max(dabs(q_i), dabs(q_ip1)), dabs(q_ip2));
    T rho = T(0);
    if (dabs(qd) > T(1.0e-12) * nn_max(qa, qb)) rho = qe / qd;
    T tmp_m = q_i - rho * dqf_minus;
    T tmp_p = q_i + rho * dqf_plus;
    T tmp2_m = q_i - T(2) * dqf_plus;
    T tmp2_p = q_i + T(2) * dqf_minus;
    if (qa_tmp <= T(0) || qb_tmp <= T(0)) {
        if (rho <= T(1) - T(1.0e-12)) {
            qr_i = tmp_m;
            ql_ip1 = tmp_p;
        }
    } else {
        if (dabs(dqf_minus) >= T(2) * dabs(dqf_plus)) qr_i = tmp2_m;
        if (dabs(dqf_plus) >= T(2) * dabs(dqf_minus)) ql_ip1 = tmp2_p;
    }
    lo = qr_i;
    hi = ql_ip1;
}

template <int REC, class T>
__device__ __forceinline__ void recon(T qm2, T qm1, T q, T qp1, T qp2, T &lo, T &hi)
{
    if constexpr (REC == REC_WENOZ)
        wenoz(qm2, qm1, q, qp1, qp2, lo, hi);
    else if constexpr (REC == REC_PPM)
        ppm(qm2, qm1, q, qp1, qp2, lo, hi);
    else
        minmod(qm1, q, qp1, lo, hi);
}

// ---------------------------------------------------------------- conversions
// 3dim/structs.jl:84-142 (NV = 5) and 2dim/structs.jl:79-119 (NV = 4):
// lor = sqrt(sum u_i^2 + 1), w = gamma*u + rho,
// U = (lor rho, (gamma-1)u - lor^2 w, u_i lor w), F^d as function_PtoF{x,y,z}.
template <int NV, class T> __host__ __device__ __forceinline__ T lorentz(const T *P)
{
    if constexpr (NV == 5)
        return vrn_sqrt1(P[2] * P[2] + P[3] * P[3] + P[4] * P[4] + T(1));
    else
        return vrn_sqrt1(P[2] * P[2] + P[3] * P[3] + T(1));
}

template <int NV, class T> __device__ __forceinline__ void p2u(const T *P, T *U, T g, T gam)
{
    T w = g * P[1] + P[0];
    U[0] = gam * P[0];
    U[1] = (g - T(1)) * P[1] - gam * gam * w;
    U[2] = P[2] * gam * w;
    U[3] = P[3] * gam * w;
    if constexpr (NV == 5) U[4] = P[4] * gam * w;
}

template <int NV, int DIM, class T>
__device__ __forceinline__ void p2f(const T *P, T *F, T g, T gam)
{
    T w = g * P[1] + P[0];
    T pr = (g - T(1)) * P[1];
    if constexpr (DIM == 0) {
        F[0] = P[0] * P[2];
        F[1] = -w * P[2] * gam;
        F[2] = P[2] * P[2] * w + pr;
        F[3] = P[2] * P[3] * w;
        if constexpr (NV == 5) F[4] = P[2] * P[4] * w;
    } else if constexpr (DIM == 1) {
        F[0] = P[0] * P[3];
        F[1] = -w * P[3] * gam;
        F[2] = P[2] * P[3] * w;
        F[3] = P[3] * P[3] * w + pr;
        if constexpr (NV == 5) F[4] = P[4] * P[3] * w;
    } else {
        F[0] = P[0] * P[4];
        F[1] = -w * P[4] * gam;
        F[2] = P[2] * P[4] * w;
        F[3] = P[3] * P[4] * w;
        F[4] = w * (P[4] * P[4]) + pr;
    }
}

// eos.jl:15-17
template <class T> __host__ __device__ __forceinline__ T sound_speed(T rho, T u, T g)
{
    return vrn_sqrt(fdiv(g * (g - T(1)) * u, rho + g * u));
}

// 3dim/algos.jl:146-157: sigma = c^2/(lor^2 (1-c^2)),
// lam+- = (v +- sqrt(sigma (1 - v^2 + sigma)))/(1 + sigma)
template <class T>
__host__ __device__ __forceinline__ void lambdas(T v, T cs, T lor2, T &lp, T &lm)
{
    T sig = fdiv(cs * cs, lor2 * (T(1) - cs * cs));
    T rt = vrn_sqrt(sig * (T(1) - v * v + sig));
#if defined(VRN_FAST)
    T inv = fdiv(T(1), T(1) + sig);
    lp = (v + rt) * inv;
    lm = (v - rt) * inv;
#else
    lp = (v + rt) / (T(1) + sig);
    lm = (v - rt) / (T(1) + sig);
#endif
}

#if defined(VRN_FAST)
// Fast FP64 form of one face state's v_d and wave speeds.  With
// c^2 = g(g-1)u/(rho + g u) (eos.jl:15-17) substituted,
//   sigma = c^2/(lor^2 (1 - c^2)) = a/(lor^2 b),  a = g(g-1)u,  b = rho + g(2-g)u,
//   1/(1 + sigma) = d/(d + a),  d = lor^2 b,
// so v = u_d/lor, sigma and 1/(1 + sigma) share one reciprocal of lor d (d + a)
// and the sound speed's square root disappears.  Rounding differs from the
// reference order at the last bits only (lor >= 1, rho and u floored > 0).
template <int DIM>
__device__ __forceinline__ void side_speeds(const double *P, double lor, double g, double &v, double &lp,
                                            double &lm)
{
    const double a = g * (g - 1.0) * P[1];
    const double b = P[0] + g * (2.0 - g) * P[1];
    const double d = lor * lor * b;
    const double s = d + a;
    const double r = recip1(lor * d * s);
    v = P[2 + DIM] * (d * s * r);
    const double sig = a * (lor * s * r);
    const double inv = d * (lor * d * r);
    const double rt = vrn_sqrt(sig * (1.0 - v * v + sig));
    lp = (v + rt) * inv;
    lm = (v - rt) * inv;
}
#endif

// HLL (3dim/algos.jl:158-174); HLLC with HLL fallback (Mignone & Bodo 2005,
// no reference counterpart) written in Verona's variables (U[1] = -E, F[1] = -m_d).
template <int NV, int RIEM, class T>
__device__ __forceinline__ void riemann_solve(const T *UL, const T *UR, const T *FL, const T *FR,
                                              T vL, T vR, T pL, T pR, T Cmax, T Cmin, int nd, T *F)
{
    if (Cmax < T(0)) {
#pragma unroll
        for (int v = 0; v < NV; v++) F[v] = FR[v];
        return;
    }
    if (Cmin < T(0)) {
#pragma unroll
        for (int v = 0; v < NV; v++) F[v] = FL[v];
        return;
    }
    if constexpr (RIEM == RIEM_HLLC) {
        T SL = -Cmin, SR = Cmax;
        T inv = recip1(Cmax + Cmin);
        T E_h = -(Cmax * UR[1] + Cmin * UL[1] + FL[1] - FR[1]) * inv;
        T m_h = (Cmax * UR[nd] + Cmin * UL[nd] + FL[nd] - FR[nd]) * inv;
        T FE_h = -(FR[1] * Cmin + FL[1] * Cmax - Cmax * Cmin * (UR[1] - UL[1])) * inv;
        T Fm_h = (FR[nd] * Cmin + FL[nd] * Cmax - Cmax * Cmin * (UR[nd] - UL[nd])) * inv;
        T b = E_h + Fm_h;
        T disc = b * b - T(4) * FE_h * m_h;
        bool ok = disc >= T(0);
        T lam = T(0), ps = T(0);
        if (ok) {
            T qq = b + vrn_sqrt(disc);
            lam = fdiv(T(2) * m_h, qq);
            ps = Fm_h - lam * FE_h;
            ok = __builtin_isfinite(lam) && lam > SL && lam < SR && ps > T(0);
        }
        if (ok) {
            bool left = lam >= T(0);
            T S = left ? SL : SR;
            T vK = left ? vL : vR;
            T pK = left ? pL : pR;
            T fac = recip1(S - lam);
            T sv = S - vK;
            T Us[NV];
            T UK[NV], FK[NV];
#pragma unroll
            for (int v = 0; v < NV; v++) {
                UK[v] = left ? UL[v] : UR[v];
                FK[v] = left ? FL[v] : FR[v];
            }
            Us[0] = UK[0] * sv * fac;
            T Es = (-UK[1] * sv + ps * lam - pK * vK) * fac;
            Us[1] = -Es;
#pragma unroll
            for (int v = 2; v < NV; v++)
                Us[v] = (v == nd) ? (UK[v] * sv + ps - pK) * fac : UK[v] * sv * fac;
            if (Us[0] > T(0) && Es > T(0)) {
#pragma unroll
                for (int v = 0; v < NV; v++) F[v] = FK[v] + S * (Us[v] - UK[v]);
                return;
            }
        }
    }
#if defined(VRN_FAST)
    const T invs = recip1(Cmax + Cmin);
#pragma unroll
    for (int v = 0; v < NV; v++)
        F[v] = (FR[v] * Cmin + FL[v] * Cmax - Cmax * Cmin * (UR[v] - UL[v])) * invs;
#else
#pragma unroll
    for (int v = 0; v < NV; v++)
        F[v] = (FR[v] * Cmin + FL[v] * Cmax - Cmax * Cmin * (UR[v] - UL[v])) / (Cmax + Cmin);
#endif
}

// Face flux along DIM from floored face states PL (left) and PR (right):
// 3dim/algos.jl:117-174 (NV = 5), 2dim/algos.jl:115-166 (NV = 4).
template <int NV, int DIM, int RIEM, class T>
__device__ __forceinline__ void face_flux(const T *PL, const T *PR, T g, T *F)
{
    T UL[NV], UR[NV], FL[NV], FR[NV];
    T lorL = lorentz<NV>(PL), lorR = lorentz<NV>(PR);
    p2u<NV>(PR, UR, g, lorR);
    p2u<NV>(PL, UL, g, lorL);
    p2f<NV, DIM>(PR, FR, g, lorR);
    p2f<NV, DIM>(PL, FL, g, lorL);
    T vL, vR, lpL, lmL, lpR, lmR;
#if defined(VRN_FAST)
    if constexpr (std::is_same<T, double>::value) {
        side_speeds<DIM>(PL, lorL, g, vL, lpL, lmL);
        side_speeds<DIM>(PR, lorR, g, vR, lpR, lmR);
    } else
#endif
    {
        vL = fdiv(PL[2 + DIM], lorL);
        vR = fdiv(PR[2 + DIM], lorR);
        T CL = sound_speed(PL[0], PL[1], g);
        T CR = sound_speed(PR[0], PR[1], g);
        lambdas(vL, CL, lorL * lorL, lpL, lmL);
        lambdas(vR, CR, lorR * lorR, lpR, lmR);
    }
    T Cmax = jmax(lpL, lpR);
    T Cmin = -jmin(lmL, lmR);
    riemann_solve<NV, RIEM>(UL, UR, FL, FR, vL, vR, (g - T(1)) * PL[1], (g - T(1)) * PR[1], Cmax,
                            Cmin, 2 + DIM, F);
}

// Cell-centred signal-speed rate sum_d max(|lam+_d|, |lam-_d|)/dx_d used by
// the CFL reduction (wave speeds as in 3dim/algos.jl:130-157).
template <int NV>
__host__ __device__ __forceinline__ double cell_rate(const double *P, double g, const double *inv_dx)
{
    double lor = lorentz<NV>(P);
    double r = 0.0;
#if defined(VRN_FAST)
    // sigma and 1/(1 + sigma) do not depend on the direction (algebra as in
    // side_speeds); max(|v + rt|, |v - rt|) = |v| + rt for rt >= 0
    const double a = g * (g - 1.0) * P[1];
    const double b = P[0] + g * (2.0 - g) * P[1];
    const double dd = lor * lor * b;
    const double s = dd + a;
    const double rr = recip(lor * dd * s);
    const double ilor = dd * s * rr, sig = a * (lor * s * rr), inv = dd * (lor * dd * rr);
#pragma unroll
    for (int d = 0; d < NV - 2; d++) {
        const double v = P[2 + d] * ilor;
        const double rt = vrn_sqrt(sig * (1.0 - v * v + sig));
        r = r + ((dabs(v) + rt) * inv) * inv_dx[d];
    }
#else
    double cs = sound_speed(P[0], P[1], g);
#pragma unroll
    for (int d = 0; d < NV - 2; d++) {
        double lp, lm;
        lambdas(fdiv(P[2 + d], lor), cs, lor * lor, lp, lm);
        r = r + jmax(dabs(lp), dabs(lm)) * inv_dx[d];
    }
#endif
    return r;
}

// ---------------------------------------------------------------- con2prim
// 3dim/structs.jl:273-343  1DW Newton (Noble et al. 2006) from the warm-start
// guess Pg.  Returns true and fills P (recovery + floors + Gamma cap,
// :598-609) on convergence; false means the cascade must continue.
__device__ __forceinline__ bool c2p_1dw(const double *U, const double *Pg, double g, int n_iter,
                                        double tol, double *P, int &iters)
{
    const double D = U[0], Qt = U[1], Qx = U[2], Qy = U[3], Qz = U[4];
    const double c = (g - 1.0) / g;
    double S2 = Qx * Qx + Qy * Qy + Qz * Qz;
    double gm = vrn_sqrt1(Pg[2] * Pg[2] + Pg[3] * Pg[3] + Pg[4] * Pg[4] + 1.0);
    double W = (Pg[0] + g * Pg[1]) * (gm * gm);
    const double W_max = 1e30;
    double W_min = vrn_sqrt(S2) * (1.0 + 1e-15);
#if defined(VRN_FAST)
    // S2 / W^2 >= 1 without the division: equal except where the quotient
    // rounds across 1 (then only the starting W differs by the factor 10);
    // S2 = W = 0 gives NaN >= 1, false, in the reference
    for (int t10 = 0; t10 < 1100 && S2 >= W * W && S2 > 0.0 && W < W_max; t10++) W *= 10.0;
#else
    for (int t10 = 0; t10 < 1100 && S2 / (W * W) >= 1.0 && W < W_max; t10++) W *= 10.0; /* unbounded in the reference; W = 0 would never exit */
#endif
    bool conv = false;
    for (int it = 0; it < n_iter; it++) {
        iters++;
        if (W < W_min) W = W_min;
        double v2 = fdiv(S2, W * W);
        if (v2 < 0.0)
            v2 = 0.0;
        else if (v2 > 1.0 - 1e-15)
            v2 = 1.0 - 1e-15;
        double sv = vrn_sqrt(1.0 - v2);
        double fun = Qt + W - c * (W * (1.0 - v2) - D * sv);
        double W3 = W * W * W;
#if defined(VRN_FAST)
        // S2/W3 shared by both Jacobian terms: one reciprocal of W3 sv
        const double q3 = fdiv(S2, W3 * sv);
        double jac = 1.0 - c * (1.0 - v2) - c * (W * (2.0 * (q3 * sv)) - D * q3);
#else
        double jac = 1.0 - c * (1.0 - v2) - c * (W * fdiv(2.0 * S2, W3) - D * fdiv(S2, W3 * sv));
#endif
        double dW = fdiv(fun, jac);
        double Wp = W - dW;
        W = (Wp < W_min) ? 0.5 * (W + W_min) : Wp;
        if (dW < 0.0) dW = -dW;
        if (dW * dW < tol * tol) {
            conv = true;
            break;
        }
    }
    if (!conv) return false;
    double v2 = fdiv(S2, W * W);
#if defined(VRN_FAST)
    // in t = sqrt(1 - v^2) = 1/Gamma, capped below at 1/50 (Gamma <= 50):
    // D/Gamma = D t, W/Gamma^2 - D/Gamma = (W t - D) t, Gamma Q/W = Q/(W t);
    // a NaN t (v^2 > 1) propagates as in the reference
    double t = vrn_sqrt(1.0 - v2);
    t = t < 0.02 ? 0.02 : t;
    P[0] = floor_max(1e-8, D * t);
    P[1] = floor_max(1e-8, (W * t - D) * t * frcp(g));
    const double iW = frcp(W * t);
    P[2] = Qx * iW;
    P[3] = Qy * iW;
    P[4] = Qz * iW;
#else
    double gam = fdiv(1.0, vrn_sqrt(1.0 - v2));
    gam = cap_min(gam, 50.0);
    P[0] = floor_max(1e-8, fdiv(D, gam));
    P[1] = floor_max(1e-8, fdiv(fdiv(W, gam * gam) - fdiv(D, gam), g));
    P[2] = gam * Qx / W;
    P[3] = gam * Qy / W;
    P[4] = gam * Qz / W;
#endif
    return true;
}

// 3dim/structs.jl:346-628: the rest of the cascade after 1DW failed --
// 2D Newton (W, v^2) with backtracking, 5D Newton on primitives, bisection on
// W, recovery, or the NaN marker.  P holds the guess on entry.  Returns the
// method that finished the cell (2, 3, 4) or 0 for failure.
__device__ __noinline__ int c2p_cascade_tail(const double *U, double *P, double g, int n_iter,
                                             double tol, int &iters)
{
    const double D = U[0], Qt = U[1], Qx = U[2], Qy = U[3], Qz = U[4];
    const double c = (g - 1.0) / g;
    double S2 = Qx * Qx + Qy * Qy + Qz * Qz;
    double W, v2;
    bool conv2 = false, conv5 = false, convb = false;
    {
        double gm = dsqrt(P[2] * P[2] + P[3] * P[3] + P[4] * P[4] + 1.0);
        W = (P[0] + g * P[1]) * (gm * gm);
        const double W_max = 1e30;
        for (int t10 = 0; t10 < 1100 && S2 / (W * W) >= 1.0 && W < W_max; t10++) W *= 10.0; /* unbounded in the reference; W = 0 would never exit */
        v2 = jmin(S2 / (W * W), 1.0 - 1e-15);
        for (int it = 0; it < n_iter; it++) {
            iters++;
            double W_old = W, v2_old = v2;
            double f0 = v2 - S2 / (W * W);
            double f1 = Qt + W - c * (W * (1.0 - v2) - D * dsqrt(1.0 - v2));
            double J0 = 2.0 * S2 / (W * W * W);
            double J1 = 1.0 - c * (1.0 - v2);
            double J2 = 1.0;
            double J3 = c * (W - D / (2.0 * dsqrt(1.0 - v2)));
            // 3dim/structs.jl:181-214  LU_dec_2D!
            J1 /= J0;
            J3 -= J1 * J2;
            double x1 = f1;
            double x0 = f0;
            x1 -= J1 * x0;
            x1 /= J3;
            x0 -= J2 * x1;
            x0 /= J0;
            double dW = x0, dv2 = x1, alpha = 1.0;
            bool accepted = false;
            for (int b = 0; b < 10; b++) {
                double sqv = dsqrt(W - alpha * dW);
                double Wc = sqv * sqv;
                double v2c = v2 - alpha * dv2;
                if (Wc <= 0.0 || Wc > W_max) Wc = W_old;
                if (v2c < 0.0)
                    v2c = 0.0;
                else if (v2c >= 1.0)
                    v2c = 1.0 - 1e-15;
                double r1 = v2c - S2 / (Wc * Wc);
                double r2 = Qt + Wc - c * (Wc * (1.0 - v2c) - D * dsqrt(1.0 - v2c));
                double errc = r1 * r1 + r2 * r2;
                double r1o = v2 - S2 / (W * W);
                double r2o = Qt + W - c * (W * (1.0 - v2) - D * dsqrt(1.0 - v2));
                double erro = r1o * r1o + r2o * r2o;
                if (errc <= erro) {
                    W = Wc;
                    v2 = v2c;
                    accepted = true;
                    break;
                }
                alpha *= 0.5;
            }
            if (!accepted) {
                W = W_old;
                v2 = v2_old;
            }
            double relW = dsqrt((W - W_old) * (W - W_old)) / jmax(dabs(W_old), tol * tol);
            double relv = dsqrt((v2 - v2_old) * (v2 - v2_old)) / jmax(dabs(v2_old), tol * tol);
            if (relW + relv < tol) {
                conv2 = true;
                break;
            }
        }
    }
    if (!conv2) {
        for (int it = 0; it < n_iter; it++) {
            iters++;
            double gam = dsqrt(P[2] * P[2] + P[3] * P[3] + P[4] * P[4] + 1.0);
            double w = g * P[1] + P[0];
            double f[5], J[25], o[5];
            f[0] = gam * P[0] - U[0];
            f[1] = (g - 1.0) * P[1] - gam * gam * w - U[1];
            f[2] = P[2] * gam * w - U[2];
            f[3] = P[3] * gam * w - U[3];
            f[4] = P[4] * gam * w - U[4];
            J[0] = gam;
            J[5] = 0.0;
            J[10] = P[2] * P[0] / gam;
            J[15] = P[3] * P[0] / gam;
            J[20] = P[4] * P[0] / gam;
            J[1] = -(gam * gam);
            J[6] = g * (-(gam * gam)) + g - 1.0;
            J[11] = -2.0 * P[2] * w;
            J[16] = -2.0 * P[3] * w;
            J[21] = -2.0 * P[4] * w;
            J[2] = P[2] * gam;
            J[7] = g * P[2] * gam;
            J[12] = P[2] * P[2] * w / gam + w * gam;
            J[17] = P[2] * P[3] * w / gam;
            J[22] = P[2] * P[4] * w / gam;
            J[3] = P[3] * gam;
            J[8] = g * P[3] * gam;
            J[13] = P[2] * P[3] * w / gam;
            J[18] = P[3] * P[3] * w / gam + w * gam;
            J[23] = P[3] * P[4] * w / gam;
            J[4] = P[4] * gam;
            J[9] = g * P[4] * gam;
            J[14] = P[2] * P[4] * w / gam;
            J[19] = P[3] * P[4] * w / gam;
            J[24] = P[4] * P[4] * w / gam + w * gam;
            // 3dim/structs.jl:144-179  LU_dec_5D! (no pivoting)
#pragma unroll
            for (int k = 0; k < 5; k++)
#pragma unroll
                for (int i = k + 1; i < 5; i++) {
                    J[k * 5 + i] /= J[k * 5 + k];
#pragma unroll
                    for (int j = k + 1; j < 5; j++) J[j * 5 + i] -= J[k * 5 + i] * J[j * 5 + k];
                }
#pragma unroll
            for (int i = 0; i < 5; i++) {
                o[i] = f[i];
#pragma unroll
                for (int j = 0; j < i; j++) o[i] -= J[j * 5 + i] * o[j];
            }
#pragma unroll
            for (int i = 4; i >= 0; i--) {
#pragma unroll
                for (int j = i + 1; j < 5; j++) o[i] -= J[j * 5 + i] * o[j];
                o[i] /= J[i * 5 + i];
            }
            if (o[0] * o[0] + o[1] * o[1] + o[2] * o[2] + o[3] * o[3] + o[4] * o[4] < tol * tol) {
                if (__builtin_isfinite(P[0]) && __builtin_isfinite(P[1]) &&
                    __builtin_isfinite(P[2]) && __builtin_isfinite(P[3]) &&
                    __builtin_isfinite(P[4])) {
                    conv5 = true;
                    P[0] = jmax(1e-8, P[0]);
                    P[1] = jmax(1e-8, P[1]);
                    double gm = dsqrt(P[2] * P[2] + P[3] * P[3] + P[4] * P[4] + 1.0);
                    P[2] = P[2] / gm;
                    P[3] = P[3] / gm;
                    P[4] = P[4] / gm;
                    gm = jmin(gm, 50.0);
                    P[2] = P[2] * gm;
                    P[3] = P[3] * gm;
                    P[4] = P[4] * gm;
                }
                break;
            }
#pragma unroll
            for (int v = 0; v < 5; v++) P[v] = P[v] - o[v];
        }
    }
    if (!conv2 && !conv5) {
        auto res = [&](double Wv) {
            return Qt + Wv - c * (Wv * (1.0 - S2 / (Wv * Wv)) - D * dsqrt(jmax(0.0, 1.0 - S2 / (Wv * Wv))));
        };
        double W_min = jmax(1e-15, dsqrt(S2));
        double W_max = W_min * 10.0;
        double fmin = res(W_min), fmax = res(W_max);
        int count = 0;
        while (fmin * fmax > 0.0 && count < 100) {
            W_max *= 10.0;
            fmax = res(W_max);
            count++;
        }
        if (fmin * fmax <= 0.0) {
            const double thr = 1e-6 * 1e-6;
            for (int it = 0; it < n_iter * 10; it++) {
                iters++;
                double W_mid = 0.5 * (W_min + W_max);
                double fmid = res(W_mid);
                if (fmid * fmid < thr) {
                    convb = true;
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
    if (conv2 || convb) {
        v2 = S2 / (W * W);
        double gam = 1.0 / dsqrt(1.0 - v2);
        gam = jmin(gam, 50.0);
        P[0] = jmax(1e-8, D / gam);
        P[1] = jmax(1e-8, (W / (gam * gam) - D / gam) / g);
        P[2] = gam * Qx / W;
        P[3] = gam * Qy / W;
        P[4] = gam * Qz / W;
    }
    if (!conv2 && !conv5 && !convb) {
        const double nan = __builtin_nan("");
#pragma unroll
        for (int v = 0; v < 5; v++) P[v] = nan;
        return 0;
    }
    return conv2 ? 2 : (conv5 ? 3 : 4);
}

// One Newton step of the reference's 4x4 con2prim (2dim/structs.jl:249-330):
// residual f and Jacobian A of P -> U at P, LU with row-scaled partial
// pivoting and explicit fma (:121-208), o = A^-1 f.  Returns false if a row
// of A is identically zero (the reference raises).
__device__ __forceinline__ bool newton_step_2d_lu(const double *U, const double *P, double g, double *o)
{
    double gam = vrn_sqrt1(P[2] * P[2] + P[3] * P[3] + 1.0);
    double w = g * P[1] + P[0];
    double f[4], A[16], rn[4];
#if defined(VRN_FAST)
    // fast form: one reciprocal of gamma, of each row maximum and of each
    // pivot instead of the 19 divisions per iteration
    const double ig = recip(gam);
    double rp[4];
#endif
    f[0] = gam * P[0] - U[0];
    f[1] = (g - 1.0) * P[1] - gam * gam * w - U[1];
    f[2] = P[2] * gam * w - U[2];
    f[3] = P[3] * gam * w - U[3];
    A[0] = gam;
    A[4] = 0.0;
#if defined(VRN_FAST)
    A[8] = P[0] * P[2] * ig;
    A[12] = P[0] * P[3] * ig;
#else
    A[8] = P[0] * P[2] / gam;
    A[12] = P[0] * P[3] / gam;
#endif
    A[1] = -(gam * gam);
    A[5] = (g - 1.0) - gam * gam * g;
    A[9] = -2.0 * P[2] * w;
    A[13] = -2.0 * P[3] * w;
    A[2] = gam * P[2];
    A[6] = gam * P[2] * g;
#if defined(VRN_FAST)
    A[10] = (2.0 * (P[2] * P[2]) + P[3] * P[3] + 1.0) * ig * w;
    A[14] = P[2] * P[3] * ig * w;
#else
    A[10] = (2.0 * (P[2] * P[2]) + P[3] * P[3] + 1.0) / gam * w;
    A[14] = P[2] * P[3] / gam * w;
#endif
    A[3] = gam * P[3];
    A[7] = gam * P[3] * g;
#if defined(VRN_FAST)
    A[11] = P[2] * P[3] * ig * w;
    A[15] = (2.0 * (P[3] * P[3]) + P[2] * P[2] + 1.0) * ig * w;
#else
    A[11] = P[2] * P[3] / gam * w;
    A[15] = (2.0 * (P[3] * P[3]) + P[2] * P[2] + 1.0) / gam * w;
#endif
#pragma unroll
    for (int i = 0; i < 4; i++) {
        double am = 0.0;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            double t = dabs(A[j * 4 + i]);
            // `if (t > am) am = t` of the reference: am starts at +0 and
            // t = |A| is >= +0 or a quiet NaN, which both forms skip, so
            // IEEE maxNum gives the same value in one instruction
            am = __builtin_fmax(am, t);
        }
        if (am == 0.0) return false;
        rn[i] = recip(am);
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
        int piv = k;
        double pv = dabs(A[k * 4 + k]) * rn[k];
#pragma unroll
        for (int i = k + 1; i < 4; i++) {
            double t = dabs(A[k * 4 + i]) * rn[i];
            if (t > pv) {
                pv = t;
                piv = i;
            }
        }
        // Row swap written as selects so the arrays stay in registers.
        // The right-hand side is eliminated along with the matrix (the
        // forward substitution's fma(-L_ij, x_j, x_i) for j = 0..i-1, in
        // that order, as each step finishes), so the multipliers of the
        // columns before k are dead and only columns k..3 move; rn[k] is
        // not read after this step.
#pragma unroll
        for (int i = k + 1; i < 4; i++) {
            bool s = (piv == i);
#pragma unroll
            for (int j = k; j < 4; j++) {
                double a = A[j * 4 + k], b = A[j * 4 + i];
                A[j * 4 + k] = s ? b : a;
                A[j * 4 + i] = s ? a : b;
            }
            double a = f[k], b = f[i];
            f[k] = s ? b : a;
            f[i] = s ? a : b;
            rn[i] = s ? rn[k] : rn[i];
        }
#if defined(VRN_FAST)
        rp[k] = recip(A[k * 4 + k]);
#endif
#pragma unroll
        for (int i = k + 1; i < 4; i++) {
#if defined(VRN_FAST)
            const double m = A[k * 4 + i] * rp[k];
#else
            const double m = A[k * 4 + i] / A[k * 4 + k];
#endif
#pragma unroll
            for (int j = k + 1; j < 4; j++) A[j * 4 + i] = __builtin_fma(-m, A[j * 4 + k], A[j * 4 + i]);
            f[i] = __builtin_fma(-m, f[k], f[i]);
        }
    }
#pragma unroll
    for (int i = 3; i >= 0; i--) {
        o[i] = f[i];
#pragma unroll
        for (int j = i + 1; j < 4; j++) o[i] = __builtin_fma(-A[j * 4 + i], o[j], o[i]);
#if defined(VRN_FAST)
        o[i] *= rp[i];
#else
        o[i] /= A[i * 4 + i];
#endif
    }
    return true;
}

#if defined(VRN_FAST)
// Fast FP64 form: the same Newton step solved in closed form.  With
// c = (g-1)/g, q2 = ux^2 + uy^2 = Gamma^2 - 1, a = d_rho + g d_u and
// s = ux d_ux + uy d_uy, rows 2 and 3 of A o = f read
//   Gamma u_i a + (w/Gamma)(Gamma^2 d_ui + u_i s) = f_i,
// row 0 gives d_rho = (f0 - rho s/Gamma)/Gamma, and ux * row 2 + uy * row 3
// together with row 1 leave the 2x2 system
//   (c - Gamma^2) a + (c rho/Gamma^2 - 2w) s = f1 + c f0/Gamma
//   Gamma q2 a + (w/Gamma)(2 Gamma^2 - 1) s = ux f2 + uy f3
// whose determinant is Dn/Gamma with
//   Dn = w ((2c - 1) Gamma^2 - c) - c rho q2 <= -c w < 0
// for 1 < g <= 2, rho > 0, w > 0 (the Gamma^4 terms cancel algebraically, so
// there is no cancellation in floating point).  Then d_u = (a - d_rho)/g and
// d_ui = (Gamma (f_i - Gamma u_i a)/w - u_i s)/Gamma^2.  On such states no row
// of A is zero (row 0 holds Gamma >= 1, row 1 -Gamma^2, row 2 Gamma ux or
// Gamma w), so the reference's singular exit cannot occur.  Against a 50-digit
// solve of the same systems it is as accurate as the pivoted LU (median
// 7e-14 vs 2e-13, p99 1e-8 vs 3e-8 component-wise relative, Gamma up to 200).
__device__ __forceinline__ void newton_step_2d_closed(const double *U, const double *P, double g, double c,
                                                      double ig_g, double q2, double g2, double w, double *o)
{
    const double rho = P[0], x = P[2], y = P[3];
    const double gam = vrn_sqrt1(g2);
    const double ig = recip(gam);
    const double f0 = gam * rho - U[0];
    const double f1 = (g - 1.0) * P[1] - g2 * w - U[1];
    const double f2 = x * gam * w - U[2];
    const double f3 = y * gam * w - U[3];
    const double r1 = f1 + c * f0 * ig;
    const double r2 = x * f2 + y * f3;
    const double iD = recip(w * ((2.0 * c - 1.0) * g2 - c) - c * rho * q2);
    const double a = (r1 * w * (2.0 * g2 - 1.0) + (2.0 * w * gam - c * rho * ig) * r2) * iD;
    const double s = gam * ((c - g2) * r2 - gam * q2 * r1) * iD;
    const double iw = recip(w), ig2 = ig * ig;
    o[0] = (f0 - rho * ig * s) * ig;
    o[1] = (a - o[0]) * ig_g;
    o[2] = (gam * (f2 - gam * x * a) * iw - x * s) * ig2;
    o[3] = (gam * (f3 - gam * y * a) * iw - y * s) * ig2;
}
#endif

// 2dim/structs.jl:249-330  4x4 Newton on (rho, u, ux, uy), warm start in P,
// break before the update when |o| < tol.  Returns false if a row of the
// Jacobian is identically zero (the reference raises).  The fast FP64 form
// takes the closed-form step on regular states (1 < g <= 2, rho > 0,
// w = g u + rho > 1e-300, Gamma^2 w < 1e300) and the LU elsewhere.
__device__ __forceinline__ bool c2p_2d(const double *U, double *P, double g, int n_iter, double tol,
                                       int &iters)
{
#if defined(VRN_FAST)
    const double c = (g - 1.0) / g, ig_g = 1.0 / g;
#endif
    for (int it = 0; it < n_iter; it++) {
        iters++;
        double o[4];
#if defined(VRN_FAST)
        const double q2 = P[2] * P[2] + P[3] * P[3], g2 = q2 + 1.0;
        const double w = g * P[1] + P[0];
        if (g <= 2.0 && P[0] > 0.0 && w > 1e-300 && g2 * w < 1e300) {
            newton_step_2d_closed(U, P, g, c, ig_g, q2, g2, w, o);
        } else
#endif
        if (!newton_step_2d_lu(U, P, g, o)) {
            return false;
        }
#if defined(VRN_FAST)
        if (o[0] * o[0] + o[1] * o[1] + o[2] * o[2] + o[3] * o[3] < tol * tol) break;
#else
        if (dsqrt(o[0] * o[0] + o[1] * o[1] + o[2] * o[2] + o[3] * o[3]) < tol) break;
#endif
#pragma unroll
        for (int v = 0; v < 4; v++) P[v] = P[v] - o[v];
    }
    return true;
}

}  // namespace VRN_NS
