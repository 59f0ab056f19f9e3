// vrn_sqrt.hpp -- FP64 square root for the hot paths, built on v_rsq_f64.
//
// sqrt_core(x): y = rsq(x), one Goldschmidt step for g ~ sqrt(x) and
// h ~ 1/(2 sqrt(x)), then two Newton corrections g += (x - g^2) h (one in
// the fast FP64 form, compiled with VRN_FAST).  For
// x >= 2^-767 this is exactly the arithmetic of the compiler's correctly
// rounded sqrt() on gfx950 (LLVM's f64 sqrt lowering), which additionally
// rescales arguments below 2^-767 by 2^256 and patches +-0 and +inf with a
// class test: 18 instructions against 10 (13 with the patch).  Results are
// therefore identical wherever the argument is 0, +inf, NaN, negative or at
// least 2^-767 (about 1.6e-231).  Below that the missing rescaling loses
// accuracy; the call sites only pass quantities bounded far away from it
// (1 + |u|^2, sound speeds and wave-speed discriminants of floored states,
// 1 - v^2 >= 1e-15, |S|^2 of non-vacuum momenta).
#pragma once

#include <hip/hip_runtime.h>

#if defined(__HIP_DEVICE_COMPILE__)
__device__ __forceinline__ double vrn_sqrt_core(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y;
    double h = 0.5 * y;
    const double r = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, r, g);
    h = __builtin_fma(h, r, h);
    g = __builtin_fma(__builtin_fma(-g, g, x), h, g);
#if !defined(VRN_FAST)
    // the second correction makes the rounding correct for every argument;
    // the fast FP64 form drops it (over 2^22 random arguments on gfx950 the
    // first correction already gave the correctly rounded value every time)
    g = __builtin_fma(__builtin_fma(-g, g, x), h, g);
#endif
    return g;
}
#endif

// sqrt(x) for x = +-0, +inf, NaN, x < 0 or x >= 2^-767.
__host__ __device__ __forceinline__ double vrn_sqrt(double x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const double g = vrn_sqrt_core(x);
    return __builtin_amdgcn_class(x, 0x260) ? x : g;  // 0x260: -0, +0, +inf
#else
    return __builtin_sqrt(x);
#endif
}

// sqrt(x) for x >= 1, +inf or NaN (Lorentz factors sqrt(1 + |u|^2)).  +inf
// is reached when |u|^2 overflows (|u| > 1.3e154, e.g. a diverging Newton
// iterate); the core gives NaN there (inf * rsq(inf) = inf * 0).  z = x 2^-1022
// is +inf for x = +inf, NaN for NaN x, and for finite x >= 0 it is at most
// sqrt(x) (x 2^-1022 <= sqrt(x) <=> x <= 2^2044), hence at most the correctly
// rounded sqrt(x); so IEEE maxNum(core, z) restores sqrt(+inf) = +inf and
// returns the core's value everywhere else (two instructions and no compare
// mask: a class-test select made the 128-register x sweep spill).  Negative
// arguments are outside the domain (they would return z, not NaN).
__host__ __device__ __forceinline__ double vrn_sqrt1(double x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_fmax(vrn_sqrt_core(x), x * 0x1p-1022);
#else
    return __builtin_sqrt(x);
#endif
}

// FP32 (the FP32 reconstruction/Riemann form): the hardware instruction,
// about 1 ulp, on the device.
__host__ __device__ __forceinline__ float vrn_sqrt(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_sqrtf(x);
#else
    return __builtin_sqrtf(x);
#endif
}
__host__ __device__ __forceinline__ float vrn_sqrt1(float x) { return vrn_sqrt(x); }
