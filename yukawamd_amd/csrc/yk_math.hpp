// yk_math.hpp -- device special functions for the pair kernels (gfx950).
//
// The FP64 forms are written for the argument ranges the pair kernels use
// (r^2 > 0, exp of a non-positive argument) and return full double accuracy
// (a few ulp), which is what the 1e-12 force / 1e-13 energy parity needs;
// they skip the special-case handling of the general library routines.
//
//   rsqrt64   hardware v_rsq_f64 seed + one third-order Newton-Raphson step
//             (y' = y + y*e*(1/2 + 3e/8), e = 1 - x*y^2).
//   exp_neg64 Cody-Waite reduction x = k*ln2 + r, |r| <= ln2/2, with the
//             two-part ln2 of fdlibm, then the degree-12 Taylor polynomial of
//             e^r in Horner form and a 2^k scale by ldexp.  Truncation error
//             |r|^13/13! < 1.8e-16 relative.
//   FP32      v_rsq_f32 and v_exp_f32 (2^x) hardware transcendentals.
#pragma once
#include <hip/hip_runtime.h>

namespace yk {

__device__ __forceinline__ double rsqrt64(double x)
{
    double y = __builtin_amdgcn_rsq(x);
    double t = x * y;
    double e = __builtin_fma(-t, y, 1.0);
    double p = __builtin_fma(e, 0.375, 0.5);
    double m = y * e;
    return __builtin_fma(m, p, y);
}

// e^x for x <= 0 (x = -0 and underflow to 0 included).
__device__ __forceinline__ double exp_neg64(double x)
{
    const double LOG2E = 1.4426950408889634074;
    const double LN2_HI = 6.93147180369123816490e-01;
    const double LN2_LO = 1.90821492927058770002e-10;
    x = x < -745.5 ? -745.5 : x;   // keeps k inside ldexp's exact range
    double k = __builtin_rint(x * LOG2E);
    double r = __builtin_fma(-k, LN2_HI, x);
    r = __builtin_fma(-k, LN2_LO, r);
    // Taylor coefficients 1/m!, m = 12 .. 0
    double p = 2.08767569878680989792e-09;              // 1/12!
    p = __builtin_fma(p, r, 2.50521083854417187751e-08); // 1/11!
    p = __builtin_fma(p, r, 2.75573192239858906526e-07); // 1/10!
    p = __builtin_fma(p, r, 2.75573192239858906526e-06); // 1/9!
    p = __builtin_fma(p, r, 2.48015873015873015873e-05); // 1/8!
    p = __builtin_fma(p, r, 1.98412698412698412698e-04); // 1/7!
    p = __builtin_fma(p, r, 1.38888888888888888889e-03); // 1/6!
    p = __builtin_fma(p, r, 8.33333333333333333333e-03); // 1/5!
    p = __builtin_fma(p, r, 4.16666666666666666667e-02); // 1/4!
    p = __builtin_fma(p, r, 1.66666666666666666667e-01); // 1/3!
    p = __builtin_fma(p, r, 0.5);
    p = __builtin_fma(p, r, 1.0);
    p = __builtin_fma(p, r, 1.0);
    return __builtin_amdgcn_ldexp(p, (int)k);
}

__device__ __forceinline__ float rsqrt32(float x) { return __builtin_amdgcn_rsqf(x); }

// 2^x with the hardware exp2 (v_exp_f32).
__device__ __forceinline__ float exp2_32(float x) { return __builtin_amdgcn_exp2f(x); }

}  // namespace yk
