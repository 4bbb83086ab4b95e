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
//             |r|^13/13! < 1.8e-16 relative.  Coefficients are computed on
//             the host as 1/m! and passed in.
//   FP32      v_rsq_f32 and v_exp_f32 (2^x) hardware transcendentals.
#pragma once
#include <hip/hip_runtime.h>

namespace yk {

// Constants of the FP64 special functions.  Kernels receive them as kernel
// arguments so they live in SGPRs (a VOP3 FMA may read one scalar operand)
// instead of being rematerialised in VGPRs for every pair.
struct ExpCoeffs {
    double c[11];                 // Taylor coefficients 1/m!, m = 12 .. 2
    double log2e, ln2_hi, ln2_lo; // Cody-Waite reduction
    double three_eighths;         // rsqrt correction
};

inline ExpCoeffs exp_coeffs()
{
    ExpCoeffs e;
    double f = 1.0;
    for (int m = 2; m <= 12; ++m) {
        f *= m;
        e.c[12 - m] = 1.0 / f;
    }
    e.log2e = 1.4426950408889634074;
    e.ln2_hi = 6.93147180369123816490e-01;   // fdlibm split of ln 2
    e.ln2_lo = 1.90821492927058770002e-10;
    e.three_eighths = 0.375;
    return e;
}

__device__ __forceinline__ double rsqrt64(double x, const ExpCoeffs &ec)
{
    double y = __builtin_amdgcn_rsq(x);
    double t = x * y;
    double e = __builtin_fma(-t, y, 1.0);
    double p = __builtin_fma(e, ec.three_eighths, 0.5);
    double m = y * e;
    return __builtin_fma(m, p, y);
}

// e^x for x <= 0 (x = -0 included; large negative x underflows to 0 through
// ldexp, so no clamp is needed).
__device__ __forceinline__ double exp_neg64(double x, const ExpCoeffs &ec)
{
    double k = __builtin_rint(x * ec.log2e);
    double r = __builtin_fma(-k, ec.ln2_hi, x);
    r = __builtin_fma(-k, ec.ln2_lo, r);
    double p = ec.c[0];
#pragma unroll
    for (int m = 1; m < 11; ++m) p = __builtin_fma(p, r, ec.c[m]);
    p = __builtin_fma(p, r, 1.0);
    p = __builtin_fma(p, r, 1.0);
    return __builtin_amdgcn_ldexp(p, (int)k);
}

__device__ __forceinline__ float rsqrt32(float x) { return __builtin_amdgcn_rsqf(x); }

// 2^x with the hardware exp2 (v_exp_f32).
__device__ __forceinline__ float exp2_32(float x) { return __builtin_amdgcn_exp2f(x); }

}  // namespace yk
