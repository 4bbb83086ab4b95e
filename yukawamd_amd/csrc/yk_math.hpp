// yk_math.hpp -- device special functions for the pair kernels (gfx950).
//
// The FP64 forms are written for the argument ranges the pair kernels use
// (r^2 > 0, exp of a non-positive argument) and return full double accuracy
// (a few ulp), which is what the 1e-12 force / 1e-13 energy parity needs;
// they skip the special-case handling of the general library routines.
//
//   rsqrt64   hardware v_rsq_f64 seed + one third-order Newton-Raphson step
//             (y' = y + y*e*(1/2 + 3e/8), e = 1 - x*y^2).
//   exp_neg64t e^x (x <= 0) by Tang's table-driven method (P. T. P. Tang,
//             ACM TOMS 15 (1989) 144): x = (64 q + j) ln2/64 + r with
//             |r| <= ln2/128 (ln2/64 split in two parts, from fdlibm's split
//             of ln 2), e^x = 2^q * 2^(j/64) * e^r, 2^(j/64) from a 64-entry
//             table (rounded to double on the host) that the kernels keep in
//             LDS, e^r by its degree-5 Taylor polynomial (truncation
//             |r|^6/6! < 3.5e-17).  Seven fewer FP64 operations per call than
//             a Cody-Waite reduction to |r| <= ln2/2 with a degree-12
//             polynomial.
//   FP32      v_rsq_f32 and v_exp_f32 (2^x) hardware transcendentals.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

namespace yk {

// Constants of the FP64 special functions.  Kernels receive them as kernel
// arguments so they live in SGPRs (a VOP3 FMA may read one scalar operand)
// instead of being rematerialised in VGPRs for every pair.
struct ExpCoeffs {
    double three_eighths;               // rsqrt correction
    double l2e64, ln2_64hi, ln2_64lo;   // table-driven reduction (64/ln2, ln2/64 split)
    double t5[4];                       // 1/5!, 1/4!, 1/3!, 1/2!
    double t64[64];                     // 2^(j/64), j = 0 .. 63
};

inline ExpCoeffs exp_coeffs()
{
    ExpCoeffs e;
    const double ln2_hi = 6.93147180369123816490e-01;   // fdlibm split of ln 2
    const double ln2_lo = 1.90821492927058770002e-10;
    e.three_eighths = 0.375;
    e.l2e64 = 64.0 * 1.4426950408889634074;
    e.ln2_64hi = ln2_hi / 64.0;   // exact: division by a power of two
    e.ln2_64lo = ln2_lo / 64.0;
    e.t5[0] = 1.0 / 120.0;
    e.t5[1] = 1.0 / 24.0;
    e.t5[2] = 1.0 / 6.0;
    e.t5[3] = 0.5;
    for (int j = 0; j < 64; ++j) e.t64[j] = (double)std::exp2((long double)j / 64.0L);
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
// ldexp, so no clamp is needed); tab = the 64 entries of ec.t64 (LDS).
__device__ __forceinline__ double exp_neg64t(double x, const ExpCoeffs &ec,
                                             const double *__restrict__ tab)
{
    const double k = __builtin_rint(x * ec.l2e64);
    double r = __builtin_fma(-k, ec.ln2_64hi, x);
    r = __builtin_fma(-k, ec.ln2_64lo, r);
    double p = __builtin_fma(ec.t5[0], r, ec.t5[1]);
    p = __builtin_fma(p, r, ec.t5[2]);
    p = __builtin_fma(p, r, ec.t5[3]);
    p = __builtin_fma(p, r, 1.0);
    p = __builtin_fma(p, r, 1.0);
    const int ki = (int)k;                 // x <= 0: ki <= 0, ki = 64 (ki >> 6) + (ki & 63)
    return __builtin_amdgcn_ldexp(tab[ki & 63] * p, ki >> 6);
}

// Copy of the 2^(j/64) table for one wave (lane l writes entry l).
__device__ __forceinline__ void load_exp_table(double *__restrict__ dst, const ExpCoeffs &ec,
                                               int lane)
{
    dst[lane] = ec.t64[lane];
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ float rsqrt32(float x) { return __builtin_amdgcn_rsqf(x); }

// 2^x with the hardware exp2 (v_exp_f32).
__device__ __forceinline__ float exp2_32(float x) { return __builtin_amdgcn_exp2f(x); }

}  // namespace yk
