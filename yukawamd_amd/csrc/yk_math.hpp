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
//             ACM TOMS 15 (1989) 144): x = (1024 q + j) ln2/1024 + r with
//             |r| <= ln2/2048, e^x = 2^q * 2^(j/1024) * e^r, 2^(j/1024) from a
//             1024-entry table (rounded to double on the host; the kernels
//             keep it in LDS), e^r by its degree-3 Taylor polynomial
//             (truncation |r|^4/4! < 5.5e-16 relative).  The reduction takes
//             ln2/1024 as one double (one FMA; Tang splits it in two): its
//             rounding, 5.4e-20 at most, adds |k| * 5.4e-20 = |x| * 8e-17 to
//             r, a relative error of 1e-15 at the ion kernels' |x| = mu r <=
//             12 and proportionally less below; the force and energy parity
//             needs 1e-12 / 1e-13.  Ten fewer FP64
//             operations per call than a Cody-Waite reduction to
//             |r| <= ln2/2 with a degree-12 polynomial.  k comes from one FMA
//             against the shifter 1.5 * 2^52 (the sum is a double with unit
//             spacing whose low 32 bits are k), so neither a rounding nor a
//             conversion instruction is needed; that requires |x| <
//             EXP_ARG_MAX, which the kernels' cutoff tests guarantee (see
//             exp_cutoff2).
//   FP32      v_rsq_f32 and v_exp_f32 (2^x) hardware transcendentals.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

namespace yk {

// Constants of the FP64 pair loop.  Kernels receive them as kernel
// arguments so they live in SGPRs (a VOP3 FMA may read one scalar operand)
// instead of being rematerialised in VGPRs for every pair.
constexpr int EXP_TBITS = 10;
constexpr int EXP_TN = 1 << EXP_TBITS;   // entries of the 2^(j/N) table

struct ExpCoeffs {
    double three_eighths;         // rsqrt correction
    double r2_floor;              // 1e-200, added to r^2 by the FP64 list kernels
    double r2_floor_mixed;        // 1e-20, the same for the mixed form (FP32 1/r^3 stays finite)
    double l2eN, ln2N;            // table-driven reduction (N/ln2, ln2/N)
    double shifter;               // 1.5 * 2^52 (exp_neg64t's rounding; kernels pin it in a VGPR)
    double t3[1];                 // 1/3! (1/2! is written inline)
};

inline ExpCoeffs exp_coeffs()
{
    ExpCoeffs e;
    e.three_eighths = 0.375;
    e.r2_floor = 1e-200;
    e.r2_floor_mixed = 1e-20;
    e.l2eN = EXP_TN * 1.4426950408889634074;
    e.ln2N = 0.69314718055994530942 / EXP_TN;   // exact division by a power of two
    e.shifter = 6755399441055744.0;
    e.t3[0] = 1.0 / 6.0;
    return e;
}

// Largest |x| passed to exp_neg64t.  e^x underflows to zero below x = -745,
// so a pair whose exponent lies beyond -EXP_ARG_MAX contributes exactly zero
// to force, energy (whose shift e^{-mu rc}/rc is then zero too), virial and
// pressure tensor; the kernels therefore test pairs against a cutoff no larger
// than EXP_ARG_MAX / rate and never evaluate such an exponent.  (At 1e6 the
// shifter's k stays within 32 bits: |k| <= 1e6 * 1024 / ln2 < 2^31.)
constexpr double EXP_ARG_MAX = 1.0e6;

// Square of the cutoff the kernels test for an exponent rate * d (d = r, or
// d = r^2 for the nucleon Gaussian): min(cut, EXP_ARG_MAX / rate), squared
// when d = r.  Equivalent to cut for every pair that contributes.
inline double exp_cutoff2(double cut2, double rate, bool of_r)
{
    if (!(rate > 0.0)) return cut2;
    const double dmax = EXP_ARG_MAX / rate;
    const double lim2 = of_r ? dmax * dmax : dmax;
    return cut2 < lim2 ? cut2 : lim2;
}

// The nucleon model's Gaussian factor g = e^{-r^2 xp2i} is below 2^-80 beyond
// r^2 = 80 ln2 / xp2i.  A pair there adds at most xpi (2 |aa| g + |b +- c|) g r
// to a force and (|aa| g + |b +- c|) g to the energy: for the pasta parameters
// (aa = 110, b +- c = -2 / -50 MeV, xpacket = 1.25 fm^2, cutoff 20 fm) below
// 7e-22 MeV/fm and 4e-23 MeV per pair, 9e-19 MeV/fm over the 1,300 such
// neighbours of a nucleon.  The kernels evaluate the nuclear term only inside
// that radius, and the cluster lists hold sources that interact only through
// it (neutrons) only out to it.
constexpr double GAUSS_NEGLIGIBLE_LOG = 80.0 * 0.69314718055994530942;

inline double gauss_cutoff2(double cut2, double xp2i)
{
    if (!(xp2i > 0.0)) return cut2;
    const double lim2 = GAUSS_NEGLIGIBLE_LOG / xp2i;
    return cut2 < lim2 ? cut2 : lim2;
}

// The 2^(j/N) table, j = 0 .. N-1, rounded to double from long double.
inline void exp_table(double *t)
{
    for (int j = 0; j < EXP_TN; ++j) t[j] = (double)std::exp2((long double)j / EXP_TN);
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

// e^x for -EXP_ARG_MAX < x <= 0 (x = -0 included; below -745 the result
// underflows to 0 through ldexp); tab = the 2^(j/N) table (LDS).
__device__ __forceinline__ double exp_neg64t(double x, const ExpCoeffs &ec,
                                             const double *__restrict__ tab)
{
    // t = x N/ln2 + 1.5 * 2^52 rounded once: k = t - 1.5 * 2^52 is the integer
    // nearest x N/ln2 (exact subtraction), and t's low 32 bits hold k in two's
    // complement
    const double t = __builtin_fma(x, ec.l2eN, ec.shifter);
    const double k = t - ec.shifter;
    const double r = __builtin_fma(-k, ec.ln2N, x);
    double p = __builtin_fma(r, ec.t3[0], 0.5);   // 1/2 as an inline constant
    p = __builtin_fma(p, r, 1.0);
    p = __builtin_fma(p, r, 1.0);
    // ki = N (ki >> TBITS) + (ki & (N-1)), also for ki < 0
    const int ki = (int)(unsigned)__builtin_bit_cast(unsigned long long, t);
    return __builtin_amdgcn_ldexp(tab[ki & (EXP_TN - 1)] * p, ki >> EXP_TBITS);
}

// Keeps a loop-invariant double in a VGPR pair for the rest of the kernel (an
// empty asm with a VGPR operand; no instruction).  A VALU instruction reads at
// most one SGPR operand, so without it the pair loop copies the shifter from
// its kernel-argument SGPRs into a VGPR for every exp_neg64t call.
__device__ __forceinline__ void pin_vgpr(double &v) { asm volatile("" : "+v"(v)); }

// The workgroup's LDS copy of the table (every thread of the block calls it,
// before any of them may return).
__device__ __forceinline__ void load_exp_table(double *__restrict__ dst,
                                               const double *__restrict__ src)
{
    for (int j = threadIdx.x; j < EXP_TN; j += blockDim.x) dst[j] = src[j];
    __syncthreads();
}

__device__ __forceinline__ float rsqrt32(float x) { return __builtin_amdgcn_rsqf(x); }

// 2^x with the hardware exp2 (v_exp_f32).
__device__ __forceinline__ float exp2_32(float x) { return __builtin_amdgcn_exp2f(x); }

}  // namespace yk
