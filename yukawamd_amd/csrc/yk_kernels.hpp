// yk_kernels.hpp -- HIP kernels of the Yukawa pair-force / velocity-Verlet
// hot path for gfx950 (CDNA4, wave64).
//
// Layout in HBM (structure of arrays, one system per handle):
//   pos[n]    double4 {x, y, z, Z}   -- position and charge number, 32 B
//   vel[3n]   component-major vx[n] | vy[n] | vz[n]
//   vold[3n]  predictor copy for the vel_verlet_2 form
//   acc[3n]   accelerations, component-major
//   aii[n], den[n]   mass number, and the acceleration denominator
//                    (A_i*xmass for ions, xmass for nucleons)
//   part[S][NC][n]   per-chunk partial sums of the pair kernel
//
// Pair kernel: each workgroup owns BLOCK consecutive targets i (one per
// lane) and one chunk of the source range j; j tiles of TILE particles are
// staged in LDS and read as wave-uniform broadcasts.  Every lane sums its own
// target in registers (full N x N matrix, no atomics), so results are
// deterministic; the S chunk partials are summed in fixed order by the
// finalize kernel.  Minimum image is d - L*rint(d/L) per axis; the cutoff
// mask is r^2 <= rc^2 and j != i.
#pragma once
#include <hip/hip_runtime.h>
#include "yk_math.hpp"

namespace yk {

constexpr int BLOCK = 256;   // targets per workgroup (4 wave64)
constexpr int TILE = 256;    // sources per LDS tile
constexpr int NQ = 14;       // reduced scalars: E, W, PP[6], KE[6]

enum Kind { PURE = 0, MIX = 1, NN = 2 };

struct ForceConsts {
    double L[3], invL[3];
    double rc2, rn2, mu, evcut;
    double xp2i, xpi, aa, aa2, bpc, bmc, vc;
    float muf, mul2ef, xp2il2ef, xpif, aaf, aa2f, bpcf, bmcf, vcf;
    int n, coul_on, nuc_on, pad;
    ExpCoeffs ec;
};

struct StepConsts {
    double dt, afac, hdt, bfield_c;  // bfield_c = 3.218404e-20 * bfield
    double L[3];
    double cnum;                     // frp*vc (mix), frp*zi*zi*vc (pure), 1 (nn)
    int n, ref_f1, kind, has_bfield;
};

// ---------------------------------------------------------------------------
// Per-lane accumulator.
template <bool MEAS> struct Acc {
    double fx = 0, fy = 0, fz = 0;
};
template <> struct Acc<true> {
    double fx = 0, fy = 0, fz = 0, e = 0, w = 0;
    double pxx = 0, pxy = 0, pxz = 0, pyy = 0, pyz = 0, pzz = 0;
};

template <bool MEAS>
__device__ __forceinline__ void accumulate(Acc<MEAS> &a, double s, double dx,
                                           double dy, double dz, double r2,
                                           double phi)
{
    a.fx = __builtin_fma(s, dx, a.fx);   // identical in both templates
    a.fy = __builtin_fma(s, dy, a.fy);
    a.fz = __builtin_fma(s, dz, a.fz);
    if constexpr (MEAS) {
        double sx = s * dx, sy = s * dy, sz = s * dz;
        a.e += phi;
        a.w = __builtin_fma(s, r2, a.w);
        a.pxx = __builtin_fma(dx, sx, a.pxx);
        a.pxy = __builtin_fma(dx, sy, a.pxy);
        a.pxz = __builtin_fma(dx, sz, a.pxz);
        a.pyy = __builtin_fma(dy, sy, a.pyy);
        a.pyz = __builtin_fma(dy, sz, a.pyz);
        a.pzz = __builtin_fma(dz, sz, a.pzz);
    } else {
        (void)r2; (void)phi;
    }
}

__device__ __forceinline__ double min_image(double d, double L, double invL)
{
    return __builtin_fma(-L, __builtin_rint(d * invL), d);
}

// One (i, j) pair.  PREC 0 = FP64, 1 = mixed.
template <int KIND, int PREC, bool MEAS>
__device__ __forceinline__ void pair(const ForceConsts &c, const double4 &pi,
                                     const double4 &pj, bool self, Acc<MEAS> &acc)
{
    double dx = min_image(pi.x - pj.x, c.L[0], c.invL[0]);
    double dy = min_image(pi.y - pj.y, c.L[1], c.invL[1]);
    double dz = min_image(pi.z - pj.z, c.L[2], c.invL[2]);
    double r2 = __builtin_fma(dx, dx, __builtin_fma(dy, dy, dz * dz));
    if constexpr (KIND != NN) {
        if (r2 <= c.rc2 && !self) {
            double s, phi = 0.0;
            if constexpr (PREC == 0) {
                double rinv = rsqrt64(r2, c.ec);
                double r = r2 * rinv;
                double mr = c.mu * r;
                double ex = exp_neg64(-mr, c.ec);
                double exr = ex * rinv;
                double fc = (1.0 + mr) * exr * (rinv * rinv);
                s = (KIND == MIX) ? pj.w * fc : fc;
                if constexpr (MEAS) {
                    phi = exr - c.evcut;
                    if (KIND == MIX) phi *= pj.w;
                }
            } else {
                float r2f = (float)r2;
                float rinv = rsqrt32(r2f);
                float r = r2f * rinv;
                float ex = exp2_32(-c.mul2ef * r);
                float exr = ex * rinv;
                float fc = __builtin_fmaf(c.muf, r, 1.0f) * exr * (rinv * rinv);
                float zj = (float)pj.w;
                s = (double)((KIND == MIX) ? zj * fc : fc);
                if constexpr (MEAS) {
                    phi = (double)exr - c.evcut;
                    if (KIND == MIX) phi *= pj.w;
                }
            }
            accumulate<MEAS>(acc, s, dx, dy, dz, r2, phi);
        }
    } else {
        bool hn = c.nuc_on && r2 <= c.rn2;
        bool hc = c.coul_on && r2 <= c.rc2 && pi.w == 1.0 && pj.w == 1.0;
        if ((hn || hc) && !self) {
            double s = 0.0, phi = 0.0;
            const bool same = (pi.w == pj.w);
            if constexpr (PREC == 0) {
                if (hn) {
                    double g = exp_neg64(-r2 * c.xp2i, c.ec);
                    double bc = same ? c.bpc : c.bmc;
                    s = c.xpi * __builtin_fma(c.aa2, g, bc) * g;
                    if constexpr (MEAS) phi = __builtin_fma(c.aa, g, bc) * g;
                }
                if (hc) {
                    double rinv = rsqrt64(r2, c.ec);
                    double r = r2 * rinv;
                    double mr = c.mu * r;
                    double ex = exp_neg64(-mr, c.ec);
                    double exr = c.vc * ex * rinv;
                    s += (1.0 + mr) * exr * (rinv * rinv);
                    if constexpr (MEAS) phi += exr - c.evcut;
                }
            } else {
                float r2f = (float)r2;
                float sf = 0.0f;
                if (hn) {
                    float g = exp2_32(-r2f * c.xp2il2ef);
                    float bc = same ? c.bpcf : c.bmcf;
                    sf = c.xpif * __builtin_fmaf(c.aa2f, g, bc) * g;
                    if constexpr (MEAS) phi = (double)(__builtin_fmaf(c.aaf, g, bc) * g);
                }
                if (hc) {
                    float rinv = rsqrt32(r2f);
                    float r = r2f * rinv;
                    float ex = exp2_32(-c.mul2ef * r);
                    float exr = c.vcf * ex * rinv;
                    sf += __builtin_fmaf(c.muf, r, 1.0f) * exr * (rinv * rinv);
                    if constexpr (MEAS) phi += (double)exr - c.evcut;
                }
                s = (double)sf;
            }
            accumulate<MEAS>(acc, s, dx, dy, dz, r2, phi);
        }
    }
}

// ---------------------------------------------------------------------------
// All-pairs partial sums over one j chunk.  grid = (ceil(n/BLOCK), S).
template <int KIND, int PREC, bool MEAS>
__global__ __launch_bounds__(BLOCK) void k_pairs(const double4 *__restrict__ pos,
                                                 ForceConsts c, int tiles_per_chunk,
                                                 double *__restrict__ part)
{
    __shared__ double4 sj[TILE];
    const int n = c.n;
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    const int chunk = blockIdx.y;
    const int ntiles = (n + TILE - 1) / TILE;
    const double nanv = __builtin_nan("");
    const double4 pi = (i < n) ? pos[i] : make_double4(nanv, nanv, nanv, 0.0);
    Acc<MEAS> acc;
    const int t0 = chunk * tiles_per_chunk;
    const int t1 = min(t0 + tiles_per_chunk, ntiles);
    for (int t = t0; t < t1; ++t) {
        const int jb = t * TILE;
        const int jl = jb + threadIdx.x;
        // padded sources are NaN: every comparison with them is false
        sj[threadIdx.x] = (jl < n) ? pos[jl] : make_double4(nanv, nanv, nanv, 0.0);
        __syncthreads();
        const int self_jj = i - jb;   // in [0, TILE) only for the diagonal tile
#pragma unroll 4
        for (int jj = 0; jj < TILE; ++jj)
            pair<KIND, PREC, MEAS>(c, pi, sj[jj], jj == self_jj, acc);
        __syncthreads();
    }
    if (i < n) {
        const size_t nn = (size_t)n;
        double *o = part + (size_t)chunk * (MEAS ? 11 : 3) * nn + i;
        o[0] = acc.fx; o[nn] = acc.fy; o[2 * nn] = acc.fz;
        if constexpr (MEAS) {
            o[3 * nn] = acc.e; o[4 * nn] = acc.w;
            o[5 * nn] = acc.pxx; o[6 * nn] = acc.pxy; o[7 * nn] = acc.pxz;
            o[8 * nn] = acc.pyy; o[9 * nn] = acc.pyz; o[10 * nn] = acc.pzz;
        }
    }
}

// ---------------------------------------------------------------------------
// Deterministic block reduction of NQ values per thread (fixed tree order).
__device__ __forceinline__ void block_reduce_store(double (&q)[NQ], double *__restrict__ out)
{
    __shared__ double red[NQ][BLOCK];
#pragma unroll
    for (int k = 0; k < NQ; ++k) red[k][threadIdx.x] = q[k];
    __syncthreads();
    for (int s = BLOCK / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
#pragma unroll
            for (int k = 0; k < NQ; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x < NQ) out[threadIdx.x] = red[threadIdx.x][0];
}

// MODE 0: accel only.  MODE 1: vel_verlet_2 corrector (src/vel_verlet_2.f:61-79).
// MODE 2: kick-drift-kick closing half kick (src/vel_verlet_1.f).
template <int KIND, bool MEAS, int MODE>
__global__ __launch_bounds__(BLOCK) void k_finalize(
    const double *__restrict__ part, int nchunks, const double4 *__restrict__ pos,
    const double *__restrict__ den, const double *__restrict__ aii,
    double *__restrict__ acc, double *__restrict__ vel,
    const double *__restrict__ vold, StepConsts sc, double *__restrict__ blockpart)
{
    const int n = sc.n;
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    const size_t nn = (size_t)n;
    constexpr int NC = MEAS ? 11 : 3;
    double q[NQ];
#pragma unroll
    for (int k = 0; k < NQ; ++k) q[k] = 0.0;
    if (i < n) {
        double f[NC];
#pragma unroll
        for (int k = 0; k < NC; ++k) f[k] = 0.0;
        for (int c = 0; c < nchunks; ++c) {
            const double *p = part + (size_t)c * NC * nn + i;
#pragma unroll
            for (int k = 0; k < NC; ++k) f[k] += p[k * nn];
        }
        double zi = pos[i].w;
        double fz_i = (KIND == MIX && !(sc.ref_f1 && i == n - 1)) ? zi : 1.0;
        double d = den[i];
        double a[3];   // fz_i is exactly 1 unless this is a mixture
#pragma unroll
        for (int k = 0; k < 3; ++k) a[k] = (sc.cnum * (fz_i * f[k])) / d;
        if constexpr (MODE == 1) {
            if (sc.has_bfield) {   // src/vel_verlet_2.f:61-69
                double omega = sc.bfield_c * zi / aii[i];
                a[0] = a[0] + vel[nn + i] * omega;
                a[1] = a[1] - vel[i] * omega;
            }
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) acc[k * nn + i] = a[k];
        double v[3] = {0, 0, 0};
        if constexpr (MODE == 1) {   // v = 0.5*(v + (vold + dt*a))
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                double vp = vold[k * nn + i] + sc.dt * a[k];
                v[k] = 0.5 * (vel[k * nn + i] + vp);
                vel[k * nn + i] = v[k];
            }
        } else if constexpr (MODE == 2) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                v[k] = vel[k * nn + i] + sc.hdt * a[k];
                vel[k * nn + i] = v[k];
            }
        }
        if constexpr (MEAS) {
            double zs = (KIND == MIX) ? zi : 1.0;   // per-i zii factor (:199-207)
            q[0] = 0.5 * zs * f[3];
            q[1] = 0.5 * zs * f[4];
#pragma unroll
            for (int k = 0; k < 6; ++k) q[2 + k] = 0.5 * zs * f[5 + k];
            if constexpr (MODE != 0) {
                double ai = aii[i];
                q[8] = ai * v[0] * v[0];
                q[9] = ai * v[0] * v[1];
                q[10] = ai * v[0] * v[2];
                q[11] = ai * v[1] * v[1];
                q[12] = ai * v[1] * v[2];
                q[13] = ai * v[2] * v[2];
            }
        }
    }
    if constexpr (MEAS) block_reduce_store(q, blockpart + (size_t)blockIdx.x * NQ);
}

// Drift (+ velocity predictor or opening half kick).  MODE 1: vel_verlet_2
// (src/vel_verlet_2.f:33-53); MODE 2: drift + half kick (src/vel_verlet_1.f:28-48).
template <int MODE>
__global__ __launch_bounds__(BLOCK) void k_drift(double4 *__restrict__ pos,
                                                 double *__restrict__ vel,
                                                 double *__restrict__ vold,
                                                 const double *__restrict__ acc,
                                                 StepConsts sc)
{
    const int n = sc.n;
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const size_t nn = (size_t)n;
    double4 p = pos[i];
    double x[3] = {p.x, p.y, p.z};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        double v = vel[k * nn + i], a = acc[k * nn + i], xk;
        if constexpr (MODE == 1) {
            xk = x[k] + v * sc.dt + sc.afac * a;   // contraction is off in this TU
            vold[k * nn + i] = v;
            vel[k * nn + i] = v + sc.dt * a;
        } else {   // src/vel_verlet_1.f:30-48
            xk = x[k] + v * sc.dt + a * sc.afac;
            vel[k * nn + i] = v + sc.hdt * a;
        }
        if (xk < 0.0) xk = xk + sc.L[k];
        if (xk > sc.L[k]) xk = xk - sc.L[k];
        x[k] = xk;
    }
    pos[i] = make_double4(x[0], x[1], x[2], p.w);
}

// Kinetic energy sum A*(vx^2+vy^2+vz^2) per block (ttot, src/gen_utils.f:40-55).
__global__ __launch_bounds__(BLOCK) void k_ttot(const double *__restrict__ vel,
                                                const double *__restrict__ aii, int n,
                                                double *__restrict__ blockpart)
{
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    const size_t nn = (size_t)n;
    double q[NQ];
#pragma unroll
    for (int k = 0; k < NQ; ++k) q[k] = 0.0;
    if (i < n) {
        double vx = vel[i], vy = vel[nn + i], vz = vel[2 * nn + i];
        q[0] = aii[i] * (vx * vx + vy * vy + vz * vz);
    }
    block_reduce_store(q, blockpart + (size_t)blockIdx.x * NQ);
}

// Final fixed-order reduction of block partials into the scalars.
struct MeasOut {
    double ev, ek, px, vir, pp[6];
};

struct MeasConsts {
    double cnum, rho, kT, xmass, vol;
    float three_n;   // 3.0*n as the reference's real(4) product
    float n_f;       // float(n)
    int mode;        // 0 accel, 1/2 newton, 3 ttot
};

__global__ __launch_bounds__(BLOCK) void k_meas_final(const double *__restrict__ blockpart,
                                                      int nblocks, MeasConsts mc,
                                                      MeasOut *__restrict__ out)
{
    double q[NQ];
#pragma unroll
    for (int k = 0; k < NQ; ++k) q[k] = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += BLOCK) {
#pragma unroll
        for (int k = 0; k < NQ; ++k) q[k] += blockpart[(size_t)b * NQ + k];
    }
    __shared__ double tot[NQ];
    {
        __shared__ double red[NQ][BLOCK];
#pragma unroll
        for (int k = 0; k < NQ; ++k) red[k][threadIdx.x] = q[k];
        __syncthreads();
        for (int s = BLOCK / 2; s > 0; s >>= 1) {
            if (threadIdx.x < s) {
#pragma unroll
                for (int k = 0; k < NQ; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + s];
            }
            __syncthreads();
        }
        if (threadIdx.x < NQ) tot[threadIdx.x] = red[threadIdx.x][0];
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    MeasOut m;
    if (mc.mode == 3) {
        m.ev = 0; m.px = 0; m.vir = 0;
        for (int k = 0; k < 6; ++k) m.pp[k] = 0;
        m.ek = (0.5 * mc.xmass * tot[0]) / (double)mc.n_f;
        *out = m;
        return;
    }
    double ev = mc.cnum * tot[0];
    double vir = mc.cnum * tot[1];
    m.ev = ev / (double)mc.n_f;
    m.vir = vir;
    m.px = mc.rho * (mc.kT + vir / (double)mc.three_n);
    for (int k = 0; k < 6; ++k) m.pp[k] = mc.cnum * tot[2 + k];
    m.ek = 0.0;
    if (mc.mode != 0) {   // src/vel_verlet_2.f:96-106
        m.ek = (0.5 * mc.xmass) * (tot[8] + tot[11] + tot[13]) / (double)mc.n_f;
        for (int k = 0; k < 6; ++k) m.pp[k] = (m.pp[k] + mc.xmass * tot[8 + k]) / mc.vol;
    }
    *out = m;
}

}  // namespace yk
