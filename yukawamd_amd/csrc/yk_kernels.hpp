// yk_kernels.hpp -- HIP kernels of the Yukawa pair-force / velocity-Verlet
// hot path for gfx950 (CDNA4, wave64).
//
// Layout in HBM (structure of arrays, one system per handle, particles kept
// in Hilbert-curve order of a 2^g x 2^g x 2^g cell grid so that 64 consecutive
// particles -- one wave, one "tile" -- occupy a compact region):
//   pos[n]    double4 {x, y, z, Z}   -- position and charge number, 32 B
//   vel[3n]   component-major vx[n] | vy[n] | vz[n]
//   vold[3n]  predictor copy for the vel_verlet_2 form
//   acc[3n]   accelerations, component-major
//   aii[n], den[n]   mass number, and the acceleration denominator
//                    (A_i*xmass for ions, xmass for nucleons)
//   orig[n]   caller's index of the particle in each slot
//   box[T]    per tile of 64 particles: AABB centre and half-extent
//   part[S][NC][n]   per-chunk partial sums of the pair kernel
//
// Pair kernel: a workgroup is 4 waves = 4 target tiles; chunk y of the grid
// covers a range of source tiles, staged 4 tiles (256 particles) at a time in
// LDS and read as wave-uniform broadcasts.  For each (target tile, source
// tile) a wave first tests the two AABBs: if their minimum-image gap exceeds
// the cutoff the whole tile is skipped.  If every pair of the two tiles has
// the same periodic image along each axis (the AABB spans leave room), the
// wave shifts its own coordinates by that image once and each pair costs one
// subtraction per axis; otherwise each pair takes d - L*rint(d/L).  Every lane
// sums its own target in registers (full N x N matrix, no atomics), and the
// chunk partials are summed in fixed order by the finalize kernel, so results
// are deterministic.
#pragma once
#include <hip/hip_runtime.h>
#include <vector>

#include "yk_layout.hpp"
#include "yk_math.hpp"
#include "yk_sfc.hpp"

namespace yk {

constexpr int STAGE = 256;   // sources staged in LDS at a time (4 tiles)

enum Kind { PURE = 0, MIX = 1, NN = 2 };

struct TileBox {
    double c[3];   // AABB centre
    double h[3];   // AABB half extent
};

struct ForceConsts {
    double L[3], invL[3], halfL_m[3];   // halfL_m: L/2 minus a rounding margin
    double rc2, rn2, mu, evcut;
    double cull2;                        // (largest cutoff)^2 used for tile culling
    double margin;                       // absolute rounding margin for AABB tests
    double xp2i, xpi, aa, aa2, bpc, bmc, vc;
    float muf, mul2ef, xp2il2ef, xpif, aaf, aa2f, bpcf, bmcf, vcf;
    // mu log2 e - mul2ef: the rounding error of mu log2 e in float.  The
    // exponent -(mul2ef r + mul2e_lo r) is formed by one FMA, so the rounding
    // of mul2ef does not bias every pair at distance r by the same relative
    // amount (mu r x 6e-8 at most); one rounding of the exponent remains
    float mul2e_lo;
    // the mixed ion form's far pairs (FP32 throughout) work in lengths scaled
    // by fs = mu log2 e, so that e^{-mu r} = 2^{-r'} and 1 + mu r = 1 + r' ln 2;
    // fs2 = fs^2 converts their force sums back.  rc2f_lo, rc2f_hi: the scaled
    // cutoff^2 narrowed and widened by 4e-6 relative (an FP32 r'^2 within that
    // band takes the FP64 test, so a pair is in or out of the cutoff exactly as
    // in FP64); r2_floor_f: the scaled r^2 floor that keeps a self pair's
    // 1/r'^3 finite (1e-20, below the last bit of any physical r'^2 = fs^2 r^2)
    double fs, fs2;
    float rc2f_lo, rc2f_hi, r2_floor_f;
    int n, ntiles, coul_on, nuc_on;
    ExpCoeffs ec;
    const double *etab_g; // the 2^(j/N) table of exp_neg64t (device memory)
    const double *etab;   // the workgroup's LDS copy of it (set in the kernel)
};

struct StepConsts {
    double dt, afac, hdt;
    double bconst, bfield;           // omega_i = bconst * Z_i * bfield / A_i
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

// Source charge as staged in LDS.  The mixed-precision ion forms read it as
// two floats: the staging copy keeps (float)Z in the low 32 bits of the w slot
// and (float)Z * muf in the high 32 bits, once per staged source, so that a
// mixture pair forms Z_j (1 + mu r) with one FMA.  Every other form keeps the
// double.
template <int KIND, int PREC>
__device__ __forceinline__ double stage_w(double z, float muf)
{
    if constexpr (PREC == 1 && KIND != NN) {
        const float zf = (float)z;
        return __builtin_bit_cast(
            double, ((unsigned long long)__builtin_bit_cast(unsigned, zf * muf) << 32) |
                        (unsigned long long)__builtin_bit_cast(unsigned, zf));
    } else {
        return z;
    }
}

__device__ __forceinline__ float staged_zf(double w)
{
    return __builtin_bit_cast(float, (unsigned)__builtin_bit_cast(unsigned long long, w));
}

__device__ __forceinline__ float staged_zmuf(double w)
{
    return __builtin_bit_cast(float, (unsigned)(__builtin_bit_cast(unsigned long long, w) >> 32));
}

// Pair interaction given the minimum-image displacement d = x_i - x_j.
// PREC 0 = FP64, 1 = mixed (FP64 displacement, r^2, cutoff test and
// accumulation; FP32 1/r, exp and Gaussians).  zj is the source's staged w
// (stage_w).  Callers never pass padded sources.
// SELF_BY_R2: the self pair is recognised by r^2 == 0 instead of the caller's
// flag (pair-list kernel: a source is its own target exactly when d = 0).
template <int KIND, int PREC, bool MEAS, bool SELF_BY_R2 = false>
__device__ __forceinline__ void interact(const ForceConsts &c, double zi, double zj,
                                         double dx, double dy, double dz, bool self,
                                         Acc<MEAS> &acc)
{
    // Ion pairs from lists: r^2 carries a floor below the last bit of any
    // physical r^2 (1e-200 in FP64; 1e-20 in the mixed form, whose FP32 1/r^3
    // must stay finite: the closest ions are fm apart, r^2 >= 1), so every
    // other pair's r^2 is unchanged.  The self pair (d = 0 exactly) then has
    // r^2 = floor and a finite factor times d = 0: an exact zero force
    // without a compare; a measuring step still excludes it by r^2 for the
    // energy.
    constexpr bool SOFT_SELF = SELF_BY_R2 && KIND != NN;
    double r2;
    if constexpr (SOFT_SELF) {
        const double fl = (PREC == 0) ? c.ec.r2_floor : c.ec.r2_floor_mixed;
        r2 = __builtin_fma(dx, dx, __builtin_fma(dy, dy, __builtin_fma(dz, dz, fl)));
        if constexpr (MEAS) self = !(r2 > 10.0 * fl);
    } else {
        r2 = __builtin_fma(dx, dx, __builtin_fma(dy, dy, dz * dz));
        if constexpr (SELF_BY_R2) self = self || !(r2 > 0.0);
    }
    if constexpr (KIND != NN) {
        // Guarded: a wave whose lanes are all outside the cutoff skips the
        // evaluation (measured faster than a branch-free select form, since
        // culled tiles and list entries leave many such waves).
        if (r2 <= c.rc2 && !self) {
            double s, phi = 0.0;
            if constexpr (PREC == 0) {
                double rinv = rsqrt64(r2, c.ec);
                double r = r2 * rinv;
                double mr = c.mu * r;
                double ex = exp_neg64t(-mr, c.ec, c.etab);
                // (1 + mu r) e^{-mu r} / r^3, the same operations whether or
                // not the step measures
                double fc = (__builtin_fma(mr, ex, ex) * rinv) * (rinv * rinv);
                s = (KIND == MIX) ? zj * fc : fc;
                if constexpr (MEAS) {
                    phi = ex * rinv - c.evcut;
                    if (KIND == MIX) phi *= zj;
                }
            } else {
                const float r2f = (float)r2;
                const float rinv = rsqrt32(r2f);
                const float r = r2f * rinv;
                const float ex = exp2_32(__builtin_fmaf(-c.mul2ef, r, -c.mul2e_lo * r));
                const float exr = ex * rinv;
                // Z_j (1 + mu r) for a mixture source (staged Z_j mu and Z_j)
                const float zl = (KIND == MIX) ? __builtin_fmaf(staged_zmuf(zj), r, staged_zf(zj))
                                               : __builtin_fmaf(c.muf, r, 1.0f);
                const float fc = zl * exr * (rinv * rinv);
                s = (double)fc;
                if constexpr (MEAS) {
                    phi = (double)exr - c.evcut;
                    if constexpr (KIND == MIX) phi *= (double)staged_zf(zj);
                }
            }
            accumulate<MEAS>(acc, s, dx, dy, dz, r2, phi);
        }
    } else {
        bool hn = c.nuc_on && r2 <= c.rn2;
        bool hc = c.coul_on && r2 <= c.rc2 && zi == 1.0 && zj == 1.0;
        if ((hn || hc) && !self) {
            double s = 0.0, phi = 0.0;
            const bool same = (zi == zj);
            if constexpr (PREC == 0) {
                if (hn) {
                    double g = exp_neg64t(-r2 * c.xp2i, c.ec, c.etab);
                    double bc = same ? c.bpc : c.bmc;
                    s = c.xpi * __builtin_fma(c.aa2, g, bc) * g;
                    if constexpr (MEAS) phi = __builtin_fma(c.aa, g, bc) * g;
                }
                if (hc) {
                    double rinv = rsqrt64(r2, c.ec);
                    double r = r2 * rinv;
                    double mr = c.mu * r;
                    double ex = exp_neg64t(-mr, c.ec, c.etab);
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
                    float ex = exp2_32(__builtin_fmaf(-c.mul2ef, r, -c.mul2e_lo * r));
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

// A value equal in every lane of the wave, moved to an SGPR pair.
__device__ __forceinline__ double uniform(double v)
{
    int lo = __builtin_amdgcn_readfirstlane(__double2loint(v));
    int hi = __builtin_amdgcn_readfirstlane(__double2hiint(v));
    return __hiloint2double(hi, lo);
}

// ---------------------------------------------------------------------------
// Tiled all-pairs partial sums.  grid = (ceil(n/BLOCK), S chunks).
template <int KIND, int PREC, bool MEAS>
__global__ __launch_bounds__(BLOCK) void k_pairs(const double4 *__restrict__ pos,
                                                 const TileBox *__restrict__ box,
                                                 ForceConsts c, int stages_per_chunk,
                                                 double *__restrict__ part,
                                                 unsigned long long *__restrict__ swept)
{
    int nswept = 0;   // source tiles this wave evaluated
    __shared__ double4 sj[STAGE];
    const int n = c.n;
    if constexpr (PREC == 0) pin_vgpr(c.ec.shifter);
    const int lane = threadIdx.x & (TILE - 1);
    if constexpr (PREC == 0) {
        __shared__ double etab_s[EXP_TN];
        load_exp_table(etab_s, c.etab_g);
        c.etab = etab_s;
    }
    const int itile = __builtin_amdgcn_readfirstlane(blockIdx.x * (BLOCK / TILE) +
                                                     (int)(threadIdx.x / TILE));
    const int i = itile * TILE + lane;
    const int chunk = blockIdx.y;
    const int nstages = (n + STAGE - 1) / STAGE;
    const double nanv = __builtin_nan("");
    const double4 pi = (i < n) ? pos[i] : make_double4(nanv, nanv, nanv, 0.0);
    const bool wave_live = itile < c.ntiles;
    double ic[3] = {0, 0, 0}, ih[3] = {0, 0, 0};
    if (wave_live) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            ic[k] = uniform(box[itile].c[k]);
            ih[k] = uniform(box[itile].h[k]);
        }
    }
    Acc<MEAS> acc;
    const int s0 = chunk * stages_per_chunk;
    const int s1 = min(s0 + stages_per_chunk, nstages);
    for (int st = s0; st < s1; ++st) {
        const int jb = st * STAGE;
        const int jl = jb + threadIdx.x;
        // padded sources are NaN and never evaluated (loops stop at n)
        double4 pjs = (jl < n) ? pos[jl] : make_double4(nanv, nanv, nanv, 0.0);
        pjs.w = stage_w<KIND, PREC>(pjs.w, c.muf);
        sj[threadIdx.x] = pjs;
        __syncthreads();
        if (wave_live) {
#pragma unroll 1
            for (int q = 0; q < STAGE / TILE; ++q) {
                const int jt = st * (STAGE / TILE) + q;
                if (jt >= c.ntiles) break;
                // AABB culling and image selection, wave-uniform
                double gap2 = 0.0, shift[3];
                bool fast = true;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    double cd = ic[k] - box[jt].c[k];
                    double h = ih[k] + box[jt].h[k] + c.margin;
                    double s = c.L[k] * __builtin_rint(cd * c.invL[k]);
                    double cm = __builtin_fabs(cd - s);
                    double g = cm - h;
                    g = g > 0.0 ? g : 0.0;
                    gap2 = __builtin_fma(g, g, gap2);
                    fast = fast && (cm + h < c.halfL_m[k]);
                    shift[k] = s;
                }
                gap2 = uniform(gap2);
                if (gap2 > c.cull2) continue;
                ++nswept;
                const int self_jj = (jt == itile) ? lane : -1;
                const double4 *src = sj + q * TILE;
                const int jcount = min(TILE, n - jt * TILE);
                if (uniform(fast ? 1.0 : 0.0) != 0.0) {
                    const double xs = pi.x - uniform(shift[0]);
                    const double ys = pi.y - uniform(shift[1]);
                    const double zs = pi.z - uniform(shift[2]);
#pragma unroll 4
                    for (int jj = 0; jj < jcount; ++jj) {
                        const double4 pj = src[jj];
                        interact<KIND, PREC, MEAS>(c, pi.w, pj.w, xs - pj.x, ys - pj.y,
                                                   zs - pj.z, jj == self_jj, acc);
                    }
                } else {
#pragma unroll 2
                    for (int jj = 0; jj < jcount; ++jj) {
                        const double4 pj = src[jj];
                        interact<KIND, PREC, MEAS>(
                            c, pi.w, pj.w, min_image(pi.x - pj.x, c.L[0], c.invL[0]),
                            min_image(pi.y - pj.y, c.L[1], c.invL[1]),
                            min_image(pi.z - pj.z, c.L[2], c.invL[2]), jj == self_jj, acc);
                    }
                }
            }
        }
        __syncthreads();
    }
    if (lane == 0 && nswept > 0) atomicAdd(swept, (unsigned long long)nswept * TILE * TILE);
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
// Cluster-pair lists.  For every target tile t the list holds the source
// slots whose minimum-image distance to the tile's AABB is within the list
// radius rl = cutoff + skin.  An entry is the slot index (bits 0-25) and, per
// axis, the periodic image of the source relative to the tile centre
// (bits 26+2k: 0 = -L, 1 = 0, 2 = +L).  "Fast" entries -- every target of the
// tile sees the source through that one image as long as no pair's
// separation has changed by more than the skin -- are stored from the front
// of the tile's region; the rest ("general", minimum image per pair) from the
// back.  The order is deterministic (by slot).  A list stays valid while the
// two largest displacements since its build add up to at most the skin
// (lists_stale); the drift kernel tracks them and the build kernel runs only
// when that bound is exceeded (or after a re-sort).
struct ListConsts {
    double L[3], invL[3], halfL_m[3];
    double rl2;          // (cutoff + skin)^2
    double rl2n;         // the same for a pair not both charged (nuclear radius)
    double skin;
    int n, ntiles, cap;  // cap = entries per tile region (= n)
    int pad;
    // FP32 copies for the cluster-list candidate test, widened so that the
    // float test keeps every candidate and only calls an entry fast when the
    // double test would: rl2f = (rl (1 + 5e-6) + 1e-6 max L)^2 (the float
    // offsets are within a few ulp of L); halfLf_m = halfL_m - 1e-5 L
    float Lf[3], invLf[3], halfLf_m[3], rl2f, skinf;
    // the same for sources that interact only through the nucleon model's
    // nuclear term (neutrons): (nuclear radius + skin)^2, widened alike
    // (rl2f for ions)
    float rl2f_n;
    // cluster lists of the mixed ion form: sources whose distance to the
    // cluster's AABB is at most sqrt(rn2f) (the near field) are stored with
    // the general entries, which the mixed pair kernels evaluate in FP64;
    // rn2f < 0: no near field
    float rn2f;
};


// Near field of the mixed ion form's cluster lists: sources within the near
// radius (plus the skin, so those pairs stay near until the next build) of a
// cluster's AABB are evaluated in FP64.  Per-pair FP32 errors are about 1e-7
// of a pair force and add up in a particle's force like a random walk, with
// weight sum over shells of r^2 F(r)^2 (F = (1 + mu r) e^{-mu r} / r^2 in a
// uniform fluid); the nearest pairs carry most of it, and a particle whose net
// force is a small fraction of them would otherwise see its relative error
// grow past 1e-6.  The near radius leaves a fraction NEAR_TAIL of that weight
// (between 0.5 a and the cutoff, a = (3 / (4 pi rho))^(1/3)) to FP32:
// 0.0019 is the fraction beyond 4 a at mu a = 0.7, where 4 a brings the
// largest relative force error of 27,648 ions to about 3e-7; the radius is
// then 5.1 a at mu a = 0.5, 3.1 a at 1 and 1.9 a at 2.  Returns -1 (no near
// field) for FP64 systems, nucleons, or no density.
constexpr double NEAR_TAIL = 0.0019;

inline double near_radius(int sim_type_is_ion, int mixed, double rho, double mu, double rc,
                          double skin)
{
    if (!sim_type_is_ion || !mixed || !(rho > 0.0)) return -1.0;
    const double a = std::cbrt(3.0 / (4.0 * 3.14159265358979323846 * rho));
    const double k = mu * a, x0 = 0.5, xc = std::min(std::max(rc / a, x0 * 1.001), 1e9);
    // the integral of x^2 F(x)^2 dx = x^3 F(x)^2 d(ln x) on a logarithmic grid
    // over [x0, xc] (Simpson per interval), accumulated from the top
    const int NI = 4096;
    const double h = std::log(xc / x0) / NI;
    auto w = [&](double t) {
        const double x = x0 * std::exp(t);
        const double f = (1.0 + k * x) * std::exp(-k * x) / (x * x);
        return x * x * x * f * f;
    };
    std::vector<double> tail(NI + 1, 0.0);
    for (int i = NI - 1; i >= 0; --i) {
        const double ta = i * h, tb = ta + h;
        tail[i] = tail[i + 1] + h / 6.0 * (w(ta) + 4.0 * w(0.5 * (ta + tb)) + w(tb));
    }
    const double want = NEAR_TAIL * tail[0];
    int i = 0;
    while (i < NI && tail[i] > want) ++i;
    return x0 * std::exp(i * h) * a + skin;
}

__device__ __forceinline__ unsigned long long as_bits(double v)
{
    return (unsigned long long)__double_as_longlong(v);
}

// Pair-list validity.  The lists hold every pair within cutoff + skin at
// their build; since then a pair's separation has changed by at most
// |d_i| + |d_j| <= d1 + d2, the two largest displacements of distinct
// particles since the build, so the lists stay valid while d1 + d2 <= skin.
// md[0] and md[1] hold d1^2 and d2^2 as the bit patterns of non-negative
// doubles (whose integer order is their numeric order); md[0] = +inf forces
// a rebuild.  The sum is formed in FP32 and enlarged by 1e-5, far more than
// its rounding, so the test can only call lists stale early, never late
// (every kernel uses this one test, so they agree).
__device__ __forceinline__ bool lists_stale(const unsigned long long *md, double skin)
{
    // hardware square roots (within a few ulp)
    const float s = __builtin_amdgcn_sqrtf((float)__longlong_as_double((long long)md[0])) +
                    __builtin_amdgcn_sqrtf((float)__longlong_as_double((long long)md[1]));
    return !(s * 1.00001f <= (float)skin);
}

// Adds the squared displacements of one workgroup (d2 >= 0 per thread) to
// the top-two pair md[0] >= md[1] that its wave names (md == nullptr: the
// wave adds nothing).  Each wave reduces its values to their largest m1 and
// second-largest m2, and the first thread merges the waves that name the same
// md into one such pair (b1, b2) before it updates md: b1 enters md[0];
// whichever of b1 and md[0]'s previous value is smaller enters md[1], as does
// b2.  Every value but one copy of the final maximum is at most the second
// largest, and the second largest always reaches md[1] (either with its own
// pair, or as the loser against the maximum), so md[1] ends as the
// second-largest displacement.  One update per workgroup and md keeps the
// atomics few.  Called by every thread of a BLOCK-thread workgroup.
__device__ __forceinline__ void record_disp(unsigned long long *md, double d2)
{
    constexpr int NW = BLOCK / 64;
    __shared__ double top_s[NW][2];
    __shared__ unsigned long long *md_s[NW];
    double m1 = d2;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(m1, off, 64);
        m1 = o > m1 ? o : m1;
    }
    const int lane = threadIdx.x & 63;
    const unsigned long long at = __ballot(d2 == m1);
    double m2 = (at != 0ull && lane == __ffsll((unsigned long long)at) - 1) ? 0.0 : d2;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(m2, off, 64);
        m2 = o > m2 ? o : m2;
    }
    const int w = threadIdx.x >> 6;
    if (lane == 0) {
        top_s[w][0] = m1;
        top_s[w][1] = m2;
        md_s[w] = md;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    for (int w0 = 0; w0 < NW; ++w0) {
        unsigned long long *m = md_s[w0];
        bool first = m != nullptr;
        for (int k = 0; k < w0; ++k) first = first && md_s[k] != m;
        if (!first) continue;   // none, or merged with an earlier wave's
        double b1 = 0.0, b2 = 0.0;
        for (int k = w0; k < NW; ++k) {
            if (md_s[k] != m) continue;
            const double a1 = top_s[k][0], a2 = top_s[k][1];
            if (a1 > b1) {
                b2 = a2 > b1 ? a2 : b1;
                b1 = a1;
            } else {
                b2 = a1 > b2 ? a1 : b2;
            }
        }
        if (b1 > 0.0) {
            const unsigned long long u1 = as_bits(b1);
            const unsigned long long old = atomicMax(&m[0], u1);
            const unsigned long long lose = old < u1 ? old : u1;
            if (lose) atomicMax(&m[1], lose);
            if (b2 > 0.0) atomicMax(&m[1], as_bits(b2));
        }
    }
}

// One workgroup per target tile; sources scanned in slot order, 256 at a time.
static __global__ __launch_bounds__(BLOCK) void k_list_build(const double4 *__restrict__ pos,
                                                      const TileBox *__restrict__ box,
                                                      ListConsts lc,
                                                      const unsigned long long *__restrict__ maxdisp,
                                                      unsigned *__restrict__ list,
                                                      int2 *__restrict__ counts,
                                                      double4 *__restrict__ xref)
{
    if (!lists_stale(maxdisp, lc.skin)) return;   // list still valid (uniform)
    const int t = blockIdx.x;
    const int lane = threadIdx.x & (TILE - 1), wv = threadIdx.x / TILE;
    __shared__ int wf[BLOCK / TILE], wg[BLOCK / TILE];
    double c[3], h[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) { c[k] = box[t].c[k]; h[k] = box[t].h[k]; }
    unsigned *mine = list + (size_t)t * lc.cap;
    int nf = 0, ng = 0;
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int jb = 0; jb < lc.n; jb += BLOCK) {
        const int j = jb + threadIdx.x;
        bool cand = false, fast = true;
        unsigned code = 0;
        if (j < lc.n) {
            const double4 pj = pos[j];
            const double xj[3] = {pj.x, pj.y, pj.z};
            double d2 = 0.0;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                double cd = c[k] - xj[k];
                double m = __builtin_rint(cd * lc.invL[k]);
                double e = __builtin_fabs(cd - lc.L[k] * m);
                double g = e - h[k];
                g = g > 0.0 ? g : 0.0;
                d2 = __builtin_fma(g, g, d2);
                // positions outside [0, L) can give |m| > 1: those take the
                // per-pair minimum image and keep a zero image code
                const bool m_ok = (m >= -1.0) && (m <= 1.0);
                fast = fast && m_ok && (e + h[k] + lc.skin < lc.halfL_m[k]);
                code |= (unsigned)(m_ok ? (int)m + 1 : 1) << (26 + 2 * k);
            }
            cand = d2 <= lc.rl2;
        }
        const unsigned long long bf = __ballot(cand && fast);
        const unsigned long long bg = __ballot(cand && !fast);
        if (lane == 0) {
            wf[wv] = __popcll(bf);
            wg[wv] = __popcll(bg);
        }
        __syncthreads();
        int of = nf, og = ng, tf = 0, tg = 0;
#pragma unroll
        for (int w = 0; w < BLOCK / TILE; ++w) {
            if (w < wv) { of += wf[w]; og += wg[w]; }
            tf += wf[w];
            tg += wg[w];
        }
        if (cand && fast) mine[of + __popcll(bf & lt)] = (unsigned)j | code;
        if (cand && !fast) mine[lc.cap - 1 - (og + __popcll(bg & lt))] = (unsigned)j | code;
        nf += tf;
        ng += tg;
        __syncthreads();
    }
    if (threadIdx.x == 0) counts[t] = make_int2(nf, ng);
    if (threadIdx.x < TILE) {
        const int i = t * TILE + threadIdx.x;
        if (i < lc.n) xref[i] = pos[i];
    }
}

// Pair sums over the lists.  Wave = one (tile, item) with item m of M
// covering an equal share of the tile's list.  Sources of each batch of 64
// entries are gathered by the lanes into the wave's own LDS slice, already
// shifted to their image, then read as broadcasts.  Self pairs are the
// pairs with r^2 = 0 (the source is the target itself); coincident distinct
// particles, where the reference divides by zero, are skipped the same way.
template <int KIND, int PREC, bool MEAS>
__global__ __launch_bounds__(BLOCK) void k_pairs_list(const double4 *__restrict__ pos,
                                                      const unsigned *__restrict__ list,
                                                      const int2 *__restrict__ counts,
                                                      ForceConsts c, int cap, int M,
                                                      double *__restrict__ part,
                                                      unsigned long long *__restrict__ maxdisp,
                                                      double skin)
{
    __shared__ double4 sj[BLOCK / TILE][TILE];
    const int n = c.n;
    if constexpr (PREC == 0) pin_vgpr(c.ec.shifter);
    // the build kernel (previous launch) has consumed the displacement flag
    if (blockIdx.x == 0 && threadIdx.x == 0 && lists_stale(maxdisp, skin)) {
        maxdisp[0] = 0ull;
        maxdisp[1] = 0ull;
    }
    const int lane = threadIdx.x & (TILE - 1);
    const int wv = threadIdx.x / TILE;
    if constexpr (PREC == 0) {
        __shared__ double etab_s[EXP_TN];
        load_exp_table(etab_s, c.etab_g);
        c.etab = etab_s;
    }
    const int item = __builtin_amdgcn_readfirstlane(blockIdx.x * (BLOCK / TILE) + wv);
    const int t = item / M, m = item % M;
    if (t >= c.ntiles) return;   // whole wave; this kernel has no block barrier
    const int i = t * TILE + lane;
    const double nanv = __builtin_nan("");
    const double4 pi = (i < n) ? pos[i] : make_double4(nanv, nanv, nanv, 0.0);
    const int2 cnt = counts[t];
    const int nf = __builtin_amdgcn_readfirstlane(cnt.x);
    const int total = __builtin_amdgcn_readfirstlane(cnt.x + cnt.y);
    const int per = (total + M - 1) / M;
    const int lo = min(m * per, total), hi = min(lo + per, total);
    const unsigned *mine = list + (size_t)t * cap;
    Acc<MEAS> acc;
    double4 *sw = sj[wv];
    // Software pipeline over batches of 64 entries: while batch k is computed
    // from LDS, the positions of batch k+1 and the entries of batch k+2 are
    // in flight in registers.
    auto entry_at = [&](int v) -> unsigned {
        return (v < hi) ? ((v < nf) ? mine[v] : mine[cap - 1 - (v - nf)]) : 0u;
    };
    unsigned e_cur = entry_at(lo + lane);
    double4 pj_cur = pos[e_cur & SLOT_MASK];
    unsigned e_nxt = entry_at(lo + TILE + lane);
    for (int v0 = lo; v0 < hi; v0 += TILE) {
        const int v = v0 + lane;
        if (v < hi) {
            double4 pj = pj_cur;
            if (v < nf) {   // apply the entry's periodic image
                pj.x += c.L[0] * (double)((int)((e_cur >> 26) & 3u) - 1);
                pj.y += c.L[1] * (double)((int)((e_cur >> 28) & 3u) - 1);
                pj.z += c.L[2] * (double)((int)((e_cur >> 30) & 3u) - 1);
            }
            pj.w = stage_w<KIND, PREC>(pj.w, c.muf);
            sw[lane] = pj;
        }
        __builtin_amdgcn_wave_barrier();
        // issue the next batch's loads before computing this one
        e_cur = e_nxt;
        pj_cur = pos[e_cur & SLOT_MASK];
        e_nxt = entry_at(v0 + 2 * TILE + lane);
        const int cntb = min(TILE, hi - v0);
        const int nfast = max(0, min(cntb, nf - v0));
        // the next source is read from LDS before the current pair is
        // computed, so the LDS latency overlaps the arithmetic
        double4 pn = sw[0];
#pragma unroll 4
        for (int jj = 0; jj < nfast; ++jj) {
            const double4 pj = pn;
            pn = sw[min(jj + 1, TILE - 1)];
            interact<KIND, PREC, MEAS, true>(c, pi.w, pj.w, pi.x - pj.x, pi.y - pj.y,
                                             pi.z - pj.z, false, acc);
        }
#pragma unroll 2
        for (int jj = nfast; jj < cntb; ++jj) {
            const double4 pj = pn;
            pn = sw[min(jj + 1, TILE - 1)];
            interact<KIND, PREC, MEAS, true>(c, pi.w, pj.w,
                                             min_image(pi.x - pj.x, c.L[0], c.invL[0]),
                                             min_image(pi.y - pj.y, c.L[1], c.invL[1]),
                                             min_image(pi.z - pj.z, c.L[2], c.invL[2]),
                                             false, acc);
        }
        __builtin_amdgcn_wave_barrier();
    }
    constexpr int NC = MEAS ? 11 : 3;
    double *o = part + ((size_t)t * M + m) * NC * TILE + lane;
    o[0] = acc.fx; o[TILE] = acc.fy; o[2 * TILE] = acc.fz;
    if constexpr (MEAS) {
        o[3 * TILE] = acc.e; o[4 * TILE] = acc.w;
        o[5 * TILE] = acc.pxx; o[6 * TILE] = acc.pxy; o[7 * TILE] = acc.pxz;
        o[8 * TILE] = acc.pyy; o[9 * TILE] = acc.pyz; o[10 * TILE] = acc.pzz;
    }
}

// Directed pairs a k_pairs_list launch over the current lists evaluates: per
// item 64 x its share of the tile's list; one thread per item, a block sum,
// one atomic per block.
static __global__ __launch_bounds__(BLOCK) void k_list_swept(const int2 *__restrict__ counts,
                                                            int ntiles, int M,
                                                            unsigned long long *__restrict__ swept)
{
    const int item = blockIdx.x * BLOCK + threadIdx.x;
    unsigned long long v = 0;
    if (item < ntiles * M) {
        const int t = item / M, m = item % M;
        const int2 cnt = counts[t];
        int lo = 0, hi = 0;
        list_share(cnt.x + cnt.y, M, m, lo, hi);
        v = (unsigned long long)(hi - lo) * TILE;
    }
    __shared__ unsigned long long red[BLOCK];
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = BLOCK / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0 && red[0] > 0) atomicAdd(swept, red[0]);
}

static __global__ void k_set_u64(unsigned long long *p, unsigned long long v) { *p = v; }

// ---------------------------------------------------------------------------
// Wave-level min/max and the per-tile AABB of 64 consecutive particles.
__device__ __forceinline__ void wave_minmax(double &lo, double &hi)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        double l2 = __shfl_xor(lo, off, 64);
        double h2 = __shfl_xor(hi, off, 64);
        lo = l2 < lo ? l2 : lo;
        hi = h2 > hi ? h2 : hi;
    }
}

__device__ __forceinline__ void tile_box_store(const double x[3], bool valid, int tile,
                                               TileBox *__restrict__ box)
{
    double c[3], h[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        double lo = valid ? x[k] : __builtin_inf();
        double hi = valid ? x[k] : -__builtin_inf();
        wave_minmax(lo, hi);
        c[k] = 0.5 * (lo + hi);
        h[k] = 0.5 * (hi - lo);
    }
    if ((threadIdx.x & (TILE - 1)) == 0) {
        TileBox b;
#pragma unroll
        for (int k = 0; k < 3; ++k) { b.c[k] = c[k]; b.h[k] = h[k]; }
        box[tile] = b;
    }
}

static __global__ __launch_bounds__(BLOCK) void k_tile_boxes(const double4 *__restrict__ pos, int n,
                                                      int ntiles, TileBox *__restrict__ box)
{
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    const int tile = i / TILE;
    double x[3] = {0, 0, 0};
    const bool valid = i < n;
    if (valid) { double4 p = pos[i]; x[0] = p.x; x[1] = p.y; x[2] = p.z; }
    if (tile < ntiles) tile_box_store(x, valid, tile, box);
}

// Whether each 64-slot tile holds a particle of nonzero charge (one wave per
// tile; slots keep their particles between re-sorts, so this is computed
// after each one).
static __global__ __launch_bounds__(BLOCK) void k_tile_charge(const double4 *__restrict__ pos,
                                                              int n, int ntiles,
                                                              int *__restrict__ tchg)
{
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    const int tile = i / TILE;
    const bool q = i < n && pos[i].w != 0.0;
    const unsigned long long m = __ballot(q);
    if (tile < ntiles && (threadIdx.x & (TILE - 1)) == 0) tchg[tile] = m != 0ull ? 1 : 0;
}

// ---------------------------------------------------------------------------
// Cluster lists.  A cluster is 16 consecutive sorted slots; its list holds
// the source slots whose minimum-image distance to the cluster's AABB is
// within the list radius rl = cutoff + skin, with the same entry encoding as
// the tile lists (slot | per-axis image code).  A 16-particle cluster's AABB
// is much smaller than a 64-particle tile's: about 30 % fewer candidate
// pairs.  Fast entries (one image for every target of the cluster while no
// pair's separation has changed by more than the skin) fill the cluster's
// region from the front, general ones from the back, both in slot order.
// Cluster size, segments and the size limit: yk_layout.hpp.

// Candidate test of the cluster-list builds.  A workgroup stages every
// candidate source once: its offset from the target tile's centre (the FP64
// difference rounded to float, then reduced to the minimum image in FP32: an
// error of a few float ulps of L, far inside the 1e-6 L and 1e-5 L widening of
// the test's bounds), and in .w the entry's
// per-axis image code for the shift that reduction applied (bits 26-31; the
// source is seen at x + (code - 1) L), bit 0 set when a shift lies beyond one
// box (such a source gives general entries only), bit 1 set for a source of
// charge 0 (a neutron of the nucleon model, listed only out to the nuclear
// radius).  An absent source (a slot beyond the system) is staged FAR_AWAY
// on every axis with bit 0 set.
constexpr unsigned CODE_MASK = ~SLOT_MASK;
constexpr unsigned STAGE_NEUTRAL = 2u;
constexpr float FAR_AWAY = 3.0e38f;

__device__ __forceinline__ float4 clist_stage(const double4 &p, bool present, const double tc[3],
                                              const float Lf[3], const float invLf[3])
{
    const double x[3] = {p.x, p.y, p.z};
    float o[3];
    unsigned w = 0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float d = (float)(x[k] - tc[k]);
        const float m = __builtin_rintf(d * invLf[k]);
        o[k] = present ? __builtin_fmaf(-Lf[k], m, d) : FAR_AWAY;
        const bool ok = __builtin_fabsf(m) <= 1.0f;
        w |= (unsigned)(ok ? 1 - (int)m : 1) << (26 + 2 * k);
        w |= ok ? 0u : 1u;
    }
    w |= p.w == 0.0 ? STAGE_NEUTRAL : 0u;
    return make_float4(o[0], o[1], o[2], __uint_as_float(present ? w : 1u));
}

// A cluster as the candidate test sees it: centre offset from the tile centre
// (cr), widened half extents (hf), the largest |offset| per axis at which the
// staged image still serves every target for the list's lifetime
// (lim = L/2 - margin - hf - skin), and a wave-uniform tier:
//   2  |cr| + hf + list radius < L/2 on every axis (with margin): a source
//      whose offset d from the cluster centre exceeds L/2 (its nearer image
//      is at L - |d|) is beyond the list radius either way, so |d| serves
//      as the distance;
//   1  |cr| < L/4: every |d| < L, the minimum image distance is
//      min(|d|, L - |d|);
//   0  otherwise (a cluster far from its tile's centre): full minimum image,
//      general entries only.
struct ClusterView {
    float cr[3], hf[3], lim[3];
    // list radius^2 of charged and of neutral sources: ListConsts::rl2f and
    // rl2f_n, or rl2f_n for both when the cluster holds no charged particle
    // (in the nucleon model only proton pairs interact beyond the nuclear
    // radius)
    float rl2c, rl2n;
    int tier;
};

__device__ __forceinline__ ClusterView cluster_view(const double c[3], const double h[3],
                                                    bool live, const double tc[3],
                                                    const float Lf[3], const float halfLf_m[3],
                                                    float skinf, float rl2f, float rl2nf,
                                                    bool charged)
{
    ClusterView v;
    v.rl2c = charged ? rl2f : rl2nf;
    v.rl2n = rl2nf;
    const float rlf = __builtin_sqrtf(rl2f);
    bool t2 = live, t1 = live;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        // half extents in float, widened by 1e-6 (more than one rounding), so
        // the float box contains the double one
        v.hf[k] = live ? (float)h[k] * 1.000001f : 0.0f;
        v.cr[k] = live ? (float)(c[k] - tc[k]) : 0.0f;
        v.lim[k] = halfLf_m[k] - v.hf[k] - skinf;
        const float acr = __builtin_fabsf(v.cr[k]);
        t2 = t2 && acr + v.hf[k] + rlf + 1e-5f * Lf[k] < 0.5f * Lf[k];
        t1 = t1 && acr < 0.25f * Lf[k];
    }
    v.tier = __builtin_amdgcn_readfirstlane(t2 ? 2 : t1 ? 1 : 0);   // wave-uniform
    return v;
}

// The candidate test of the staged source s against the cluster box (FP32
// with the widened bounds of ListConsts): cand, the source is within the list
// radius (v.rl2c or, for a neutral source, v.rl2n); fast, its staged image
// serves every target of the cluster for the list's lifetime; and the same as
// this wave's lane masks mcand, mfast.  Each
// mask is the ballot of a single compare (which compiles to the compare's own
// result), combined with scalar ANDs.  Absent sources fail at tiers 1 and 2
// (FAR_AWAY); tier 0 callers test presence.
template <int TIER>
__device__ __forceinline__ void clist_test(const ClusterView &v, const float4 &s, const float Lf[3],
                                           const float invLf[3], float rl2f, float rn2f,
                                           bool &cand, bool &fast, unsigned long long &mcand,
                                           unsigned long long &mfast)
{
    const float sx[3] = {s.x, s.y, s.z};
    float d2 = 0.0f;
    const bool wok = (__float_as_uint(s.w) & 1u) == 0u;
    fast = TIER > 0 && wok;
    mfast = TIER > 0 ? __builtin_amdgcn_ballot_w64(wok) : 0ull;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float cd = v.cr[k] - sx[k];
        const float a = __builtin_fabsf(cd);
        float ev;
        if constexpr (TIER == 2) {
            ev = a;
        } else if constexpr (TIER == 1) {
            ev = __builtin_fminf(a, __builtin_fabsf(Lf[k] - a));
        } else {
            ev = __builtin_fabsf(__builtin_fmaf(-Lf[k], __builtin_rintf(cd * invLf[k]), cd));
        }
        if constexpr (TIER > 0) {
            const bool in = a < v.lim[k];
            fast = fast & in;
            mfast &= __builtin_amdgcn_ballot_w64(in);
        }
        float gk = ev - v.hf[k];
        gk = gk > 0.0f ? gk : 0.0f;
        d2 = __builtin_fmaf(gk, gk, d2);
    }
    cand = d2 <= ((__float_as_uint(s.w) & STAGE_NEUTRAL) ? v.rl2n : v.rl2c);
    mcand = __builtin_amdgcn_ballot_w64(cand);
    // a near-field source is stored with the general entries (evaluated in
    // FP64 by the mixed kernels)
    const bool near = d2 <= rn2f;
    fast = fast & !near;
    mfast &= ~__builtin_amdgcn_ballot_w64(near);
}

// Lanes below this one whose bit is set in m.
__device__ __forceinline__ unsigned lanes_below(unsigned long long m)
{
    return __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

// One round of a cluster-list build for one wave: its cluster against the W
// staged source tiles of the round (tile slots base[w] + lane; slots at or
// beyond n are absent), entries appended to the segment at clist + seg_off
// (fast from the front, general from the back of its segcap entries; seg_off
// is wave-uniform).  The entry is the slot with the staged image code, which
// the pair kernels apply to fast entries only (general ones take the per-pair
// minimum image).
template <int TIER, int W>
__device__ __forceinline__ void clist_round(const ClusterView &v, const float4 (*st)[TILE],
                                            const int *base, int n, bool live, int lane,
                                            const float Lf[3], const float invLf[3], float rl2f,
                                            float rn2f, unsigned *__restrict__ clist,
                                            size_t seg_off, unsigned segcap, unsigned &nf,
                                            unsigned &ng)
{
#pragma unroll
    for (int w = 0; w < W; ++w) {
        const int j = base[w] + lane;
        const float4 sv = st[w][lane];
        bool cand, fast;
        unsigned long long mcand, mfast;
        clist_test<TIER>(v, sv, Lf, invLf, rl2f, rn2f, cand, fast, mcand, mfast);
        if constexpr (TIER == 0) {
            const bool present = j < n;
            cand = cand & live & present;
            mcand &= live ? __builtin_amdgcn_ballot_w64(present) : 0ull;
        }
        const unsigned long long bf = mcand & mfast, bg = mcand & ~mfast;
        if (cand) {
            const unsigned at = fast ? nf + lanes_below(bf) : segcap - 1u - (ng + lanes_below(bg));
            clist[seg_off + at] = (unsigned)j | (__float_as_uint(sv.w) & CODE_MASK);
        }
        nf += (unsigned)__popcll(bf);
        ng += (unsigned)__popcll(bg);
    }
}

// The scan of a cluster-list build for one wave: its cluster against every
// candidate source tile (ctiles[0..ncand), slots tile * 64 + lane of pos,
// those at or beyond n absent), the workgroup staging W tiles per round as
// clist_stage offsets (double buffered, the loads of the next two rounds in
// flight).  nf, ng: the fast and general entries written to the segment at dst.
template <int W>
__device__ __forceinline__ void clist_scan(const double4 *__restrict__ pos, int n,
                                           const int *ctiles, int ncand, const ClusterView &cv,
                                           bool live, const double tc[3], const float Lf[3],
                                           const float invLf[3], float rl2f, float rn2f,
                                           float4 (*stage)[W][TILE], unsigned *__restrict__ clist,
                                           size_t seg_off, int segcap, int lane, int wv, int &nf,
                                           int &ng)
{
    auto slot_of = [&](int r) -> int {
        const int idx = r * W + wv;
        return (idx < ncand) ? ctiles[idx] * TILE + lane : n;
    };
    auto load = [&](int j) -> double4 { return (j < n) ? pos[j] : make_double4(0, 0, 0, 0); };
    const int nrounds = (ncand + W - 1) / W;
    // wave-uniform segment offset (scalar registers); readfirstlane returns
    // int, so each half is widened as unsigned (a low half of 2^31 or more
    // sign-extended would put the segment 16 EiB away)
    const unsigned off_hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(seg_off >> 32));
    const unsigned off_lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)seg_off);
    seg_off = ((size_t)off_hi << 32) | (size_t)off_lo;
    int ja = slot_of(0), jb = slot_of(1);
    double4 pa = load(ja), pb = load(jb);
    unsigned uf = 0, ug = 0;
    for (int r = 0; r < nrounds; ++r) {
        stage[r & 1][wv][lane] = clist_stage(pa, ja < n, tc, Lf, invLf);
        pa = pb;
        ja = jb;
        jb = slot_of(r + 2);
        pb = load(jb);
        __syncthreads();
        int base[W];
#pragma unroll
        for (int w = 0; w < W; ++w) base[w] = (r * W + w < ncand) ? ctiles[r * W + w] * TILE : n;
        if (cv.tier == 2)
            clist_round<2, W>(cv, stage[r & 1], base, n, live, lane, Lf, invLf, rl2f, rn2f,
                              clist, seg_off, (unsigned)segcap, uf, ug);
        else if (cv.tier == 1)
            clist_round<1, W>(cv, stage[r & 1], base, n, live, lane, Lf, invLf, rl2f, rn2f,
                              clist, seg_off, (unsigned)segcap, uf, ug);
        else
            clist_round<0, W>(cv, stage[r & 1], base, n, live, lane, Lf, invLf, rl2f, rn2f,
                              clist, seg_off, (unsigned)segcap, uf, ug);
    }
    nf = (int)uf;
    ng = (int)ug;
}

// One workgroup per tile, wave w builds the list of the tile's cluster w.
// Source tiles are culled by the minimum-image gap between their AABB
// (current positions, from k_drift / k_tile_boxes) and the tile's; the four
// waves then load the surviving source tiles four at a time into LDS (double
// buffered, the loads of the next two rounds in flight) and each tests its
// own cluster against every staged slot, so one read of a source tile serves
// the tile's four clusters (staged once per workgroup as clist_stage
// offsets).  Also records the positions the list was built from (xref).
// tchg[t]: whether tile t holds a charged particle (k_tile_charge); a tile
// pair not charged on both sides is culled at the nuclear radius (lc.rl2n).
// nullptr: every tile counts as charged.
static __global__ __launch_bounds__(BLOCK) void k_clist_build(
    const double4 *__restrict__ pos, const TileBox *__restrict__ box, ListConsts lc,
    int segcap, const unsigned long long *__restrict__ maxdisp, unsigned *__restrict__ clist,
    int2 *__restrict__ ccounts, double4 *__restrict__ xref, const int *__restrict__ tchg)
{
    if (!lists_stale(maxdisp, lc.skin)) return;   // lists still valid (uniform)
    constexpr int W = BLOCK / TILE;
    __shared__ float4 stage[2][W][TILE];
    __shared__ int ctiles[CLIST_MAX_N / TILE / NSEG + 1];
    const int t = blockIdx.x;
    const int seg = blockIdx.y;   // source tiles st with st % NSEG == seg
    const int lane = threadIdx.x & (TILE - 1), wv = threadIdx.x / TILE;
    const int cl = t * CPW + wv;
    const bool cvalid = cl * CL < lc.n;
    double tc[3], th[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) { tc[k] = box[t].c[k]; th[k] = box[t].h[k]; }
    // cluster AABB (lanes 0..15 hold its slots) and whether it holds a
    // charged particle
    double c[3], h[3];
    bool charged;
    {
        const int i = cl * CL + lane;
        const bool v = lane < CL && i < lc.n;
        const double4 p = v ? pos[i] : make_double4(0, 0, 0, 0);
        if (v && seg == 0) xref[i] = p;
        charged = __ballot(v && p.w != 0.0) != 0ull;
        const double xi[3] = {p.x, p.y, p.z};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            double lo = v ? xi[k] : __builtin_inf(), hi = v ? xi[k] : -__builtin_inf();
            wave_minmax(lo, hi);
            c[k] = 0.5 * (lo + hi);
            h[k] = 0.5 * (hi - lo);
        }
    }
    const bool live = cvalid && h[0] >= 0.0;   // the cluster holds a particle
    const ClusterView cv = cluster_view(c, h, live, tc, lc.Lf, lc.halfLf_m, lc.skinf, lc.rl2f,
                                        lc.rl2f_n, charged);
    // candidate source tiles of the whole tile, in ascending order
    int ncand = 0;
    {
        const double rl = __builtin_sqrt(lc.rl2), rln = __builtin_sqrt(lc.rl2n);
        const bool tch = tchg ? tchg[t] != 0 : true;
        const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
        for (int tb = 0; tb < lc.ntiles; tb += TILE * NSEG) {
            const int st = tb + lane * NSEG + seg;
            bool tcand = false;
            if (st < lc.ntiles) {
                double gap2 = 0.0;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    double d = tc[k] - box[st].c[k];
                    d = __builtin_fma(-lc.L[k], __builtin_rint(d * lc.invL[k]), d);
                    double gk = __builtin_fabs(d) - th[k] - box[st].h[k];
                    gk = gk > 0.0 ? gk : 0.0;
                    gap2 = __builtin_fma(gk, gk, gap2);
                }
                const bool both = tch && (tchg ? tchg[st] != 0 : true);
                tcand = gap2 <= (both ? rl * rl : rln * rln);
            }
            const unsigned long long tm = __ballot(tcand);
            if (wv == 0 && tcand) ctiles[ncand + __popcll(tm & lt)] = st;
            ncand += __popcll(tm);
        }
    }
    __syncthreads();
    int nf, ng;
    clist_scan<W>(pos, lc.n, ctiles, ncand, cv, live, tc, lc.Lf, lc.invLf, lc.rl2f, lc.rn2f,
                  stage, clist, ((size_t)(cvalid ? cl : 0) * NSEG + seg) * segcap, segcap, lane,
                  wv, nf, ng);
    if (cvalid && lane == 0) ccounts[cl * NSEG + seg] = make_int2(nf, ng);
}

// Cluster order for the pair kernel: by list length, longest first (ties by
// index), so the four clusters a wave takes have nearly equal lists.  One
// workgroup of 1024 threads, bitonic sort of up to 4096 keys in LDS; larger
// systems sort k_corder_keys' keys with a device-wide radix sort.
constexpr int CORDER_THREADS = 1024;
constexpr int CORDER_MAX = 4096;
static __global__ __launch_bounds__(CORDER_THREADS) void k_corder(
    const int2 *__restrict__ ccounts, int ncl, ListConsts lc,
    const unsigned long long *__restrict__ maxdisp, int *__restrict__ order)
{
    if (!lists_stale(maxdisp, lc.skin)) return;
    __shared__ unsigned key[CORDER_MAX];
    int np = 1;
    while (np < ncl) np <<= 1;
    for (int k = threadIdx.x; k < np; k += CORDER_THREADS) {
        unsigned kk = 0xffffffffu;
        if (k < ncl) {
            unsigned len = 0;
            for (int sg = 0; sg < NSEG; ++sg) {
                const int2 cc = ccounts[k * NSEG + sg];
                len += (unsigned)(cc.x + cc.y);
            }
            kk = ((0x1ffffu - (len > 0x1ffffu ? 0x1ffffu : len)) << 12) | (unsigned)k;
        }
        key[k] = kk;
    }
    __syncthreads();
    for (int size = 2; size <= np; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int k = threadIdx.x; k < np; k += CORDER_THREADS) {
                const int j = k ^ stride;
                if (j > k) {
                    const bool up = (k & size) == 0;
                    const unsigned a = key[k], b = key[j];
                    if ((a > b) == up) { key[k] = b; key[j] = a; }
                }
            }
            __syncthreads();
        }
    }
    for (int k = threadIdx.x; k < ncl; k += CORDER_THREADS) order[k] = (int)(key[k] & 0xfffu);
}

// The same order for more than CORDER_MAX clusters: key 2^18 - 1 - list
// length (clamped) and value the cluster index, for a stable radix sort over
// 18 bits (ties keep index order, as in k_corder).
static __global__ __launch_bounds__(BLOCK) void k_corder_keys(const int2 *__restrict__ ccounts,
                                                              int ncl,
                                                              unsigned *__restrict__ key,
                                                              int *__restrict__ val)
{
    const int k = blockIdx.x * BLOCK + threadIdx.x;
    if (k >= ncl) return;
    unsigned len = 0;
    for (int sg = 0; sg < NSEG; ++sg) {
        const int2 cc = ccounts[k * NSEG + sg];
        len += (unsigned)(cc.x + cc.y);
    }
    key[k] = 0x3ffffu - (len > 0x3ffffu ? 0x3ffffu : len);
    val[k] = k;
}
constexpr int CORDER_KEY_BITS = 18;

// A cluster's list as its NSEG segments, read back entries first: each
// segment's region holds its fast entries from the front and the others
// (general ones and, in the mixed ion form, the near field) from the back; the
// list is the back entries of segments 0..3, then their fast entries, so that
// the entries the mixed kernels evaluate in FP64 come first in every list.
// entry(v): entry v and whether it is a back entry (0 at or past hi).
// SegList and seg_list (a cluster's list as its segments' entries, back
// entries first): yk_layout.hpp.

// The far pairs of the mixed ion form, FP32 throughout, in lengths scaled by
// fs = mu log2 e: target and source as FP32 scaled offsets from the cluster's
// centre (each the rounding of an FP64 product), displacement d', r'^2 (with a
// floor that keeps a self pair's 1/r'^3 finite), cutoff test, 1/r', exp and
// the force factor.  With r' = fs r, e^{-mu r} = 2^{-r'} and
// (1 + mu r) / r'^3 = (1/r')^2 (1/r' + ln 2), so the factor
// q = Z_j 2^{-r'} (1/r')^2 (1/r' + ln 2) takes no multiply by mu and q d' is
// the force over fs^2.  The force accumulates in FP32 over one 16-source batch
// (the caller adds fs^2 times it to the FP64 sums), the measured quantities in
// FP64 in unscaled units.  s.w: the source's Z.  An FP32 r'^2 within the band
// [rc2f_lo, rc2f_hi] of the scaled cutoff decides by the FP64 r^2 of the
// target pi and the source's FP64 image sd.
template <int KIND, bool MEAS>
__device__ __forceinline__ void interact_far(const ForceConsts &c, float xi, float yi, float zi,
                                             const float4 &s, const double4 &pi,
                                             const double4 *__restrict__ sd, float &fx, float &fy,
                                             float &fz, Acc<MEAS> &acc)
{
    constexpr float LN2F = 0.693147180559945309f;
    const float dx = xi - s.x, dy = yi - s.y, dz = zi - s.z;
    const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, __builtin_fmaf(dz, dz, c.r2_floor_f)));
    bool in = r2 <= c.rc2f_hi;
    if (in && r2 > c.rc2f_lo) {   // rare: FP32 cannot tell
        const double4 pj = *sd;
        const double ex = pi.x - pj.x, ey = pi.y - pj.y, ez = pi.z - pj.z;
        in = __builtin_fma(ex, ex, __builtin_fma(ey, ey, __builtin_fma(ez, ez, c.ec.r2_floor))) <=
             c.rc2;
    }
    if (in) {
        const float rinv = rsqrt32(r2);
        const float ex = exp2_32(-(r2 * rinv));
        float q = ((rinv * rinv) * (rinv + LN2F)) * ex;
        if constexpr (KIND == MIX) q *= s.w;
        fx = __builtin_fmaf(q, dx, fx);
        fy = __builtin_fmaf(q, dy, fy);
        fz = __builtin_fmaf(q, dz, fz);
        if constexpr (MEAS) {
            if (r2 > 10.0f * c.r2_floor_f) {   // not the self pair
                // Z_j e^{-mu r} / r = fs Z_j 2^{-r'} / r'; with d = d' / fs, the
                // virial and tensor terms F r^2 and d F d are fs q r'^2 and
                // fs d' q d'
                double phi = c.fs * (double)(ex * rinv) - c.evcut;
                if constexpr (KIND == MIX) phi *= (double)s.w;
                const double sd = (double)q * c.fs, ddx = dx, ddy = dy, ddz = dz;
                const double sx = sd * ddx, sy = sd * ddy, sz = sd * ddz;
                acc.e += phi;
                acc.w = __builtin_fma(sd, (double)r2, acc.w);
                acc.pxx = __builtin_fma(ddx, sx, acc.pxx);
                acc.pxy = __builtin_fma(ddx, sy, acc.pxy);
                acc.pxz = __builtin_fma(ddx, sz, acc.pxz);
                acc.pyy = __builtin_fma(ddy, sy, acc.pyy);
                acc.pyz = __builtin_fma(ddy, sz, acc.pyz);
                acc.pzz = __builtin_fma(ddz, sz, acc.pzz);
            }
        }
    }
}

// The pairs of a 16-lane group's target (pi, one per lane; NaN for none)
// with its share [lo, hi) of its cluster's list (sl), the wave stepping
// through its four groups' shares together for iters entries; sources are
// slots of pos.  Each group stages 16 of its sources per batch in its part of
// the wave's LDS slice sw (the next batch's entries and positions in flight);
// lanes past their share stage NaN sources, which every pair test rejects.
// A batch holding a back entry (general, or near field) takes the per-pair
// minimum image -- in FP64 also in the mixed ion form (PREC 1) --; the other
// batches of the mixed ion form run interact_far on FP32 offsets, scaled by
// fs, from the group's cluster centre cc.  Nucleons keep the double staging in
// both precisions.
template <int KIND, int PREC, bool MEAS>
__device__ __forceinline__ void clist_share(const ForceConsts &c, const double4 &pi,
                                            const double cc[3], const SegList &sl, int lo,
                                            int iters, const double4 *__restrict__ pos,
                                            double4 *__restrict__ sw, float4 *__restrict__ sf,
                                            SegList *__restrict__ hdr, int lane, Acc<MEAS> &acc)
{
    constexpr bool FAR32 = PREC == 1 && KIND != NN;
    const int g = lane / CL, gl = lane % CL;
    const int hi = sl.hi;
    const unsigned *const base = sl.base;
    // each lane walks its entries lo + gl, lo + gl + CL, ... with a cursor:
    // the offset steps by +-CL within a segment's run; where a run ends it is
    // recomputed from the list's header, which waits in the group's LDS slot
    // hdr[g] (every lane of the group writes the same value) and is read
    // there behind a compiler barrier, so that the segment counts hold no
    // registers across the loop
    hdr[g] = sl;
    __builtin_amdgcn_wave_barrier();
    SegList::Cursor cu = sl.cursor(lo + gl, CL);
    auto advance = [&](int vn) {
        if (vn >= cu.vend) {
            __asm__ volatile("" ::: "memory");
            cu = hdr[g].cursor(vn, CL);
        } else {
            cu.off += (unsigned)cu.step;
        }
    };
    auto fetch = [&](int vn, bool &back) -> unsigned {
        back = false;
        if (vn >= hi) return 0u;   // entry 0 (slot 0, a valid source) past the share
        back = cu.step < 0;
        return base[cu.off];
    };
    // the far batches' common origin: the cluster centre, scaled by fs and
    // rounded to FP32 (any point common to the group's targets and sources
    // serves; an FP32 one takes half the registers); offsets from it are
    // formed by one FP64 FMA each and rounded to FP32
    float xi_f = 0.f, yi_f = 0.f, zi_f = 0.f;
    float ccs[3] = {0.f, 0.f, 0.f};
    if constexpr (FAR32) {
        for (int k = 0; k < 3; ++k) ccs[k] = (float)(cc[k] * c.fs);
        xi_f = (float)__builtin_fma(pi.x, c.fs, -(double)ccs[0]);
        yi_f = (float)__builtin_fma(pi.y, c.fs, -(double)ccs[1]);
        zi_f = (float)__builtin_fma(pi.z, c.fs, -(double)ccs[2]);
    }
    bool g_cur, g_nxt;
    unsigned e_cur = fetch(lo + gl, g_cur);
    double4 pj_cur = pos[e_cur & SLOT_MASK];
    advance(lo + CL + gl);
    unsigned e_nxt = fetch(lo + CL + gl, g_nxt);
    for (int b0 = 0; b0 < iters; b0 += CL) {
        const int v = lo + b0 + gl;
        const bool live = v < hi;
        // a batch holding a back entry takes the per-pair minimum image for all
        // of its pairs (a no-op for the already shifted ones)
        const bool gen = __ballot(live && g_cur) != 0ull;
        double4 pj = pj_cur;
        if (live && !g_cur) {   // apply the entry's periodic image
            pj.x += c.L[0] * (double)((int)((e_cur >> 26) & 3u) - 1);
            pj.y += c.L[1] * (double)((int)((e_cur >> 28) & 3u) - 1);
            pj.z += c.L[2] * (double)((int)((e_cur >> 30) & 3u) - 1);
        }
        const double q = __builtin_nan("");
        if (FAR32 && !gen) {
            const float qf = (float)q;
            // the origin's FP64 values are formed here, not held across the
            // loop (the empty asm keeps the compiler from hoisting them)
            float o0 = ccs[0], o1 = ccs[1], o2 = ccs[2];
            __asm__("" : "+v"(o0), "+v"(o1), "+v"(o2));
            sf[lane] = live ? make_float4((float)__builtin_fma(pj.x, c.fs, -(double)o0),
                                          (float)__builtin_fma(pj.y, c.fs, -(double)o1),
                                          (float)__builtin_fma(pj.z, c.fs, -(double)o2),
                                          (float)pj.w)
                            : make_float4(qf, qf, qf, 0.f);
        }
        // the FP64 images: general batches, and the far batches' cutoff band
        pj.w = FAR32 ? pj.w : stage_w<KIND, PREC>(pj.w, c.muf);
        sw[lane] = live ? pj : make_double4(q, q, q, 0.0);
        __builtin_amdgcn_wave_barrier();
        e_cur = e_nxt;
        g_cur = g_nxt;
        pj_cur = pos[e_cur & SLOT_MASK];
        advance(v + 2 * CL);
        e_nxt = fetch(v + 2 * CL, g_nxt);
        const int kb = min(CL, iters - b0);
        // FP64 (the mixed ion form's too), minimum image per pair
        auto general_batch = [&]() {
            constexpr int GP = FAR32 ? 0 : PREC;
            const double4 *gs = sw + g * CL;
#pragma unroll 2
            for (int k = 0; k < kb; ++k) {
                const double4 pjk = gs[k];
                interact<KIND, GP, MEAS, true>(c, pi.w, pjk.w,
                                               min_image(pi.x - pjk.x, c.L[0], c.invL[0]),
                                               min_image(pi.y - pjk.y, c.L[1], c.invL[1]),
                                               min_image(pi.z - pjk.z, c.L[2], c.invL[2]),
                                               false, acc);
            }
        };
        if constexpr (FAR32) {
            if (!gen) {
                const float4 *gs = sf + g * CL;
                const double4 *gd = sw + g * CL;
                float fx = 0.f, fy = 0.f, fz = 0.f;
                float4 pn = gs[0];
                if (kb == CL) {
#pragma unroll
                    for (int k = 0; k < CL; ++k) {
                        const float4 pf = pn;
                        if (k + 1 < CL) pn = gs[k + 1];
                        interact_far<KIND, MEAS>(c, xi_f, yi_f, zi_f, pf, pi, gd + k, fx, fy, fz,
                                                 acc);
                    }
                } else {
#pragma unroll 4
                    for (int k = 0; k < kb; ++k) {
                        const float4 pf = pn;
                        pn = gs[min(k + 1, CL - 1)];
                        interact_far<KIND, MEAS>(c, xi_f, yi_f, zi_f, pf, pi, gd + k, fx, fy, fz,
                                                 acc);
                    }
                }
                acc.fx = __builtin_fma(c.fs2, (double)fx, acc.fx);
                acc.fy = __builtin_fma(c.fs2, (double)fy, acc.fy);
                acc.fz = __builtin_fma(c.fs2, (double)fz, acc.fz);
            } else {
                general_batch();
            }
        } else {
            if (!gen) {
                const double4 *gs = sw + g * CL;
                double4 pn = gs[0];
                if (kb == CL) {
                    // a full batch (all but a share's last): fully unrolled, so
                    // the LDS reads take immediate offsets and a pair needs no
                    // address arithmetic; the next source is read one pair ahead
#pragma unroll
                    for (int k = 0; k < CL; ++k) {
                        const double4 pjk = pn;
                        if (k + 1 < CL) pn = gs[k + 1];
                        interact<KIND, PREC, MEAS, true>(c, pi.w, pjk.w, pi.x - pjk.x,
                                                         pi.y - pjk.y, pi.z - pjk.z, false, acc);
                    }
                } else {
#pragma unroll 4
                    for (int k = 0; k < kb; ++k) {
                        const double4 pjk = pn;
                        pn = gs[min(k + 1, CL - 1)];
                        interact<KIND, PREC, MEAS, true>(c, pi.w, pjk.w, pi.x - pjk.x,
                                                         pi.y - pjk.y, pi.z - pjk.z, false, acc);
                    }
                }
            } else {
                general_batch();
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// The centre of the AABB of a 16-lane group's valid targets (every lane of
// the group gets it; 0 for a group without targets).
__device__ __forceinline__ void group_centre(const double4 &pi, bool valid, double cc[3])
{
    const double x[3] = {pi.x, pi.y, pi.z};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        double lo = valid ? x[k] : __builtin_inf(), hi = valid ? x[k] : -__builtin_inf();
#pragma unroll
        for (int off = 1; off < CL; off <<= 1) {
            const double l2 = __shfl_xor(lo, off, 64), h2 = __shfl_xor(hi, off, 64);
            lo = l2 < lo ? l2 : lo;
            hi = h2 > hi ? h2 : hi;
        }
        cc[k] = (lo <= hi) ? 0.5 * (lo + hi) : 0.0;
    }
}

// Pair sums over the cluster lists.  A wave takes four clusters (order[4s ..
// 4s+3]), 16 lanes each, one target per lane, and item m of M: an equal share
// of every cluster's list (clist_share).  The kernel issues no atomics: on
// this chip an atomic on one address from every CU is served one at a time
// beyond the XCDs' own L2s (about 25 ns each), so a per-wave counter or work
// queue costs as much as a short wave's pairs.  The pairs evaluated are
// counted by k_clist_swept, from the same list counts, when asked for.
template <int KIND, int PREC, bool MEAS>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(PREC == 1 && !MEAS ? 6 : 1)))
void k_pairs_clist(const double4 *__restrict__ pos,
                                                       const unsigned *__restrict__ clist,
                                                       const int2 *__restrict__ ccounts,
                                                       const int *__restrict__ order,
                                                       ForceConsts c, int segcap, int ncl, int M,
                                                       double *__restrict__ part,
                                                       unsigned long long *__restrict__ maxdisp,
                                                       double skin)
{
    __shared__ double4 sj[BLOCK / TILE][TILE];
    __shared__ float4 sfj[BLOCK / TILE][(PREC == 1 && KIND != NN) ? TILE : 1];
    __shared__ SegList hdr_s[BLOCK / TILE][CPW];
    constexpr bool TAB = PREC == 0 || KIND != NN;   // an FP64 pair path with the exp table
    const int n = c.n;
    if constexpr (TAB) pin_vgpr(c.ec.shifter);
    // the build kernels (previous launches) have consumed the displacement flag
    if (blockIdx.x == 0 && threadIdx.x == 0 && lists_stale(maxdisp, skin)) {
        maxdisp[0] = 0ull;
        maxdisp[1] = 0ull;
    }
    const int lane = threadIdx.x & (TILE - 1);
    const int wv = threadIdx.x / TILE;
    if constexpr (TAB) {
        __shared__ double etab_s[EXP_TN];
        load_exp_table(etab_s, c.etab_g);
        c.etab = etab_s;
    }
    const int g = lane / CL, gl = lane % CL;
    // items in launch order share-index first (every set's share 0, then share
    // 1, ...): the back entries -- FP64 in the mixed ion form -- lead every
    // list, so the first shares are the longest, and they start first
    const int item = __builtin_amdgcn_readfirstlane(blockIdx.x * (BLOCK / TILE) + wv);
    const int nsets = (ncl + CPW - 1) / CPW;
    const int m = item / nsets, set = item % nsets;
    if (m >= M) return;   // whole wave; no block barrier follows
    const int ci = set * CPW + g;
    const int cl = (ci < ncl) ? order[ci] : -1;
    const int i = cl * CL + gl;
    const bool ivalid = cl >= 0 && i < n;
    const double nanv = __builtin_nan("");
    const double4 pi = ivalid ? pos[i] : make_double4(nanv, nanv, nanv, 0.0);
    double cc[3] = {0.0, 0.0, 0.0};
    if constexpr (PREC == 1 && KIND != NN) group_centre(pi, ivalid, cc);
    // this group's share of its cluster's list
    int2 sc[NSEG];
    int total = 0;
#pragma unroll
    for (int sg = 0; sg < NSEG; ++sg) {
        sc[sg] = (cl >= 0) ? ccounts[cl * NSEG + sg] : make_int2(0, 0);
        total += sc[sg].x + sc[sg].y;
    }
    int lo = 0, hi = 0;
    list_share(total, M, m, lo, hi);
    int iters = hi - lo;
#pragma unroll
    for (int off = CL; off < TILE; off <<= 1) {
        const int o = __shfl_xor(iters, off, 64);
        iters = o > iters ? o : iters;
    }
    iters = __builtin_amdgcn_readfirstlane(iters);
    const unsigned *mine = clist + (size_t)(cl >= 0 ? cl : 0) * NSEG * segcap;
    const SegList sl = seg_list(mine, segcap, sc, hi);
    Acc<MEAS> acc;
    clist_share<KIND, PREC, MEAS>(c, pi, cc, sl, lo, iters, pos, sj[wv], sfj[wv], hdr_s[wv], lane,
                                  acc);
    if (cl < 0) return;
    constexpr int NC = MEAS ? 11 : 3;
    double *o = part + ((size_t)cl * M + m) * NC * CL + gl;
    o[0] = acc.fx; o[CL] = acc.fy; o[2 * CL] = acc.fz;
    if constexpr (MEAS) {
        o[3 * CL] = acc.e; o[4 * CL] = acc.w;
        o[5 * CL] = acc.pxx; o[6 * CL] = acc.pxy; o[7 * CL] = acc.pxz;
        o[8 * CL] = acc.pyy; o[9 * CL] = acc.pyz; o[10 * CL] = acc.pzz;
    }
}

// Directed pairs a k_pairs_clist launch over the current lists evaluates:
// per item 64 x its longest group share (the lanes a wave runs, masked or
// not), one thread per item, a block sum, one atomic per block.
static __global__ __launch_bounds__(BLOCK) void k_clist_swept(const int2 *__restrict__ ccounts,
                                                             const int *__restrict__ order,
                                                             int ncl, int M,
                                                             unsigned long long *__restrict__ swept)
{
    const int nsets = (ncl + CPW - 1) / CPW;
    const int item = blockIdx.x * BLOCK + threadIdx.x;
    unsigned long long v = 0;
    if (item < nsets * M) {
        const int m = item / nsets, set = item % nsets;
        int iters = 0;
        for (int g = 0; g < CPW; ++g) {
            const int ci = set * CPW + g;
            if (ci >= ncl) continue;
            const int cl = order[ci];
            int total = 0;
            for (int sg = 0; sg < NSEG; ++sg) {
                const int2 q = ccounts[cl * NSEG + sg];
                total += q.x + q.y;
            }
            int lo = 0, hi = 0;
            list_share(total, M, m, lo, hi);
            iters = hi - lo > iters ? hi - lo : iters;
        }
        v = (unsigned long long)iters * TILE;
    }
    __shared__ unsigned long long red[BLOCK];
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = BLOCK / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0 && red[0] > 0) atomicAdd(swept, red[0]);
}

// ---------------------------------------------------------------------------
// Hilbert keys of the current slots (cell grid 2^g per axis) and the gather
// that applies a sort permutation to every per-particle array.  by_charge:
// charged particles take bit 3g of the key, so the sort puts the neutral
// ones (neutrons of the nucleon model) first and the charged ones after,
// each in Hilbert order, and a 16-slot cluster holds one kind except where
// the two runs meet.
static __global__ __launch_bounds__(BLOCK) void k_hilbert(double4 *__restrict__ pos, int n,
                                                   double gx, double gy, double gz, int g,
                                                   double Lx, double Ly, double Lz,
                                                   int by_charge, unsigned *__restrict__ key,
                                                   int *__restrict__ val)
{
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    double4 p = pos[i];
    p.x -= Lx * __builtin_floor(p.x / Lx);
    p.y -= Ly * __builtin_floor(p.y / Ly);
    p.z -= Lz * __builtin_floor(p.z / Lz);
    pos[i] = p;
    const int gmax = (1 << g) - 1;
    int cx = (int)(p.x * gx), cy = (int)(p.y * gy), cz = (int)(p.z * gz);
    cx = cx < 0 ? 0 : (cx > gmax ? gmax : cx);
    cy = cy < 0 ? 0 : (cy > gmax ? gmax : cy);
    cz = cz < 0 ? 0 : (cz > gmax ? gmax : cz);
    const unsigned charged = (by_charge && p.w != 0.0) ? 1u << (3 * g) : 0u;
    key[i] = hilbert_key3((unsigned)cx, (unsigned)cy, (unsigned)cz, g) | charged;
    val[i] = i;
}

struct PermArrays {
    const double4 *pos; const double *vel, *vold, *acc, *aii, *den; const int *orig;
    double4 *pos2; double *vel2, *vold2, *acc2, *aii2, *den2; int *orig2;
};

static __global__ __launch_bounds__(BLOCK) void k_permute(const int *__restrict__ perm, int n,
                                                   PermArrays a)
{
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const size_t nn = (size_t)n;
    const int s = perm[i];
    a.pos2[i] = a.pos[s];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        a.vel2[k * nn + i] = a.vel[k * nn + s];
        a.vold2[k * nn + i] = a.vold[k * nn + s];
        a.acc2[k * nn + i] = a.acc[k * nn + s];
    }
    a.aii2[i] = a.aii[s];
    a.den2[i] = a.den[s];
    a.orig2[i] = a.orig[s];
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
// MODE 2: kick-drift-kick closing half kick (src/vel_verlet_1.f:52-56).
// LAYOUT 0: partials part[c][NC][n] from k_pairs (nparts chunks);
// LAYOUT 1: part[t][m][NC][TILE] from k_pairs_list (nparts items per tile);
// LAYOUT 2: part[cl][m][NC][CL] from k_pairs_clist (nparts items per cluster).
// DRIFT (vel_verlet_2 only): the next step follows at once (same dt, no
// re-sort or utility in between), so this kernel also does that step's drift
// for its own slot with k_drift<1>'s arithmetic -- positions, vold, predicted
// velocity, tile AABBs, displacement tracking -- and the next step launches
// no k_drift (DriftArgs: the arrays only the drift writes).
struct DriftArgs {
    double4 *pos;
    double *vold;
    int ntiles;
    TileBox *box;
    const double4 *xref;
    unsigned long long *maxdisp;
};

// The per-slot sums of k_pairs_clist's M item partials (part[cl][m][NC][CL])
// for lists of many items per cluster set (small systems, where k_finalize's
// one thread per slot would wait on M dependent load rounds from a few
// workgroups): one workgroup per cluster, T threads per slot, thread t
// summing items t, t + T, t + 2T, ... in order, then the T sub-sums added in
// order t = 0 .. T-1.  out[k][n] is read by k_finalize as one partial
// (LAYOUT 0, nparts = 1); the order is fixed, so results are deterministic.
template <int NC, int T>
__global__ __launch_bounds__(CL * T) void k_part_reduce(const double *__restrict__ part, int M,
                                                       int n, double *__restrict__ out)
{
    static_assert(T >= 4 && CL * T <= 1024, "T sub-sums per slot");
    __shared__ double red[T][NC][CL];
    const int cl = blockIdx.x;
    const int gl = threadIdx.x % CL, t = threadIdx.x / CL;
    double f[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) f[k] = 0.0;
    const double *p0 = part + (size_t)cl * M * NC * CL + gl;
    // 24 (force-only) or 22 (measuring) loads in flight
    constexpr int UNROLL = NC == 3 ? 8 : 2;
#pragma unroll UNROLL
    for (int m = t; m < M; m += T) {
        const double *p = p0 + (size_t)m * NC * CL;
#pragma unroll
        for (int k = 0; k < NC; ++k) f[k] += p[k * CL];
    }
#pragma unroll
    for (int k = 0; k < NC; ++k) red[t][k][gl] = f[k];
    __syncthreads();
    const int i = cl * CL + gl;
    for (int k = t; k < NC; k += T) {
        double s = red[0][k][gl];
#pragma unroll
        for (int u = 1; u < T; ++u) s += red[u][k][gl];
        if (i < n) out[(size_t)k * n + i] = s;
    }
}

template <int KIND, bool MEAS, int MODE, int LAYOUT, bool DRIFT = false>
__global__ __launch_bounds__(BLOCK) void k_finalize(
    const double *__restrict__ part, int nparts, const double4 *__restrict__ pos,
    const double *__restrict__ den, const double *__restrict__ aii,
    const int *__restrict__ orig, double *__restrict__ acc, double *__restrict__ vel,
    const double *__restrict__ vold, StepConsts sc, double *__restrict__ blockpart,
    DriftArgs dr = DriftArgs{})
{
    static_assert(!DRIFT || MODE == 1, "the fused drift is vel_verlet_2's");
    const int n = sc.n;
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    const size_t nn = (size_t)n;
    constexpr int NC = MEAS ? 11 : 3;
    double q[NQ];
#pragma unroll
    for (int k = 0; k < NQ; ++k) q[k] = 0.0;
    double xn[3] = {0, 0, 0};   // DRIFT: the next step's positions
    if (i < n) {
        double f[NC];
#pragma unroll
        for (int k = 0; k < NC; ++k) f[k] = 0.0;
        // part p's NC values for this slot start at base(p) and are stride
        // apart; the sums keep part order.  Force-only sums (NC = 3) unroll by
        // 8 parts so that 24 loads are in flight; measuring sums (NC = 11)
        // keep two parts in flight (unrolled by 8 they ran four times slower).
        auto sum_parts = [&](auto base, int stride) {
            if constexpr (NC == 3) {
#pragma unroll 8
                for (int m = 0; m < nparts; ++m) {
                    const double *p = base(m);
#pragma unroll
                    for (int k = 0; k < NC; ++k) f[k] += p[k * stride];
                }
            } else {
#pragma unroll 2
                for (int m = 0; m < nparts; ++m) {
                    const double *p = base(m);
#pragma unroll
                    for (int k = 0; k < NC; ++k) f[k] += p[k * stride];
                }
            }
        };
        if constexpr (LAYOUT == 0) {
            sum_parts([&](int c) { return part + (size_t)c * NC * nn + i; }, n);
        } else if constexpr (LAYOUT == 1) {
            const int t = i / TILE, lane = i % TILE;
            sum_parts([&](int m) { return part + ((size_t)t * nparts + m) * NC * TILE + lane; },
                      TILE);
        } else {
            const int cl = i / CL, gl = i % CL;
            sum_parts([&](int m) { return part + ((size_t)cl * nparts + m) * NC * CL + gl; }, CL);
        }
        // DRIFT reads and writes positions and vold through dr only
        const double4 p0 = DRIFT ? dr.pos[i] : pos[i];
        double zi = p0.w;
        // the reference's mixture kernel never applies zii to the caller's last
        // particle (src/int_ion_mix.f:158, :201, :250); YK_FLAG_REF_F1 keeps that
        double fz_i = (KIND == MIX && !(sc.ref_f1 && orig[i] == n - 1)) ? zi : 1.0;
        double d = den[i];
        double a[3];   // fz_i is exactly 1 unless this is a mixture
#pragma unroll
        for (int k = 0; k < 3; ++k) a[k] = (sc.cnum * (fz_i * f[k])) / d;
        if constexpr (MODE == 1) {
            if (sc.has_bfield) {   // src/vel_verlet_2.f:61-69
                double omega = sc.bconst * zi * sc.bfield / aii[i];
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
                double vp = (DRIFT ? dr.vold[k * nn + i] : vold[k * nn + i]) + sc.dt * a[k];
                v[k] = 0.5 * (vel[k * nn + i] + vp);
                if constexpr (!DRIFT) vel[k * nn + i] = v[k];
            }
        } else if constexpr (MODE == 2) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                v[k] = vel[k * nn + i] + sc.hdt * a[k];
                vel[k * nn + i] = v[k];
            }
        } else if constexpr (MODE == 3) {   // src/vel_verlet_3.f:66-71
            const double omega = sc.bconst * zi * sc.bfield / aii[i];
            const double omegap = sc.hdt * omega;
            v[0] = (vel[i] + sc.hdt * (a[0] + omegap * a[1])) / (1.0 + omegap * omegap);
            v[1] = (vel[nn + i] + sc.hdt * (a[1] - omegap * a[0])) / (1.0 + omegap * omegap);
            v[2] = vel[2 * nn + i] + sc.hdt * a[2];
#pragma unroll
            for (int k = 0; k < 3; ++k) vel[k * nn + i] = v[k];
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
        if constexpr (DRIFT) {   // k_drift<1> (src/vel_verlet_2.f:33-53)
            const double xo[3] = {p0.x, p0.y, p0.z};
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                xn[k] = xo[k] + v[k] * sc.dt + sc.afac * a[k];   // contraction is off
                dr.vold[k * nn + i] = v[k];
                vel[k * nn + i] = v[k] + sc.dt * a[k];
            }
            dr.pos[i] = make_double4(xn[0], xn[1], xn[2], p0.w);
        }
    }
    if constexpr (MEAS) block_reduce_store(q, blockpart + (size_t)blockIdx.x * NQ);
    if constexpr (DRIFT) {
        const int tile = i / TILE;
        if (tile < dr.ntiles) tile_box_store(xn, i < n, tile, dr.box);
        if (dr.xref) {   // largest displacement since the list build
            double d2 = 0.0;
            if (i < n) {
                const double4 r = dr.xref[i];
                const double xr[3] = {r.x, r.y, r.z};
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    double d = xn[k] - xr[k];
                    d2 = __builtin_fma(d, d, d2);
                }
            }
            record_disp(dr.maxdisp, d2);
        }
    }
}

// Drift (+ velocity predictor or opening half kick), then the tile AABBs of
// the new positions and the largest displacement since the pair lists were
// built.  MODE 1: vel_verlet_2 (src/vel_verlet_2.f:33-53); MODE 2: drift +
// half kick (src/vel_verlet_1.f:28-48); MODE 3: drift + first Spreiter-Walter
// velocity stage (src/vel_verlet_3.f:31-54).
template <int MODE>
__global__ __launch_bounds__(BLOCK) void k_drift(double4 *__restrict__ pos,
                                                 double *__restrict__ vel,
                                                 double *__restrict__ vold,
                                                 const double *__restrict__ acc,
                                                 const double *__restrict__ aii,
                                                 StepConsts sc, int ntiles,
                                                 TileBox *__restrict__ box,
                                                 const double4 *__restrict__ xref,
                                                 unsigned long long *__restrict__ maxdisp)
{
    const int n = sc.n;
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    const size_t nn = (size_t)n;
    const bool valid = i < n;
    double x[3] = {0, 0, 0};
    if (valid) {
        double4 p = pos[i];
        x[0] = p.x; x[1] = p.y; x[2] = p.z;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            double v = vel[k * nn + i], a = acc[k * nn + i], xk;
            if constexpr (MODE == 1) {
                xk = x[k] + v * sc.dt + sc.afac * a;   // contraction is off in this TU
                vold[k * nn + i] = v;
                vel[k * nn + i] = v + sc.dt * a;
            } else if constexpr (MODE == 3) {   // src/vel_verlet_3.f:35
                xk = x[k] + v * sc.dt + sc.afac * a;
            } else {   // src/vel_verlet_1.f:30-48
                xk = x[k] + v * sc.dt + a * sc.afac;
                vel[k * nn + i] = v + sc.hdt * a;
            }
            // No periodic wrap here: positions stay continuous between
            // re-sorts so that the periodic image stored with every pair-list
            // entry stays valid.  k_hilbert wraps them into the box at each
            // re-sort and the host boundary wraps them on download with the
            // reference's rule (src/vel_verlet_2.f:39-40).
            x[k] = xk;
        }
        pos[i] = make_double4(x[0], x[1], x[2], p.w);
        if constexpr (MODE == 3) {   // src/vel_verlet_3.f:47-53
            const double a0 = acc[i], a1 = acc[nn + i], a2 = acc[2 * nn + i];
            const double omega = sc.bconst * p.w * sc.bfield / aii[i];
            const double omegap = sc.hdt * omega;
            const double c1 = 1.0 - omegap * omegap;
            // the y update reads the x component updated just before it
            const double v0 = c1 * vel[i] + 2.0 * omegap * vel[nn + i] +
                              sc.hdt * (a0 + omegap * a1);
            const double v1 = c1 * vel[nn + i] - 2.0 * omegap * v0 + sc.hdt * (a1 - omegap * a0);
            vel[i] = v0;
            vel[nn + i] = v1;
            vel[2 * nn + i] = vel[2 * nn + i] + sc.hdt * a2;
        }
    }
    const int tile = i / TILE;
    if (tile < ntiles) tile_box_store(x, valid, tile, box);
    if (xref) {   // largest displacement since the list build
        double d2 = 0.0;
        if (valid) {
            const double4 r = xref[i];
            const double xr[3] = {r.x, r.y, r.z};
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                double d = x[k] - xr[k];
                d2 = __builtin_fma(d, d, d2);
            }
        }
        record_disp(maxdisp, d2);
    }
}

// Kinetic energy sum A*(vx^2+vy^2+vz^2) per block (ttot, src/gen_utils.f:40-55).
static __global__ __launch_bounds__(BLOCK) void k_ttot(const double *__restrict__ vel,
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
// ---------------------------------------------------------------------------
// Velocity utilities of the step loop on the device (src/gen_utils.f:40-107):
// the centre-of-mass velocity and the rescale to the temperature, with
// fixed-order reductions (per-block sums, then one workgroup over the blocks:
// repeatable, no host round trip).  Per block: q[0] = sum aii |v|^2,
// q[1] = sum aii, q[2..4] = sum aii v_k.
constexpr int NVQ = 5;

static __global__ __launch_bounds__(BLOCK) void k_vsums(const double *__restrict__ vel,
                                                       const double *__restrict__ aii, int n,
                                                       double *__restrict__ blockpart)
{
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    const size_t nn = (size_t)n;
    double q[NVQ] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (i < n) {
        const double a = aii[i], vx = vel[i], vy = vel[nn + i], vz = vel[2 * nn + i];
        q[0] = a * (vx * vx + vy * vy + vz * vz);   // src/gen_utils.f:50
        q[1] = a;                                   // src/gen_utils.f:96-97
        q[2] = a * vx;
        q[3] = a * vy;
        q[4] = a * vz;
    }
    __shared__ double red[NVQ][BLOCK];
#pragma unroll
    for (int k = 0; k < NVQ; ++k) red[k][threadIdx.x] = q[k];
    __syncthreads();
    for (int st = BLOCK / 2; st > 0; st >>= 1) {
        if (threadIdx.x < st) {
#pragma unroll
            for (int k = 0; k < NVQ; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + st];
        }
        __syncthreads();
    }
    if (threadIdx.x < NVQ) blockpart[(size_t)blockIdx.x * NVQ + threadIdx.x] = red[threadIdx.x][0];
}

constexpr int VEL_COM = 0;     // vstate[0..2] = vcom = ptot / mtot
constexpr int VEL_TNORM = 1;   // vstate[3] = fac = sqrt(eka / (1.5 kT))

struct VelConsts {
    double xmass, kT;
    float n_f;   // float(n), src/gen_utils.f:53
    int mode;
};

static __global__ __launch_bounds__(BLOCK) void k_vfinal(const double *__restrict__ blockpart,
                                                        int nblocks, VelConsts vc,
                                                        double *__restrict__ vstate)
{
    double q[NVQ] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < nblocks; b += BLOCK) {
#pragma unroll
        for (int k = 0; k < NVQ; ++k) q[k] += blockpart[(size_t)b * NVQ + k];
    }
    __shared__ double red[NVQ][BLOCK];
#pragma unroll
    for (int k = 0; k < NVQ; ++k) red[k][threadIdx.x] = q[k];
    __syncthreads();
    for (int st = BLOCK / 2; st > 0; st >>= 1) {
        if (threadIdx.x < st) {
#pragma unroll
            for (int k = 0; k < NVQ; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + st];
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    if (vc.mode == VEL_COM) {   // src/gen_utils.f:100
        for (int k = 0; k < 3; ++k) vstate[k] = red[2 + k][0] / red[1][0];
    } else {                    // src/gen_utils.f:53, :68
        const double eka = (0.5 * vc.xmass * red[0][0]) / (double)vc.n_f;
        vstate[3] = __builtin_sqrt(eka / (1.5 * vc.kT));
    }
}

static __global__ __launch_bounds__(BLOCK) void k_vapply(double *__restrict__ vel, int n, int mode,
                                                        const double *__restrict__ vstate)
{
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const size_t nn = (size_t)n;
    if (mode == VEL_COM) {   // src/gen_utils.f:101-105
#pragma unroll
        for (int k = 0; k < 3; ++k) vel[k * nn + i] = vel[k * nn + i] - vstate[k];
    } else {                 // src/gen_utils.f:70-73
        const double fac = vstate[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) vel[k * nn + i] = vel[k * nn + i] / fac;
    }
}

// A new box (strain, or a caller's new edge lengths): positions, continuous
// on the device since the last re-sort, first take the periodic wrap into
// the old box that the reference applies every step (src/vel_verlet_2.f:
// 39-40; a floor-based wrap for anything farther out), then strain's wrap of
// those beyond the new box (src/strain.f:76-82).
static __global__ __launch_bounds__(BLOCK) void k_box_change(double4 *__restrict__ pos, int n,
                                                            double lo0, double lo1, double lo2,
                                                            double ln0, double ln1, double ln2)
{
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    double4 p = pos[i];
    double x[3] = {p.x, p.y, p.z};
    const double lo[3] = {lo0, lo1, lo2}, ln[3] = {ln0, ln1, ln2};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        double v = x[k];
        if (v < 0.0) v = v + lo[k];
        if (v > lo[k]) v = v - lo[k];
        if (v < 0.0 || v > lo[k]) v = v - lo[k] * __builtin_floor(v / lo[k]);
        if (v > ln[k]) v = v - ln[k];
        x[k] = v;
    }
    p.x = x[0];
    p.y = x[1];
    p.z = x[2];
    pos[i] = p;
}

struct MeasOut {
    double ev, ek, px, vir, pp[6];
};

struct MeasConsts {
    double cnum, rho, kT, xmass, vol;
    float three_n;   // 3.0*n as the reference's real(4) product
    float n_f;       // float(n)
    int mode;        // MEAS_ACCEL, an integrator mode (1, 2, 3), or MEAS_TTOT
};

constexpr int MEAS_ACCEL = 0;
constexpr int MEAS_TTOT = 16;

static __global__ __launch_bounds__(BLOCK) void k_meas_final(const double *__restrict__ blockpart,
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
    if (mc.mode == MEAS_TTOT) {
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
    if (mc.mode != MEAS_ACCEL) {   // src/vel_verlet_2.f:96-106, src/vel_verlet_3.f:80-90
        m.ek = (0.5 * mc.xmass) * (tot[8] + tot[11] + tot[13]) / (double)mc.n_f;
        for (int k = 0; k < 6; ++k) m.pp[k] = (m.pp[k] + mc.xmass * tot[8 + k]) / mc.vol;
    }
    *out = m;
}

}  // namespace yk
