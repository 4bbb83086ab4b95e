// yk_lib.hip -- C-ABI host side of libyukawa_hip.so (declared in include/yk.h).
//
// One yk_system owns one particle system resident in HBM.  yk_accel replaces
// accel(do_measurements) (src/int_control.f:52-81) and yk_newton replaces
// newton(do_measurements) (src/vel_verlet_2.f:18-120).  The host-side
// transposes between the reference's Fortran (3,n) arrays and the device
// structure-of-arrays happen only in yk_put_state / yk_get_state.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstddef>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "../../include/yk.h"
#include "yk_kernels.hpp"

using namespace yk;

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define YK_HIP(call)                                                                   \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess) {                                                        \
            if (e_ == hipErrorOutOfMemory) (void)hipGetLastError();                    \
            return fail(e_ == hipErrorOutOfMemory ? YK_ERR_NOMEM : YK_ERR_HIP,          \
                        "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, \
                        __LINE__);                                                     \
        }                                                                              \
    } while (0)

}  // namespace

// Work items per cluster set of the cluster-list force kernels (-D overrides
// them for tuning builds).  Measured on MI355X at N = 27,648, FP64 step time
// (median of 6): 24 items 0.419 ms, 32 0.385, 40 0.406, 48 0.414, 64 0.449.
// At 5 resident waves per SIMD (the kernel's 87 VGPRs) 32 items per set give
// 13,824 waves, 2.7 per wave slot, where 48 left a nearly empty fifth round;
// fewer items also mean fewer partial sums for the finalize to read.
#ifndef YK_LIST_ITEMS_FP64
#define YK_LIST_ITEMS_FP64 32
#endif
#ifndef YK_LIST_ITEMS_MIXED
#define YK_LIST_ITEMS_MIXED 32
#endif
constexpr int LIST_ITEMS_FP64 = YK_LIST_ITEMS_FP64;
constexpr int LIST_ITEMS_MIXED = YK_LIST_ITEMS_MIXED;
constexpr int CLIST_SMALL_WAVES = 8192;   // waves below which a cluster-list launch gets more items

struct yk_system {
    yk_params p{};
    int n = 0;
    int device = 0;
    // caller-order host copies of the per-particle constants
    std::vector<double> zhost, ahost, dhost;
    hipStream_t stream = nullptr;
    // device state, in Hilbert-sorted slot order (second set: permutation target)
    double4 *pos = nullptr, *pos2 = nullptr;
    double *vel = nullptr, *vold = nullptr, *acc = nullptr;
    double *vel2 = nullptr, *vold2 = nullptr, *acc2 = nullptr;
    double *aii = nullptr, *den = nullptr, *aii2 = nullptr, *den2 = nullptr;
    int *orig = nullptr, *orig2 = nullptr;   // caller index of each slot
    TileBox *box = nullptr;                  // [ntiles]
    // cluster-pair lists
    bool use_list = true;
    int list_M = 64;                         // work items per target tile / cluster set
    double skin = 0.0;
    double near_r = -1.0;                    // near-field radius of the mixed ion form (fm)
    unsigned *list = nullptr;                // [ntiles][n]
    int2 *counts = nullptr;                  // [ntiles] (fast, general)
    bool use_clist = false;                  // 16-slot cluster lists on top of the tile lists
    int ncl = 0;                             // clusters
    unsigned *clist = nullptr;               // [ncl][n]
    int2 *ccounts = nullptr;                 // [ncl] (fast, general)
    int *corder = nullptr;                   // [ncl] clusters by list length
    int *tchg = nullptr;                     // [ntiles] tile holds a charged particle
    int segcap = 0;                          // entries per cluster-list segment
    bool corder_stale = true;                // slots moved: order to be recomputed
    double4 *xref = nullptr;                 // positions at the last build
    unsigned long long *maxdisp = nullptr;   // [2] two largest displacements^2 since build (bits)
    unsigned long long *swept = nullptr;     // directed pairs evaluated (counter)
    unsigned long long swept_base = 0;
    double *part = nullptr;        // [nchunks][11][n] or [ntiles][M][11][64]
    size_t part_doubles = 0;
    double *psum = nullptr;        // [11][n] cluster-list partials summed (k_part_reduce)
    double *blockpart = nullptr;   // [nblocks][NQ]
    MeasOut *d_meas = nullptr;     // device scalar records
    double *d_vstate = nullptr;    // [4] vcom, fac of the velocity utilities
    double *etab_g = nullptr;      // [EXP_TN] 2^(j/N) table of the FP64 exp
    int meas_cap = 0;
    // sort scratch
    unsigned *keys = nullptr, *keys2 = nullptr;
    int *vals = nullptr, *vals2 = nullptr;
    void *sort_tmp = nullptr;
    size_t sort_tmp_bytes = 0;
    int grid_bits = 1;
    // host staging (pinned)
    double *h_stage = nullptr;     // 4n doubles
    int *h_orig = nullptr;         // n
    MeasOut *h_meas = nullptr;
    int h_meas_cap = 0;
    // decomposition
    int nblocks = 0, ntiles = 0, nstages = 0, nchunks = 0, stages_per_chunk = 0;
    bool have_accel = false;
    int64_t steps_done = 0;        // newton steps since the last put_state
    int resort_every = 0;
    // timing
    std::vector<hipEvent_t> ev;    // pairs of (start, stop) for force launches
    size_t ev_used = 0;
    hipEvent_t step0 = nullptr, step1 = nullptr;
    yk_timing timing{};
};

namespace {

// Device positions are continuous between re-sorts; the caller sees them in
// the box: the reference's single wrap (src/vel_verlet_2.f:39-40) when that
// suffices, a floor-based wrap for anything farther out.
double wrap_into_box(double x, double L)
{
    if (x < 0.0) x = x + L;
    if (x > L) x = x - L;
    if (x < 0.0 || x > L) x = x - L * std::floor(x / L);
    return x;
}

bool finite3(const double *x, int n)
{
    for (size_t k = 0; k < (size_t)3 * n; ++k)
        if (!std::isfinite(x[k])) return false;
    return true;
}

int check_params(const yk_params *p)
{
    if (!p) return fail(YK_ERR_ARG, "params is NULL");
    if (p->sim_type < YK_PURE_ION || p->sim_type > YK_BFIELD_TEST)
        return fail(YK_ERR_ARG, "sim_type=%d: program not compiled to do this simulation type",
                    p->sim_type);
    if (p->precision != YK_FP64 && p->precision != YK_MIXED)
        return fail(YK_ERR_ARG, "precision=%d is not YK_FP64 or YK_MIXED", p->precision);
    for (int k = 0; k < 3; ++k)
        if (!(p->xl[k] > 0.0) || !std::isfinite(p->xl[k]))
            return fail(YK_ERR_ARG, "box edge xl(%d)=%g must be positive", k + 1, p->xl[k]);
    if (!(p->rccut > 0.0)) return fail(YK_ERR_ARG, "rccut=%g must be positive", p->rccut);
    if (p->xmuc < 0.0 || !std::isfinite(p->xmuc))
        return fail(YK_ERR_ARG, "xmuc=%g must be finite and >= 0", p->xmuc);
    if (!(p->xmass > 0.0)) return fail(YK_ERR_ARG, "xmass=%g must be positive", p->xmass);
    if (p->sim_type == YK_PURE_ION && !(p->ai > 0.0))
        return fail(YK_ERR_ARG, "pure-ion needs ai > 0 (got %g)", p->ai);
    if (p->sim_type == YK_NUCLEON && p->nuclear_on && !(p->xpacket > 0.0))
        return fail(YK_ERR_ARG, "xpacket=%g must be positive", p->xpacket);
    if ((p->flags & YK_FLAG_KDK) && (p->flags & YK_FLAG_SW))
        return fail(YK_ERR_ARG, "YK_FLAG_KDK and YK_FLAG_SW select different integrators");
    if (!std::isfinite(p->bfield)) return fail(YK_ERR_ARG, "bfield must be finite");
    return YK_OK;
}

// evcut switch: src/int_ion_mix.f:140-145, src/int_nn.f:176-181.
double evcut_of(const yk_params &p)
{
    double hd = std::sqrt(p.xl[0] * p.xl[0] + p.xl[1] * p.xl[1] + p.xl[2] * p.xl[2]) / 2.0;
    if (!(p.rccut <= hd)) return 0.0;
    double e = std::exp(-(p.xmuc * p.rccut)) / p.rccut;
    return (p.sim_type == YK_NUCLEON) ? p.vc * std::exp(-(p.xmuc * p.rccut)) / p.rccut : e;
}

ForceConsts force_consts_of(const yk_params &p, int n, int ntiles, const double *etab_g)
{
    ForceConsts c{};
    for (int k = 0; k < 3; ++k) {
        c.L[k] = p.xl[k];
        c.invL[k] = 1.0 / p.xl[k];
    }
    // pair cutoffs as the kernels test them (exp_cutoff2: unchanged for every
    // pair that contributes)
    c.rc2 = exp_cutoff2(p.rccut * p.rccut, p.xmuc, true);
    c.mu = p.xmuc;
    c.evcut = evcut_of(p);
    c.xp2i = (p.xpacket > 0) ? 1.0 / (2.0 * p.xpacket) : 0.0;
    c.rn2 = gauss_cutoff2(exp_cutoff2(p.rncut * p.rncut, c.xp2i, false), c.xp2i);
    c.xpi = (p.xpacket > 0) ? 1.0 / p.xpacket : 0.0;
    c.aa = p.aa;
    c.aa2 = 2.0 * p.aa;
    c.bpc = p.bb + p.cc;
    c.bmc = p.bb - p.cc;
    c.vc = p.vc;
    const double LOG2E = 1.4426950408889634074;
    c.muf = (float)p.xmuc;
    c.mul2ef = (float)(p.xmuc * LOG2E);
    c.mul2e_lo = (float)(p.xmuc * LOG2E - (double)c.mul2ef);
    c.xp2il2ef = (float)(c.xp2i * LOG2E);
    c.xpif = (float)c.xpi;
    c.aaf = (float)c.aa;
    c.aa2f = (float)c.aa2;
    c.bpcf = (float)c.bpc;
    c.bmcf = (float)c.bmc;
    c.vcf = (float)p.vc;
    c.fs = p.xmuc * LOG2E;
    c.fs2 = c.fs * c.fs;
    c.rc2f_lo = (float)(c.rc2 * c.fs2 * (1.0 - 4e-6));
    c.rc2f_hi = (float)(c.rc2 * c.fs2 * (1.0 + 4e-6));
    c.r2_floor_f = 1e-20f;
    c.n = n;
    c.ntiles = ntiles;
    c.coul_on = p.coulomb_on ? 1 : 0;
    c.nuc_on = p.nuclear_on ? 1 : 0;
    c.ec = exp_coeffs();
    c.etab_g = etab_g;
    double lmax = std::max(p.xl[0], std::max(p.xl[1], p.xl[2]));
    c.margin = 1e-12 * lmax;
    for (int k = 0; k < 3; ++k) c.halfL_m[k] = 0.5 * p.xl[k] - 4.0 * c.margin;
    double cut = p.rccut;
    if (p.sim_type == YK_NUCLEON) {
        cut = 0.0;
        if (p.coulomb_on) cut = p.rccut;
        if (p.nuclear_on) cut = std::max(cut, p.rncut);
    }
    // tiles farther apart than the cutoff (plus rounding slack) are skipped
    c.cull2 = (cut >= 1e100) ? 1e300 : cut * cut * (1.0 + 1e-12) + c.margin;
    return c;
}

ForceConsts force_consts(const yk_system *s)
{
    return force_consts_of(s->p, s->n, s->ntiles, s->etab_g);
}

// The radius inside which the nucleon kernels evaluate the nuclear term:
// rncut, or less where the Gaussian is negligible (gauss_cutoff2).
double nuclear_cutoff(const yk_params &p)
{
    const double xp2i = (p.xpacket > 0) ? 1.0 / (2.0 * p.xpacket) : 0.0;
    return std::sqrt(gauss_cutoff2(p.rncut * p.rncut, xp2i));
}

double list_cutoff(const yk_params &p)
{
    double cut = p.rccut;
    if (p.sim_type == YK_NUCLEON) {
        cut = 0.0;
        if (p.coulomb_on) cut = p.rccut;
        if (p.nuclear_on) cut = std::max(cut, nuclear_cutoff(p));
    }
    return cut;
}

// The list skin: 5 % of the list cutoff for ions; 10 % for nucleons, whose
// fastest particles move about 0.3 fm per step at T = 1 MeV, so that a 1 fm
// skin rebuilt the lists every second step (10 %: 3 % less time per step
// at N = 51,200, over 400 steps from the random start; 15 % gains nothing
// more), where the box allows it (a skin below 5 % of the box edge).
double list_skin(const yk_params &p)
{
    const double cut = list_cutoff(p);
    if (p.sim_type == YK_NUCLEON && 0.10 * cut < 0.05 * p.xl[0]) return 0.10 * cut;
    return 0.05 * cut;
}

// The list radius of sources that interact only through the nuclear term
// (neutrons of the nucleon model; every source of an ion system).
double list_cutoff_neutral(const yk_params &p)
{
    if (p.sim_type != YK_NUCLEON) return list_cutoff(p);
    return p.nuclear_on ? nuclear_cutoff(p) : 0.0;
}

ListConsts list_consts(const yk_system *s)
{
    const yk_params &p = s->p;
    ListConsts l{};
    double lmax = std::max(p.xl[0], std::max(p.xl[1], p.xl[2]));
    double margin = 1e-12 * lmax;
    for (int k = 0; k < 3; ++k) {
        l.L[k] = p.xl[k];
        l.invL[k] = 1.0 / p.xl[k];
        l.halfL_m[k] = 0.5 * p.xl[k] - 4.0 * margin;
    }
    double rl = list_cutoff(p) + s->skin + margin;
    l.rl2 = rl * rl;
    const double rln0 = list_cutoff_neutral(p) + s->skin + margin;
    l.rl2n = rln0 * rln0;
    l.skin = s->skin;
    l.n = s->n;
    l.ntiles = s->ntiles;
    l.cap = s->n;
    for (int k = 0; k < 3; ++k) {
        l.Lf[k] = (float)l.L[k];
        l.invLf[k] = (float)l.invL[k];
        l.halfLf_m[k] = (float)(l.halfL_m[k] - 1e-5 * p.xl[k]);
    }
    {   // float offsets carry errors of a few ulp of L: widen absolutely too
        const double lm = std::max(p.xl[0], std::max(p.xl[1], p.xl[2]));
        const double rlf = rl * (1.0 + 5e-6) + 1e-6 * lm;
        l.rl2f = (float)(rlf * rlf);
        const double rln = list_cutoff_neutral(p) + s->skin + margin;
        const double rlnf = rln * (1.0 + 5e-6) + 1e-6 * lm;
        l.rl2f_n = (float)(rlnf * rlnf);
    }
    l.skinf = (float)s->skin;
    l.rn2f = s->near_r < 0.0 ? -1.0f : (float)(s->near_r * s->near_r);
    return l;
}

// The near-field radius for the system's parameters and skin (computed once
// per parameter change: near_radius integrates on the host).
void set_near_radius(yk_system *s)
{
    const yk_params &p = s->p;
    s->near_r = near_radius(p.sim_type == YK_PURE_ION || p.sim_type == YK_ION_MIXTURE,
                            p.precision == YK_MIXED, p.rho, p.xmuc, p.rccut, s->skin);
}

unsigned long long bits_of(double v)
{
    unsigned long long b;
    std::memcpy(&b, &v, sizeof b);
    return b;
}

// Forces a list rebuild before the next pair evaluation.
int invalidate_list(yk_system *s)
{
    if (!s->use_list) return YK_OK;
    hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(1), 0, s->stream, s->maxdisp,
                       bits_of(__builtin_inf()));
    s->corder_stale = true;
    YK_HIP(hipGetLastError());
    return YK_OK;
}

StepConsts step_consts_of(const yk_params &p, int n, double dt)
{
    StepConsts c{};
    const bool kdk = (p.flags & YK_FLAG_KDK) != 0;
    c.dt = dt;
    c.afac = kdk ? 0.5 * dt * dt : 0.5 * (dt * dt);   // vel_verlet_1.f:28 / vel_verlet_2.f:33
    c.hdt = 0.5 * dt;
    // field constant: vel_verlet_2's double literal (src/vel_verlet_2.f:65),
    // vel_verlet_3's single-precision literal widened (src/vel_verlet_3.f:48)
    c.bconst = (p.flags & YK_FLAG_SW) ? (double)3.218404e-20f : 3.218404e-20;
    c.bfield = p.bfield;
    c.has_bfield = (p.bfield != 0.0) ? 1 : 0;
    for (int k = 0; k < 3; ++k) c.L[k] = p.xl[k];
    switch (p.sim_type) {
    case YK_PURE_ION: c.cnum = p.frp * p.zi * p.zi * p.vc; break;   // int_ion_pure.f:229
    case YK_ION_MIXTURE: c.cnum = p.frp * p.vc; break;              // int_ion_mix.f:275
    default: c.cnum = 1.0; break;                                    // int_nn.f:320
    }
    c.n = n;
    c.ref_f1 = (p.flags & YK_FLAG_REF_F1) ? 1 : 0;
    c.kind = p.sim_type;
    return c;
}

StepConsts step_consts(const yk_system *s, double dt) { return step_consts_of(s->p, s->n, dt); }

MeasConsts meas_consts_of(const yk_params &p, int n, int mode)
{
    StepConsts sc = step_consts_of(p, n, 0.0);
    MeasConsts m{};
    m.cnum = sc.cnum;
    m.rho = p.rho;
    m.kT = p.kT;
    m.xmass = p.xmass;
    m.vol = p.xl[0] * p.xl[1] * p.xl[2];
    m.three_n = 3.0f * (float)n;
    m.n_f = (float)n;
    m.mode = mode;
    return m;
}

MeasConsts meas_consts(const yk_system *s, int mode) { return meas_consts_of(s->p, s->n, mode); }

void choose_decomposition(yk_system *s)
{
    const int n = s->n;
    s->nblocks = (n + BLOCK - 1) / BLOCK;
    s->ntiles = (n + TILE - 1) / TILE;
    s->nstages = (n + STAGE - 1) / STAGE;
    // enough workgroups to fill 256 CUs several times over
    int want = (2048 + s->nblocks - 1) / s->nblocks;
    if (want < 1) want = 1;
    if (want > s->nstages) want = s->nstages;
    s->stages_per_chunk = (s->nstages + want - 1) / want;
    s->nchunks = (s->nstages + s->stages_per_chunk - 1) / s->stages_per_chunk;
    // Hilbert grid: about 1-2 particles per cell
    int g = 1;
    while (g < 10 && (double)(1 << (3 * g)) < 0.5 * n) ++g;
    s->grid_bits = g;
}

int ensure_meas_cap(yk_system *s, int cap)
{
    if (cap <= s->meas_cap) return YK_OK;
    if (s->d_meas) YK_HIP(hipFree(s->d_meas));
    s->d_meas = nullptr;
    YK_HIP(hipMalloc(&s->d_meas, sizeof(MeasOut) * (size_t)cap));
    s->meas_cap = cap;
    if (cap > s->h_meas_cap) {
        if (s->h_meas) YK_HIP(hipHostFree(s->h_meas));
        s->h_meas = nullptr;
        YK_HIP(hipHostMalloc(&s->h_meas, sizeof(MeasOut) * (size_t)cap, hipHostMallocDefault));
        s->h_meas_cap = cap;
    }
    return YK_OK;
}

int timing_begin(yk_system *s, hipEvent_t *stop_out)
{
    *stop_out = nullptr;
    if (!(s->p.flags & YK_FLAG_TIMING)) return YK_OK;
    if (s->ev_used + 2 > s->ev.size()) {
        for (int k = 0; k < 64; ++k) {
            hipEvent_t e;
            YK_HIP(hipEventCreate(&e));
            s->ev.push_back(e);
        }
    }
    hipEvent_t a = s->ev[s->ev_used], b = s->ev[s->ev_used + 1];
    s->ev_used += 2;
    YK_HIP(hipEventRecord(a, s->stream));
    *stop_out = b;
    return YK_OK;
}

int timing_collect(yk_system *s)
{
    if (s->ev_used == 0) return YK_OK;
    YK_HIP(hipEventSynchronize(s->ev[s->ev_used - 1]));
    for (size_t k = 0; k < s->ev_used; k += 2) {
        float ms = 0.f;
        YK_HIP(hipEventElapsedTime(&ms, s->ev[k], s->ev[k + 1]));
        s->timing.force_ms += ms;
        s->timing.force_launches += 1;
    }
    s->ev_used = 0;
    return YK_OK;
}

// Directed pairs evaluated by the pair kernels since the last reset
// (lanes of a swept tile or list batch count, masked or not): the list
// launches are counted under YK_FLAG_TIMING (k_clist_swept, k_list_swept),
// the all-pairs kernel counts its own culled tiles.
int read_swept(yk_system *s, unsigned long long *out)
{
    YK_HIP(hipStreamSynchronize(s->stream));
    YK_HIP(hipMemcpy(out, s->swept, sizeof *out, hipMemcpyDeviceToHost));
    return YK_OK;
}

template <int KIND, int PREC>
void launch_pairs_k(yk_system *s, const ForceConsts &c, bool meas)
{
    if constexpr (PREC == 1 && KIND != NN) {
        // the mixed ion form's far pairs on the cluster lists work in lengths
        // scaled by mu log2 e: unscreened Coulomb (mu = 0) is computed in FP64
        if (s->use_clist && !(c.mu > 0.0)) return launch_pairs_k<KIND, 0>(s, c, meas);
    }
    if (s->use_clist) {
        const ListConsts lc = list_consts(s);
        const int items = (s->ncl + CPW - 1) / CPW * s->list_M;
        dim3 grid((items + BLOCK / TILE - 1) / (BLOCK / TILE));
        if (meas)
            hipLaunchKernelGGL((k_pairs_clist<KIND, PREC, true>), grid, dim3(BLOCK), 0,
                               s->stream, s->pos, s->clist, s->ccounts, s->corder, c,
                               s->segcap, s->ncl, s->list_M, s->part, s->maxdisp, lc.skin);
        else
            hipLaunchKernelGGL((k_pairs_clist<KIND, PREC, false>), grid, dim3(BLOCK), 0,
                               s->stream, s->pos, s->clist, s->ccounts, s->corder, c,
                               s->segcap, s->ncl, s->list_M, s->part, s->maxdisp, lc.skin);
        return;
    }
    if (s->use_list) {
        const ListConsts lc = list_consts(s);
        const int items = s->ntiles * s->list_M;
        dim3 grid((items + BLOCK / TILE - 1) / (BLOCK / TILE));
        if (meas)
            hipLaunchKernelGGL((k_pairs_list<KIND, PREC, true>), grid, dim3(BLOCK), 0, s->stream,
                               s->pos, s->list, s->counts, c, s->n, s->list_M, s->part,
                               s->maxdisp, lc.skin);
        else
            hipLaunchKernelGGL((k_pairs_list<KIND, PREC, false>), grid, dim3(BLOCK), 0,
                               s->stream, s->pos, s->list, s->counts, c, s->n, s->list_M,
                               s->part, s->maxdisp, lc.skin);
        return;
    }
    dim3 grid(s->nblocks, s->nchunks);
    if (meas)
        hipLaunchKernelGGL((k_pairs<KIND, PREC, true>), grid, dim3(BLOCK), 0, s->stream,
                           s->pos, s->box, c, s->stages_per_chunk, s->part, s->swept);
    else
        hipLaunchKernelGGL((k_pairs<KIND, PREC, false>), grid, dim3(BLOCK), 0, s->stream,
                           s->pos, s->box, c, s->stages_per_chunk, s->part, s->swept);
}

// Rebuild the pair lists if the two largest displacements since the last
// build add up to more than the skin (decided on the device; the kernels exit
// at once otherwise).
int launch_list_build(yk_system *s)
{
    if (!s->use_list) return YK_OK;
    const ListConsts lc = list_consts(s);
    if (s->use_clist) {
        hipLaunchKernelGGL(k_clist_build, dim3(s->ntiles, NSEG), dim3(BLOCK), 0, s->stream,
                           s->pos, s->box, lc, s->segcap, s->maxdisp, s->clist, s->ccounts,
                           s->xref, s->tchg);
        // the cluster order is refreshed with the build that follows a re-sort
        // (or a new state); between re-sorts clusters keep their particles and
        // their list lengths change slowly, so the order stays balanced
        if (s->corder_stale) {
            if (s->ncl <= CORDER_MAX) {
                hipLaunchKernelGGL(k_corder, dim3(1), dim3(CORDER_THREADS), 0, s->stream,
                                   s->ccounts, s->ncl, lc, s->maxdisp, s->corder);
            } else {
                // the order is stale only after invalidate_list, which also
                // forces this build, so the counts are fresh here; the Hilbert
                // sort's buffers are free between re-sorts
                hipLaunchKernelGGL(k_corder_keys, dim3((s->ncl + BLOCK - 1) / BLOCK),
                                   dim3(BLOCK), 0, s->stream, s->ccounts, s->ncl, s->keys,
                                   s->vals);
                YK_HIP(hipGetLastError());
                size_t bytes = s->sort_tmp_bytes;
                YK_HIP(hipcub::DeviceRadixSort::SortPairs(s->sort_tmp, bytes, s->keys,
                                                          s->keys2, s->vals, s->corder,
                                                          s->ncl, 0, CORDER_KEY_BITS,
                                                          s->stream));
            }
            s->corder_stale = false;
        }
    } else {
        hipLaunchKernelGGL(k_list_build, dim3(s->ntiles), dim3(BLOCK), 0, s->stream, s->pos,
                           s->box, lc, s->maxdisp, s->list, s->counts, s->xref);
    }
    YK_HIP(hipGetLastError());
    return YK_OK;
}

int launch_pairs(yk_system *s, bool meas)
{
    if (s->n < 2 || s->p.sim_type == YK_BFIELD_TEST) {
        YK_HIP(hipMemsetAsync(s->part, 0, sizeof(double) * s->part_doubles, s->stream));
        return YK_OK;
    }
    ForceConsts c = force_consts(s);
    int rc = launch_list_build(s);
    if (rc) return rc;
    hipEvent_t stop = nullptr;
    if ((rc = timing_begin(s, &stop))) return rc;
    const bool mixed = (s->p.precision == YK_MIXED);
    switch (s->p.sim_type) {
    case YK_PURE_ION:
        mixed ? launch_pairs_k<PURE, 1>(s, c, meas) : launch_pairs_k<PURE, 0>(s, c, meas);
        break;
    case YK_ION_MIXTURE:
        mixed ? launch_pairs_k<MIX, 1>(s, c, meas) : launch_pairs_k<MIX, 0>(s, c, meas);
        break;
    case YK_NUCLEON:
        mixed ? launch_pairs_k<NN, 1>(s, c, meas) : launch_pairs_k<NN, 0>(s, c, meas);
        break;
    default: return fail(YK_ERR_UNSUPPORTED, "sim_type %d", s->p.sim_type);
    }
    YK_HIP(hipGetLastError());
    if (stop) {
        YK_HIP(hipEventRecord(stop, s->stream));
        // the pairs the list launch evaluated, counted from its lists outside
        // the timed launch (the all-pairs kernel counts its own culled tiles)
        if (s->use_clist) {
            const int items = (s->ncl + CPW - 1) / CPW * s->list_M;
            hipLaunchKernelGGL(k_clist_swept, dim3((items + BLOCK - 1) / BLOCK), dim3(BLOCK), 0,
                               s->stream, s->ccounts, s->corder, s->ncl, s->list_M, s->swept);
        } else if (s->use_list) {
            const int items = s->ntiles * s->list_M;
            hipLaunchKernelGGL(k_list_swept, dim3((items + BLOCK - 1) / BLOCK), dim3(BLOCK), 0,
                               s->stream, s->counts, s->ntiles, s->list_M, s->swept);
        }
        YK_HIP(hipGetLastError());
    }
    if (s->ev_used >= 1024) return timing_collect(s);
    return YK_OK;
}

// Cluster-list launches of at least this many items per cluster set (small
// systems) sum their partials with k_part_reduce before the finalize.
constexpr int PART_REDUCE_MIN_M = 64;

template <int NC>
void launch_part_reduce(yk_system *s)
{
    if (s->list_M >= 128)
        hipLaunchKernelGGL((k_part_reduce<NC, 16>), dim3(s->ncl), dim3(CL * 16), 0, s->stream,
                           s->part, s->list_M, s->n, s->psum);
    else
        hipLaunchKernelGGL((k_part_reduce<NC, 8>), dim3(s->ncl), dim3(CL * 8), 0, s->stream,
                           s->part, s->list_M, s->n, s->psum);
}

template <int KIND, bool MEAS, int MODE>
void launch_fin_k(yk_system *s, const StepConsts &sc, bool drift)
{
    if (s->use_clist && s->list_M >= PART_REDUCE_MIN_M) {
        // the partials summed by k_part_reduce, read as one LAYOUT 0 part
        launch_part_reduce<MEAS ? 11 : 3>(s);
        if constexpr (MODE == 1) {
            if (drift) {   // with the next step's drift
                const DriftArgs dr{s->pos, s->vold, s->ntiles, s->box,
                                   s->use_list ? s->xref : nullptr, s->maxdisp};
                hipLaunchKernelGGL((k_finalize<KIND, MEAS, 1, 0, true>), dim3(s->nblocks),
                                   dim3(BLOCK), 0, s->stream, s->psum, 1, s->pos, s->den, s->aii,
                                   s->orig, s->acc, s->vel, s->vold, sc, s->blockpart, dr);
                return;
            }
        }
        hipLaunchKernelGGL((k_finalize<KIND, MEAS, MODE, 0>), dim3(s->nblocks), dim3(BLOCK), 0,
                           s->stream, s->psum, 1, s->pos, s->den, s->aii, s->orig, s->acc,
                           s->vel, s->vold, sc, s->blockpart);
        return;
    }
    if constexpr (MODE == 1) {
        if (drift && s->use_clist) {   // with the next step's drift
            const DriftArgs dr{s->pos, s->vold, s->ntiles, s->box,
                               s->use_list ? s->xref : nullptr, s->maxdisp};
            hipLaunchKernelGGL((k_finalize<KIND, MEAS, 1, 2, true>), dim3(s->nblocks),
                               dim3(BLOCK), 0, s->stream, s->part, s->list_M, s->pos, s->den,
                               s->aii, s->orig, s->acc, s->vel, s->vold, sc, s->blockpart, dr);
            return;
        }
    }
    if (s->use_clist)
        hipLaunchKernelGGL((k_finalize<KIND, MEAS, MODE, 2>), dim3(s->nblocks), dim3(BLOCK), 0,
                           s->stream, s->part, s->list_M, s->pos, s->den, s->aii, s->orig,
                           s->acc, s->vel, s->vold, sc, s->blockpart);
    else if (s->use_list)
        hipLaunchKernelGGL((k_finalize<KIND, MEAS, MODE, 1>), dim3(s->nblocks), dim3(BLOCK), 0,
                           s->stream, s->part, s->list_M, s->pos, s->den, s->aii, s->orig,
                           s->acc, s->vel, s->vold, sc, s->blockpart);
    else
        hipLaunchKernelGGL((k_finalize<KIND, MEAS, MODE, 0>), dim3(s->nblocks), dim3(BLOCK), 0,
                           s->stream, s->part, s->nchunks, s->pos, s->den, s->aii, s->orig,
                           s->acc, s->vel, s->vold, sc, s->blockpart);
}

template <int KIND>
void launch_fin_kind(yk_system *s, const StepConsts &sc, bool meas, int mode, bool drift)
{
    if (meas) {
        if (mode == 0) launch_fin_k<KIND, true, 0>(s, sc, false);
        else if (mode == 1) launch_fin_k<KIND, true, 1>(s, sc, drift);
        else if (mode == 2) launch_fin_k<KIND, true, 2>(s, sc, false);
        else launch_fin_k<KIND, true, 3>(s, sc, false);
    } else {
        if (mode == 0) launch_fin_k<KIND, false, 0>(s, sc, false);
        else if (mode == 1) launch_fin_k<KIND, false, 1>(s, sc, drift);
        else if (mode == 2) launch_fin_k<KIND, false, 2>(s, sc, false);
        else launch_fin_k<KIND, false, 3>(s, sc, false);
    }
}

// finalize + optional measurement reduction into d_meas[slot]; drift: the
// finalize also does the next step's vel_verlet_2 drift (cluster lists only)
int launch_finalize(yk_system *s, double dt, bool meas, int mode, int slot, bool drift = false)
{
    StepConsts sc = step_consts(s, dt);
    switch (s->p.sim_type) {
    case YK_PURE_ION: launch_fin_kind<PURE>(s, sc, meas, mode, drift); break;
    case YK_ION_MIXTURE: launch_fin_kind<MIX>(s, sc, meas, mode, drift); break;
    default: launch_fin_kind<NN>(s, sc, meas, mode, drift); break;   // nucleon and B-field test
    }
    YK_HIP(hipGetLastError());
    if (meas) {
        MeasConsts mc = meas_consts(s, mode);
        hipLaunchKernelGGL(k_meas_final, dim3(1), dim3(BLOCK), 0, s->stream, s->blockpart,
                           s->nblocks, mc, s->d_meas + slot);
        YK_HIP(hipGetLastError());
    }
    return YK_OK;
}

int launch_drift(yk_system *s, double dt)
{
    StepConsts sc = step_consts(s, dt);
    const double4 *xr = s->use_list ? s->xref : nullptr;
    if (s->p.flags & YK_FLAG_SW)
        hipLaunchKernelGGL(k_drift<3>, dim3(s->nblocks), dim3(BLOCK), 0, s->stream, s->pos,
                           s->vel, s->vold, s->acc, s->aii, sc, s->ntiles, s->box, xr,
                           s->maxdisp);
    else if (s->p.flags & YK_FLAG_KDK)
        hipLaunchKernelGGL(k_drift<2>, dim3(s->nblocks), dim3(BLOCK), 0, s->stream, s->pos,
                           s->vel, s->vold, s->acc, s->aii, sc, s->ntiles, s->box, xr,
                           s->maxdisp);
    else
        hipLaunchKernelGGL(k_drift<1>, dim3(s->nblocks), dim3(BLOCK), 0, s->stream, s->pos,
                           s->vel, s->vold, s->acc, s->aii, sc, s->ntiles, s->box, xr,
                           s->maxdisp);
    YK_HIP(hipGetLastError());
    return YK_OK;
}

void to_meas(const MeasOut &m, yk_meas *out)
{
    out->ev = m.ev;
    out->ek = m.ek;
    out->px = m.px;
    out->vir = m.vir;
    for (int k = 0; k < 6; ++k) out->pp[k] = m.pp[k];
}

void zero_meas(yk_meas *out)
{
    if (out) std::memset(out, 0, sizeof *out);
}

// Re-establish Hilbert order of the slots: keys, stable radix sort, gather
// of every per-particle array into the second buffer set, swap, tile AABBs.
int resort(yk_system *s)
{
    const int n = s->n;
    const int g = s->grid_bits;
    const double G = (double)(1 << g);
    // nucleons: neutrons and protons in separate Hilbert runs, so that a
    // neutron cluster's list needs sources only within the nuclear radius
    // (list_cutoff_neutral) and only proton clusters reach to the Coulomb one
    const int by_charge = s->p.sim_type == YK_NUCLEON ? 1 : 0;
    hipLaunchKernelGGL(k_hilbert, dim3(s->nblocks), dim3(BLOCK), 0, s->stream, s->pos, n,
                       G / s->p.xl[0], G / s->p.xl[1], G / s->p.xl[2], g,
                       s->p.xl[0], s->p.xl[1], s->p.xl[2], by_charge, s->keys, s->vals);
    YK_HIP(hipGetLastError());
    size_t bytes = s->sort_tmp_bytes;
    YK_HIP(hipcub::DeviceRadixSort::SortPairs(s->sort_tmp, bytes, s->keys, s->keys2, s->vals,
                                              s->vals2, n, 0, 3 * g + by_charge, s->stream));
    PermArrays a{s->pos, s->vel, s->vold, s->acc, s->aii, s->den, s->orig,
                 s->pos2, s->vel2, s->vold2, s->acc2, s->aii2, s->den2, s->orig2};
    hipLaunchKernelGGL(k_permute, dim3(s->nblocks), dim3(BLOCK), 0, s->stream, s->vals2, n, a);
    YK_HIP(hipGetLastError());
    std::swap(s->pos, s->pos2);
    std::swap(s->vel, s->vel2);
    std::swap(s->vold, s->vold2);
    std::swap(s->acc, s->acc2);
    std::swap(s->aii, s->aii2);
    std::swap(s->den, s->den2);
    std::swap(s->orig, s->orig2);
    hipLaunchKernelGGL(k_tile_boxes, dim3(s->nblocks), dim3(BLOCK), 0, s->stream, s->pos, n,
                       s->ntiles, s->box);
    if (s->tchg)
        hipLaunchKernelGGL(k_tile_charge, dim3(s->nblocks), dim3(BLOCK), 0, s->stream, s->pos,
                           n, s->ntiles, s->tchg);
    YK_HIP(hipGetLastError());
    return invalidate_list(s);   // slots moved: the lists refer to old slots
}

// Velocity utilities (src/gen_utils.f:61-107) on the stream, no host sync.
int vel_util(yk_system *s, int mode)
{
    hipLaunchKernelGGL(k_vsums, dim3(s->nblocks), dim3(BLOCK), 0, s->stream, s->vel, s->aii,
                       s->n, s->blockpart);
    const VelConsts vc{s->p.xmass, s->p.kT, (float)s->n, mode};
    hipLaunchKernelGGL(k_vfinal, dim3(1), dim3(BLOCK), 0, s->stream, s->blockpart, s->nblocks,
                       vc, s->d_vstate);
    hipLaunchKernelGGL(k_vapply, dim3(s->nblocks), dim3(BLOCK), 0, s->stream, s->vel, s->n, mode,
                       s->d_vstate);
    YK_HIP(hipGetLastError());
    return YK_OK;
}

int do_tnorm(yk_system *s) { return vel_util(s, VEL_TNORM); }

int do_cancel_vcom(yk_system *s)   // src/gen_utils.f:82-107: then tnorm
{
    int rc = vel_util(s, VEL_COM);
    return rc ? rc : do_tnorm(s);
}

// New box edges: device positions wrap into the old box and beyond the new
// one (k_box_change), the pair lists are rebuilt at the next evaluation.
int change_box(yk_system *s, const double *xl)
{
    if (xl[0] == s->p.xl[0] && xl[1] == s->p.xl[1] && xl[2] == s->p.xl[2]) return YK_OK;
    hipLaunchKernelGGL(k_box_change, dim3(s->nblocks), dim3(BLOCK), 0, s->stream, s->pos, s->n,
                       s->p.xl[0], s->p.xl[1], s->p.xl[2], xl[0], xl[1], xl[2]);
    YK_HIP(hipGetLastError());
    for (int k = 0; k < 3; ++k) s->p.xl[k] = xl[k];
    return invalidate_list(s);
}

// The box after one step of strain rates deps from the gage dimensions xl0,
// with the formulas of src/strain.f:31-72 (host arithmetic, so that the
// box is the reference's to the last bit).  Returns false for no strain.
bool strain_box(double *xl, const double *xl0, const double *deps, double dt)
{
    const int c = (deps[0] != YK_XNOSTRAIN ? 1 : 0) | (deps[1] != YK_XNOSTRAIN ? 2 : 0) |
                  (deps[2] != YK_XNOSTRAIN ? 4 : 0);
    // the edge that follows when one or two are strained, as a function of the others
    switch (c) {
    case 0:
        return false;
    case 1:
        xl[0] = xl[0] + deps[0] * dt * xl0[0];
        xl[1] = xl0[1] * std::sqrt(xl0[0] / xl[0]);
        xl[2] = xl0[2] * std::sqrt(xl0[0] / xl[0]);
        break;
    case 2:
        xl[1] = xl[1] + deps[1] * dt * xl0[1];
        xl[0] = xl0[0] * std::sqrt(xl0[1] / xl[1]);
        xl[2] = xl0[2] * std::sqrt(xl0[1] / xl[1]);
        break;
    case 3:
        xl[0] = xl[0] + deps[0] * dt * xl0[0];
        xl[1] = xl[1] + deps[1] * dt * xl0[1];
        xl[2] = xl0[2] * (xl0[0] / xl[0]) * (xl0[1] / xl[1]);
        break;
    case 4:
        xl[2] = xl[2] + deps[2] * dt * xl0[2];
        xl[0] = xl0[0] * std::sqrt(xl0[2] / xl[2]);
        xl[1] = xl0[1] * std::sqrt(xl0[2] / xl[2]);
        break;
    case 5:
        xl[0] = xl[0] + deps[0] * dt * xl0[0];
        xl[2] = xl[2] + deps[2] * dt * xl0[2];
        xl[1] = xl0[1] * (xl0[0] / xl[0]) * (xl0[2] / xl[2]);
        break;
    case 6:
        xl[1] = xl[1] + deps[1] * dt * xl0[1];
        xl[2] = xl[2] + deps[2] * dt * xl0[2];
        xl[0] = xl0[0] * (xl0[1] / xl[1]) * (xl0[2] / xl[2]);
        break;
    default:
        for (int k = 0; k < 3; ++k) xl[k] = xl[k] + deps[k] * dt * xl0[k];
        break;
    }
    return true;
}

int do_strain(yk_system *s, double dt, const double *deps, const double *xl0)
{
    double xl[3] = {s->p.xl[0], s->p.xl[1], s->p.xl[2]};
    if (!strain_box(xl, xl0, deps, dt)) return YK_OK;
    for (int k = 0; k < 3; ++k)
        if (!(std::isfinite(xl[k]) && xl[k] > 0.0))
            return fail(YK_ERR_ARG, "strain gives box edge %d = %g", k, xl[k]);
    return change_box(s, xl);
}

int mdmod(int p, int q) { return q == 0 ? -1 : p % q; }   // src/md.F:248-256

// One step on the device: (periodic re-sort), drift, pairs, finalize
// (+ measurement into slot).  Inside one run loop the finalize may also do
// the next step's drift: the caller passes next_plain when the next step of
// the same loop follows with the same dt and no utility (cancel_vcom, tnorm,
// strain) before it; the step fuses when, in addition, the integrator is
// vel_verlet_2, the cluster lists are in use and no re-sort is due.  *fused
// tells the caller, which passes it back as drift_done for that next step.
int device_step(yk_system *s, double dt, bool meas, int slot, bool drift_done = false,
                bool next_plain = false, bool *fused = nullptr)
{
    int rc;
    if (fused) *fused = false;
    if (!drift_done) {   // a fused drift is only done when no re-sort is due
        if (s->resort_every > 0 && s->steps_done > 0 && s->steps_done % s->resort_every == 0)
            if ((rc = resort(s))) return rc;
    }
    s->steps_done += 1;
    if (!drift_done && (rc = launch_drift(s, dt))) return rc;
    if ((rc = launch_pairs(s, meas))) return rc;
    const int mode = (s->p.flags & YK_FLAG_SW) ? 3 : (s->p.flags & YK_FLAG_KDK) ? 2 : 1;
    const bool fuse = next_plain && mode == 1 && s->use_clist &&
                      !(s->resort_every > 0 && s->steps_done % s->resort_every == 0);
    if ((rc = launch_finalize(s, dt, meas, mode, slot, fuse))) return rc;
    if (fused) *fused = fuse;
    return YK_OK;
}

void release(yk_system *s)
{
    if (!s) return;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    for (auto e : s->ev) (void)hipEventDestroy(e);
    if (s->step0) (void)hipEventDestroy(s->step0);
    if (s->step1) (void)hipEventDestroy(s->step1);
    for (void *ptr : {(void *)s->pos, (void *)s->pos2, (void *)s->vel, (void *)s->vel2,
                      (void *)s->vold, (void *)s->vold2, (void *)s->acc, (void *)s->acc2,
                      (void *)s->aii, (void *)s->aii2, (void *)s->den, (void *)s->den2,
                      (void *)s->orig, (void *)s->orig2, (void *)s->box, (void *)s->part,
                      (void *)s->psum,
                      (void *)s->blockpart, (void *)s->d_meas, (void *)s->d_vstate, (void *)s->etab_g,
                      (void *)s->keys,
                      (void *)s->keys2, (void *)s->vals, (void *)s->vals2, s->sort_tmp,
                      (void *)s->list, (void *)s->counts, (void *)s->xref,
                      (void *)s->clist, (void *)s->ccounts, (void *)s->corder, (void *)s->tchg,
                      (void *)s->maxdisp, (void *)s->swept})
        if (ptr) (void)hipFree(ptr);
    if (s->h_stage) (void)hipHostFree(s->h_stage);
    if (s->h_orig) (void)hipHostFree(s->h_orig);
    if (s->h_meas) (void)hipHostFree(s->h_meas);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

}  // namespace

// ===========================================================================
namespace {

// ---------------------------------------------------------------------------
// yk_accel_noalloc: accel for device-resident arrays with no allocation.  Its
// work memory is one file-scope device array per GPU (`scratch` below), laid
// out for at most NA_MAX_N particles and NA_MAX_CHUNKS source chunks of the
// tiled all-pairs kernel.  Host code reaches it only through the address
// hipGetSymbolAddress returns on each device (na_scratch); every call on a
// GPU uses the same memory, so calls that overlap on one GPU race.
constexpr int NA_MAX_N = YK_NOALLOC_MAX_N;
constexpr int NA_MAX_CHUNKS = 16;

struct NaLayout {   // byte offsets inside scratch, 256-aligned
    size_t pos, den, orig, box, part, blockpart, acc, etab, swept, total;
};

constexpr size_t na_align(size_t b) { return (b + 255) / 256 * 256; }

constexpr NaLayout na_layout()
{
    constexpr size_t n = NA_MAX_N;
    NaLayout l{};
    size_t o = 0;
    l.pos = o;       o += na_align(sizeof(double4) * n);
    l.den = o;       o += na_align(sizeof(double) * n);
    l.orig = o;      o += na_align(sizeof(int) * n);
    l.box = o;       o += na_align(sizeof(TileBox) * ((n + TILE - 1) / TILE));
    l.part = o;      o += na_align(sizeof(double) * 11 * n * NA_MAX_CHUNKS);
    l.blockpart = o; o += na_align(sizeof(double) * NQ * ((n + BLOCK - 1) / BLOCK));
    l.acc = o;       o += na_align(sizeof(double) * 3 * n);
    l.etab = o;      o += na_align(sizeof(double) * EXP_TN);
    l.swept = o;     o += na_align(sizeof(unsigned long long));
    l.total = o;
    return l;
}
constexpr NaLayout NA = na_layout();

static_assert(sizeof(yk_meas) == sizeof(MeasOut) && offsetof(yk_meas, pp) == offsetof(MeasOut, pp),
              "k_meas_final writes a yk_meas");

}  // namespace

static __device__ __align__(256) unsigned char scratch[NA.total];

namespace {

std::mutex na_mutex;
void *na_ptr[256];   // scratch's device address, per device ordinal

// The address of this GPU's scratch (looked up once per device, with the
// device current).
int na_scratch(unsigned char **out)
{
    int dev = 0;
    YK_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 256) return fail(YK_ERR_UNSUPPORTED, "device ordinal %d", dev);
    std::lock_guard<std::mutex> lock(na_mutex);
    if (!na_ptr[dev]) {
        void *ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(YK_ERR_HIP, "hipGetSymbolAddress(scratch) failed: %s",
                        hipGetErrorString(e));
        na_ptr[dev] = ptr;
    }
    *out = static_cast<unsigned char *>(na_ptr[dev]);
    return YK_OK;
}

// ptr must lie in one device allocation holding at least `bytes` from ptr.
int na_check_range(const void *ptr, size_t bytes, const char *name)
{
    hipDeviceptr_t base = nullptr;
    size_t size = 0;
    hipError_t e = hipMemGetAddressRange(&base, &size, (hipDeviceptr_t)ptr);
    if (e != hipSuccess) {
        (void)hipGetLastError();   // the lookup's error is not the caller's
        return fail(YK_ERR_ARG, "%s is not device memory", name);
    }
    const char *b = static_cast<const char *>(base), *p = static_cast<const char *>(ptr);
    if (p < b || (size_t)(p - b) + bytes > size)
        return fail(YK_ERR_ARG, "%s: its device allocation holds fewer than %zu bytes from it",
                    name, bytes);
    return YK_OK;
}

// Caller's (3,n) positions and per-particle Z, A -> pos {x,y,z,Z}, the
// acceleration denominator (as yk_create forms it) and the identity slot map.
__global__ __launch_bounds__(BLOCK) void k_na_stage(const double *__restrict__ x,
                                                   const double *__restrict__ zii,
                                                   const double *__restrict__ aii, int n,
                                                   int kind, double zi, double ai, double xmass,
                                                   double4 *__restrict__ pos,
                                                   double *__restrict__ den, int *__restrict__ orig)
{
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const double z = zii ? zii[i] : zi;
    pos[i] = make_double4(x[3 * (size_t)i], x[3 * (size_t)i + 1], x[3 * (size_t)i + 2], z);
    den[i] = (kind == YK_PURE_ION) ? ai * xmass
                                   : (kind == YK_ION_MIXTURE ? aii[i] * xmass : xmass);
    orig[i] = i;
}

// The 2^(j/N) table of exp_neg64t, from the device exp2.
__global__ void k_na_exp_table(double *__restrict__ t)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < EXP_TN) t[j] = exp2((double)j / EXP_TN);
}

__global__ __launch_bounds__(BLOCK) void k_na_fill(double *__restrict__ p, size_t count)
{
    const size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i < count) p[i] = 0.0;
}

// Component-major accelerations -> the caller's (3,n) array.
__global__ __launch_bounds__(BLOCK) void k_na_out(const double *__restrict__ acc, int n,
                                                 double *__restrict__ a)
{
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const size_t nn = (size_t)n;
#pragma unroll
    for (int k = 0; k < 3; ++k) a[3 * (size_t)i + k] = acc[k * nn + i];
}

__global__ void k_na_zero_meas(yk_meas *m)
{
    if (threadIdx.x == 0) *m = yk_meas{};
}

template <int KIND, int PREC>
void na_launch_pairs(hipStream_t st, const double4 *pos, const TileBox *box, const ForceConsts &c,
                     int nblocks, int nchunks, int spc, double *part, unsigned long long *swept,
                     bool meas)
{
    const dim3 grid(nblocks, nchunks);
    if (meas)
        hipLaunchKernelGGL((k_pairs<KIND, PREC, true>), grid, dim3(BLOCK), 0, st, pos, box, c,
                           spc, part, swept);
    else
        hipLaunchKernelGGL((k_pairs<KIND, PREC, false>), grid, dim3(BLOCK), 0, st, pos, box, c,
                           spc, part, swept);
}

template <int KIND>
void na_launch_finalize(hipStream_t st, int nblocks, const double *part, int nchunks,
                        const double4 *pos, const double *den, const int *orig, double *acc,
                        const StepConsts &sc, double *blockpart, bool meas)
{
    if (meas)
        hipLaunchKernelGGL((k_finalize<KIND, true, 0, 0>), dim3(nblocks), dim3(BLOCK), 0, st,
                           part, nchunks, pos, den, (const double *)nullptr, orig, acc,
                           (double *)nullptr, (const double *)nullptr, sc, blockpart);
    else
        hipLaunchKernelGGL((k_finalize<KIND, false, 0, 0>), dim3(nblocks), dim3(BLOCK), 0, st,
                           part, nchunks, pos, den, (const double *)nullptr, orig, acc,
                           (double *)nullptr, (const double *)nullptr, sc, blockpart);
}

}  // namespace

extern "C" {

int32_t yk_abi_version(void) { return YK_ABI_VERSION; }

const char *yk_last_error(void) { return g_err.c_str(); }

int yk_create(const yk_params *params, int32_t n, const double *zii, const double *aii,
              yk_system **out)
{
    if (!out) return fail(YK_ERR_ARG, "out is NULL");
    *out = nullptr;
    int rc = check_params(params);
    if (rc) return rc;
    if (n < 1) return fail(YK_ERR_ARG, "n=%d must be >= 1", n);
    if (n > (1 << 26)) return fail(YK_ERR_ARG, "n=%d exceeds the supported 2^26", n);
    if (params->sim_type != YK_PURE_ION && (!zii || !aii))
        return fail(YK_ERR_ARG, "zii and aii are required for sim_type %d", params->sim_type);
    yk_system *s = new (std::nothrow) yk_system();
    if (!s) return fail(YK_ERR_NOMEM, "host allocation failed");
    s->p = *params;
    s->n = n;
    s->resort_every = params->resort_every > 0 ? params->resort_every : 16;
    choose_decomposition(s);
    auto bail = [&](int code) {
        release(s);
        return code;
    };
    // a failed allocation leaves the runtime's last-error state set: cleared,
    // so that a later launch check (of this or another system) does not
    // report it
#define YK_HIPB(call)                                                                  \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess) {                                                        \
            if (e_ == hipErrorOutOfMemory) (void)hipGetLastError();                    \
            return bail(fail(e_ == hipErrorOutOfMemory ? YK_ERR_NOMEM : YK_ERR_HIP,    \
                             "%s failed: %s", #call, hipGetErrorString(e_)));          \
        }                                                                              \
    } while (0)
    const size_t nn = (size_t)n;
    // per-particle Z, A and the acceleration denominator, caller order
    s->zhost.resize(nn);
    s->ahost.resize(nn);
    s->dhost.resize(nn);
    for (size_t i = 0; i < nn; ++i) {
        double zi_ = (params->sim_type == YK_PURE_ION) ? (zii ? zii[i] : params->zi) : zii[i];
        double ai_ = (params->sim_type == YK_PURE_ION) ? (aii ? aii[i] : params->ai) : aii[i];
        if (!std::isfinite(zi_) || !std::isfinite(ai_))
            return bail(fail(YK_ERR_NONFINITE, "zii/aii(%zu) not finite", i));
        s->zhost[i] = zi_;
        s->ahost[i] = ai_;
        switch (params->sim_type) {
        case YK_PURE_ION: s->dhost[i] = params->ai * params->xmass; break;
        case YK_ION_MIXTURE:
            if (!(ai_ > 0.0)) return bail(fail(YK_ERR_ARG, "aii(%zu)=%g must be > 0", i, ai_));
            s->dhost[i] = ai_ * params->xmass;
            break;
        default: s->dhost[i] = params->xmass; break;
        }
    }
    YK_HIPB(hipGetDevice(&s->device));
    YK_HIPB(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    YK_HIPB(hipMalloc(&s->pos, sizeof(double4) * nn));
    YK_HIPB(hipMalloc(&s->pos2, sizeof(double4) * nn));
    for (double **pp : {&s->vel, &s->vel2, &s->vold, &s->vold2, &s->acc, &s->acc2})
        YK_HIPB(hipMalloc(pp, sizeof(double) * 3 * nn));
    for (double **pp : {&s->aii, &s->aii2, &s->den, &s->den2})
        YK_HIPB(hipMalloc(pp, sizeof(double) * nn));
    YK_HIPB(hipMalloc(&s->orig, sizeof(int) * nn));
    YK_HIPB(hipMalloc(&s->orig2, sizeof(int) * nn));
    YK_HIPB(hipMalloc(&s->box, sizeof(TileBox) * (size_t)s->ntiles));
    s->use_list = !(params->flags & YK_FLAG_ALLPAIRS);
    // items per target tile (cluster set)
    s->list_M = (params->precision == YK_MIXED) ? LIST_ITEMS_MIXED : LIST_ITEMS_FP64;
    if (s->use_list && !(params->flags & YK_FLAG_TILE_LISTS) && n <= CLIST_MAX_N) {
        // small systems: more items per cluster set, up to 128, until the
        // launch has 8,192 waves (N = 3,456: 128 items, 6,912 waves, 30 us
        // per force launch against 33 us with 32 items)
        const int nsets = ((n + CL - 1) / CL + CPW - 1) / CPW;
        while (s->list_M < 128 && nsets * s->list_M < CLIST_SMALL_WAVES) s->list_M *= 2;
    }
    s->skin = list_skin(*params);
    if (!(s->skin < 0.05 * params->xl[0])) s->skin = 0.0;   // huge cutoff: every pair listed
    set_near_radius(s);
    s->part_doubles = 11 * nn * (size_t)s->nchunks;
    if (s->use_list) {
        s->part_doubles = std::max(s->part_doubles,
                                   (size_t)s->ntiles * s->list_M * 11 * TILE);
        YK_HIPB(hipMalloc(&s->xref, sizeof(double4) * nn));
        s->use_clist = !(params->flags & YK_FLAG_TILE_LISTS) && n <= CLIST_MAX_N;
        if (!s->use_clist) {
            YK_HIPB(hipMalloc(&s->list, sizeof(unsigned) * (size_t)s->ntiles * nn));
            YK_HIPB(hipMalloc(&s->counts, sizeof(int2) * (size_t)s->ntiles));
        } else {
            s->ncl = (n + CL - 1) / CL;
            s->segcap = (s->ntiles + NSEG - 1) / NSEG * TILE;   // slots of a segment's tiles
            YK_HIPB(hipMalloc(&s->clist,
                              sizeof(unsigned) * (size_t)s->ncl * NSEG * s->segcap));
            YK_HIPB(hipMalloc(&s->ccounts, sizeof(int2) * (size_t)s->ncl * NSEG));
            YK_HIPB(hipMalloc(&s->corder, sizeof(int) * (size_t)s->ncl));
            YK_HIPB(hipMalloc(&s->tchg, sizeof(int) * (size_t)s->ntiles));
            // nonzero (every tile charged, the conservative cull) until the
            // first re-sort computes the flags
            YK_HIPB(hipMemsetAsync(s->tchg, 1, sizeof(int) * (size_t)s->ntiles, s->stream));
            if (s->list_M >= PART_REDUCE_MIN_M)
                YK_HIPB(hipMalloc(&s->psum, sizeof(double) * 11 * nn));
        }
    }
    YK_HIPB(hipMalloc(&s->maxdisp, 2 * sizeof(unsigned long long)));
    YK_HIPB(hipMemsetAsync(s->maxdisp, 0, 2 * sizeof(unsigned long long), s->stream));
    YK_HIPB(hipMalloc(&s->swept, sizeof(unsigned long long)));
    YK_HIPB(hipMemsetAsync(s->swept, 0, sizeof(unsigned long long), s->stream));
    YK_HIPB(hipMalloc(&s->part, sizeof(double) * s->part_doubles));
    YK_HIPB(hipMalloc(&s->blockpart, sizeof(double) * NQ * (size_t)s->nblocks));
    YK_HIPB(hipMalloc(&s->d_vstate, 4 * sizeof(double)));
    {
        std::vector<double> t(EXP_TN);
        exp_table(t.data());
        YK_HIPB(hipMalloc(&s->etab_g, sizeof(double) * EXP_TN));
        YK_HIPB(hipMemcpy(s->etab_g, t.data(), sizeof(double) * EXP_TN, hipMemcpyHostToDevice));
    }
    YK_HIPB(hipMalloc(&s->keys, sizeof(unsigned) * nn));
    YK_HIPB(hipMalloc(&s->keys2, sizeof(unsigned) * nn));
    YK_HIPB(hipMalloc(&s->vals, sizeof(int) * nn));
    YK_HIPB(hipMalloc(&s->vals2, sizeof(int) * nn));
    // sized for the keys' largest width (3g bits and the charge bit)
    YK_HIPB(hipcub::DeviceRadixSort::SortPairs(nullptr, s->sort_tmp_bytes, s->keys, s->keys2,
                                               s->vals, s->vals2, n, 0, 3 * s->grid_bits + 1,
                                               s->stream));
    if (s->use_clist && s->ncl > CORDER_MAX) {   // and the cluster order's sort
        size_t cbytes = 0;
        YK_HIPB(hipcub::DeviceRadixSort::SortPairs(nullptr, cbytes, s->keys, s->keys2, s->vals,
                                                   s->corder, s->ncl, 0, CORDER_KEY_BITS,
                                                   s->stream));
        s->sort_tmp_bytes = std::max(s->sort_tmp_bytes, cbytes);
    }
    YK_HIPB(hipMalloc(&s->sort_tmp, s->sort_tmp_bytes > 0 ? s->sort_tmp_bytes : 16));
    YK_HIPB(hipHostMalloc(&s->h_stage, sizeof(double) * 4 * nn, hipHostMallocDefault));
    YK_HIPB(hipHostMalloc(&s->h_orig, sizeof(int) * nn, hipHostMallocDefault));
    if ((rc = ensure_meas_cap(s, 1))) return bail(rc);
    // all particles at the origin until the caller uploads positions
    std::vector<double> zero(3 * nn, 0.0);
    if ((rc = yk_put_state(s, zero.data(), nullptr))) return bail(rc);
#undef YK_HIPB
    *out = s;
    return YK_OK;
}

void yk_destroy(yk_system *sys) { release(sys); }

int yk_set_params(yk_system *s, const yk_params *params)
{
    if (!s) return fail(YK_ERR_ARG, "handle is NULL");
    int rc = check_params(params);
    if (rc) return rc;
    if (params->sim_type != s->p.sim_type || params->precision != s->p.precision)
        return fail(YK_ERR_ARG, "sim_type and precision are fixed at yk_create");
    if (params->sim_type == YK_PURE_ION &&
        (params->ai != s->p.ai || params->xmass != s->p.xmass))
        return fail(YK_ERR_UNSUPPORTED, "pure-ion ai/xmass are fixed at yk_create");
    if (params->xmass != s->p.xmass)
        return fail(YK_ERR_UNSUPPORTED, "xmass is fixed at yk_create");
    if ((params->flags & YK_FLAG_ALLPAIRS) != (s->p.flags & YK_FLAG_ALLPAIRS) ||
        (params->flags & YK_FLAG_TILE_LISTS) != (s->p.flags & YK_FLAG_TILE_LISTS))
        return fail(YK_ERR_UNSUPPORTED, "YK_FLAG_ALLPAIRS / YK_FLAG_TILE_LISTS are fixed at yk_create");
    if (list_cutoff(*params) + 0.0 != list_cutoff(s->p) && s->use_list)
        s->skin = list_skin(*params);
    if ((rc = change_box(s, params->xl))) return rc;
    s->p = *params;
    set_near_radius(s);
    if (params->resort_every > 0) s->resort_every = params->resort_every;
    s->have_accel = false;
    return invalidate_list(s);
}

int yk_put_state(yk_system *s, const double *x, const double *v)
{
    if (!s || !x) return fail(YK_ERR_ARG, "handle or x is NULL");
    const int n = s->n;
    const size_t nn = (size_t)n;
    if (!finite3(x, n)) return fail(YK_ERR_NONFINITE, "positions contain NaN/Inf");
    if (v && !finite3(v, n)) return fail(YK_ERR_NONFINITE, "velocities contain NaN/Inf");
    YK_HIP(hipStreamSynchronize(s->stream));
    // caller order first (orig = identity), then Hilbert-sort on the device
    std::vector<double4> p4(nn);
    std::vector<int> iota(nn);
    for (size_t i = 0; i < nn; ++i) {
        p4[i] = make_double4(x[3 * i + 0], x[3 * i + 1], x[3 * i + 2], s->zhost[i]);
        iota[i] = (int)i;
    }
    YK_HIP(hipMemcpyAsync(s->pos, p4.data(), sizeof(double4) * nn, hipMemcpyHostToDevice,
                          s->stream));
    YK_HIP(hipMemcpyAsync(s->orig, iota.data(), sizeof(int) * nn, hipMemcpyHostToDevice,
                          s->stream));
    YK_HIP(hipMemcpyAsync(s->aii, s->ahost.data(), sizeof(double) * nn, hipMemcpyHostToDevice,
                          s->stream));
    YK_HIP(hipMemcpyAsync(s->den, s->dhost.data(), sizeof(double) * nn, hipMemcpyHostToDevice,
                          s->stream));
    double *hv = s->h_stage;
    if (v) {
        for (size_t i = 0; i < nn; ++i)
            for (int k = 0; k < 3; ++k) hv[k * nn + i] = v[3 * i + k];
        YK_HIP(hipMemcpyAsync(s->vel, hv, sizeof(double) * 3 * nn, hipMemcpyHostToDevice,
                              s->stream));
    } else {
        YK_HIP(hipMemsetAsync(s->vel, 0, sizeof(double) * 3 * nn, s->stream));
    }
    YK_HIP(hipMemsetAsync(s->vold, 0, sizeof(double) * 3 * nn, s->stream));
    YK_HIP(hipMemsetAsync(s->acc, 0, sizeof(double) * 3 * nn, s->stream));
    int rc = resort(s);
    if (rc) return rc;
    YK_HIP(hipStreamSynchronize(s->stream));
    s->have_accel = false;
    s->steps_done = 0;
    return YK_OK;
}

int yk_put_accel(yk_system *s, const double *a)
{
    if (!s || !a) return fail(YK_ERR_ARG, "handle or a is NULL");
    const size_t nn = (size_t)s->n;
    if (!finite3(a, s->n)) return fail(YK_ERR_NONFINITE, "accelerations contain NaN/Inf");
    YK_HIP(hipStreamSynchronize(s->stream));
    YK_HIP(hipMemcpy(s->h_orig, s->orig, sizeof(int) * nn, hipMemcpyDeviceToHost));
    double *h = s->h_stage;
    for (size_t i = 0; i < nn; ++i)
        for (int k = 0; k < 3; ++k) h[k * nn + i] = a[3 * (size_t)s->h_orig[i] + k];
    YK_HIP(hipMemcpyAsync(s->acc, h, sizeof(double) * 3 * nn, hipMemcpyHostToDevice,
                          s->stream));
    YK_HIP(hipStreamSynchronize(s->stream));
    s->have_accel = true;
    return YK_OK;
}

int yk_get_state(yk_system *s, double *x, double *v, double *a)
{
    if (!s) return fail(YK_ERR_ARG, "handle is NULL");
    const size_t nn = (size_t)s->n;
    YK_HIP(hipStreamSynchronize(s->stream));
    YK_HIP(hipMemcpy(s->h_orig, s->orig, sizeof(int) * nn, hipMemcpyDeviceToHost));
    const int *o = s->h_orig;
    if (x) {
        std::vector<double4> p4(nn);
        YK_HIP(hipMemcpy(p4.data(), s->pos, sizeof(double4) * nn, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < nn; ++i) {
            double *xi = x + 3 * (size_t)o[i];
            const double xs[3] = {p4[i].x, p4[i].y, p4[i].z};
            for (int k = 0; k < 3; ++k) xi[k] = wrap_into_box(xs[k], s->p.xl[k]);
        }
    }
    double *h = s->h_stage;
    if (v) {
        YK_HIP(hipMemcpy(h, s->vel, sizeof(double) * 3 * nn, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < nn; ++i)
            for (int k = 0; k < 3; ++k) v[3 * (size_t)o[i] + k] = h[k * nn + i];
    }
    if (a) {
        YK_HIP(hipMemcpy(h, s->acc, sizeof(double) * 3 * nn, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < nn; ++i)
            for (int k = 0; k < 3; ++k) a[3 * (size_t)o[i] + k] = h[k * nn + i];
    }
    return YK_OK;
}

int yk_accel(yk_system *s, int32_t do_measurements, yk_meas *out)
{
    if (!s) return fail(YK_ERR_ARG, "handle is NULL");
    const bool meas = do_measurements != 0;
    int rc = launch_pairs(s, meas);
    if (rc) return rc;
    if ((rc = launch_finalize(s, 0.0, meas, 0, 0))) return rc;
    if ((rc = timing_collect(s))) return rc;
    s->have_accel = true;
    if (meas) {
        YK_HIP(hipMemcpyAsync(s->h_meas, s->d_meas, sizeof(MeasOut), hipMemcpyDeviceToHost,
                              s->stream));
        YK_HIP(hipStreamSynchronize(s->stream));
        if (out) to_meas(s->h_meas[0], out);
    } else {
        YK_HIP(hipStreamSynchronize(s->stream));
        zero_meas(out);
    }
    return YK_OK;
}

int yk_newton(yk_system *s, double dt, int32_t do_measurements, yk_meas *out)
{
    if (!s) return fail(YK_ERR_ARG, "handle is NULL");
    if (!std::isfinite(dt)) return fail(YK_ERR_ARG, "dt=%g is not finite", dt);
    if (!s->have_accel)
        return fail(YK_ERR_STATE, "newton needs accelerations of the current positions: "
                                  "call yk_accel first (src/md.F:69)");
    const bool meas = do_measurements != 0;
    int rc = device_step(s, dt, meas, 0);
    if (rc) return rc;
    if ((rc = timing_collect(s))) return rc;
    if (meas) {
        YK_HIP(hipMemcpyAsync(s->h_meas, s->d_meas, sizeof(MeasOut), hipMemcpyDeviceToHost,
                              s->stream));
        YK_HIP(hipStreamSynchronize(s->stream));
        if (out) to_meas(s->h_meas[0], out);
    } else {
        YK_HIP(hipStreamSynchronize(s->stream));
        zero_meas(out);
    }
    return YK_OK;
}

int yk_run(yk_system *s, double dt, int32_t nsteps, int32_t nind, yk_meas *records)
{
    if (!s) return fail(YK_ERR_ARG, "handle is NULL");
    if (nsteps < 0) return fail(YK_ERR_ARG, "nsteps=%d < 0", nsteps);
    if (!std::isfinite(dt)) return fail(YK_ERR_ARG, "dt=%g is not finite", dt);
    if (!s->have_accel)
        return fail(YK_ERR_STATE, "run needs accelerations of the current positions: "
                                  "call yk_accel first (src/md.F:69)");
    const int nrec = (nind > 0) ? nsteps / nind : 0;
    int rc = ensure_meas_cap(s, nrec > 0 ? nrec : 1);
    if (rc) return rc;
    const bool timing = (s->p.flags & YK_FLAG_TIMING) != 0;
    if (timing) {
        if (!s->step0) YK_HIP(hipEventCreate(&s->step0));
        if (!s->step1) YK_HIP(hipEventCreate(&s->step1));
        YK_HIP(hipEventRecord(s->step0, s->stream));
    }
    bool ahead = false;   // this step's drift was done by the previous finalize
    for (int st = 1; st <= nsteps; ++st) {
        const bool meas = (nind > 0) && (st % nind == 0);
        if ((rc = device_step(s, dt, meas, meas ? st / nind - 1 : 0, ahead, st < nsteps, &ahead)))
            return rc;
    }
    if (timing) {
        YK_HIP(hipEventRecord(s->step1, s->stream));
        YK_HIP(hipEventSynchronize(s->step1));
        float ms = 0.f;
        YK_HIP(hipEventElapsedTime(&ms, s->step0, s->step1));
        s->timing.step_ms += ms;
        s->timing.steps += nsteps;
    }
    if ((rc = timing_collect(s))) return rc;
    if (nrec > 0 && records) {
        YK_HIP(hipMemcpyAsync(s->h_meas, s->d_meas, sizeof(MeasOut) * (size_t)nrec,
                              hipMemcpyDeviceToHost, s->stream));
    }
    YK_HIP(hipStreamSynchronize(s->stream));
    if (nrec > 0 && records)
        for (int r = 0; r < nrec; ++r) to_meas(s->h_meas[r], &records[r]);
    return YK_OK;
}

int yk_ttot(yk_system *s, double *eka)
{
    if (!s || !eka) return fail(YK_ERR_ARG, "handle or eka is NULL");
    hipLaunchKernelGGL(k_ttot, dim3(s->nblocks), dim3(BLOCK), 0, s->stream, s->vel, s->aii,
                       s->n, s->blockpart);
    YK_HIP(hipGetLastError());
    MeasConsts mc = meas_consts(s, MEAS_TTOT);
    hipLaunchKernelGGL(k_meas_final, dim3(1), dim3(BLOCK), 0, s->stream, s->blockpart,
                       s->nblocks, mc, s->d_meas);
    YK_HIP(hipGetLastError());
    YK_HIP(hipMemcpyAsync(s->h_meas, s->d_meas, sizeof(MeasOut), hipMemcpyDeviceToHost,
                          s->stream));
    YK_HIP(hipStreamSynchronize(s->stream));
    *eka = s->h_meas[0].ek;
    return YK_OK;
}

int yk_tnorm(yk_system *s)
{
    if (!s) return fail(YK_ERR_ARG, "handle is NULL");
    if (!(s->p.kT > 0.0)) return fail(YK_ERR_ARG, "tnorm needs kT > 0 (kT=%g)", s->p.kT);
    return do_tnorm(s);
}

int yk_cancel_vcom(yk_system *s)
{
    if (!s) return fail(YK_ERR_ARG, "handle is NULL");
    if (!(s->p.kT > 0.0)) return fail(YK_ERR_ARG, "cancel_vcom needs kT > 0 (kT=%g)", s->p.kT);
    return do_cancel_vcom(s);
}

int yk_strain(yk_system *s, double dt, const double *deps, const double *xl0, double *xl_out)
{
    if (!s || !deps || !xl0) return fail(YK_ERR_ARG, "handle, deps or xl0 is NULL");
    int rc = do_strain(s, dt, deps, xl0);
    if (rc) return rc;
    if (xl_out)
        for (int k = 0; k < 3; ++k) xl_out[k] = s->p.xl[k];
    return YK_OK;
}

int yk_run_md(yk_system *s, const yk_loop *lp, yk_meas *records, int32_t *rec_steps,
              int32_t max_records, int32_t *nrec, double *xl_out)
{
    if (!s || !lp) return fail(YK_ERR_ARG, "handle or loop is NULL");
    if (lp->nsteps < 0 || lp->istep0 < 0 || lp->nind < 0 || lp->nptensor < 0 || lp->ncom < 0 ||
        lp->ntnorm < 0)
        return fail(YK_ERR_ARG, "negative step count or period");
    if (!std::isfinite(lp->dt)) return fail(YK_ERR_ARG, "dt=%g is not finite", lp->dt);
    if (!s->have_accel)
        return fail(YK_ERR_STATE, "run needs accelerations of the current positions: "
                                  "call yk_accel first (src/md.F:69)");
    if ((lp->ncom > 0 || lp->ntnorm > 0) && !(s->p.kT > 0.0))
        return fail(YK_ERR_ARG, "cancel_vcom/tnorm need kT > 0 (kT=%g)", s->p.kT);
    // measuring steps, known in advance
    int count = 0;
    for (int st = 0; st < lp->nsteps; ++st) {
        const int is = lp->istep0 + st;
        if (mdmod(is + 1, lp->nind) == 0 || mdmod(is + 1, lp->nptensor) == 0) ++count;
    }
    if (records && count > max_records)
        return fail(YK_ERR_ARG, "%d measuring steps, records holds %d", count, max_records);
    int rc = ensure_meas_cap(s, count > 0 ? count : 1);
    if (rc) return rc;
    const bool timing = (s->p.flags & YK_FLAG_TIMING) != 0;
    if (timing) {
        if (!s->step0) YK_HIP(hipEventCreate(&s->step0));
        if (!s->step1) YK_HIP(hipEventCreate(&s->step1));
        YK_HIP(hipEventRecord(s->step0, s->stream));
    }
    int slot = 0;
    const bool straining = lp->deps[0] != YK_XNOSTRAIN || lp->deps[1] != YK_XNOSTRAIN ||
                           lp->deps[2] != YK_XNOSTRAIN;
    bool ahead = false;   // this step's drift was done by the previous finalize
    for (int st = 0; st < lp->nsteps; ++st) {   // src/md.F:124-133
        const int is = lp->istep0 + st;
        // (with ahead set none of these is due: the previous step checked)
        if (mdmod(is, lp->ncom) == 0 && (rc = do_cancel_vcom(s))) return rc;
        if (mdmod(is, lp->ntnorm) == 0 && (rc = do_tnorm(s))) return rc;
        if ((rc = do_strain(s, lp->dt, lp->deps, lp->xl0))) return rc;
        const bool meas = mdmod(is + 1, lp->nind) == 0 || mdmod(is + 1, lp->nptensor) == 0;
        const bool next_plain = st + 1 < lp->nsteps && !straining &&
                                mdmod(is + 1, lp->ncom) != 0 && mdmod(is + 1, lp->ntnorm) != 0;
        if ((rc = device_step(s, lp->dt, meas, meas ? slot : 0, ahead, next_plain, &ahead)))
            return rc;
        if (meas) {
            if (rec_steps && slot < max_records) rec_steps[slot] = is + 1;
            ++slot;
        }
    }
    if (timing) {
        YK_HIP(hipEventRecord(s->step1, s->stream));
        YK_HIP(hipEventSynchronize(s->step1));
        float ms = 0.f;
        YK_HIP(hipEventElapsedTime(&ms, s->step0, s->step1));
        s->timing.step_ms += ms;
        s->timing.steps += lp->nsteps;
    }
    if ((rc = timing_collect(s))) return rc;
    if (count > 0 && records)
        YK_HIP(hipMemcpyAsync(s->h_meas, s->d_meas, sizeof(MeasOut) * (size_t)count,
                              hipMemcpyDeviceToHost, s->stream));
    YK_HIP(hipStreamSynchronize(s->stream));
    if (count > 0 && records)
        for (int r = 0; r < count; ++r) to_meas(s->h_meas[r], &records[r]);
    if (nrec) *nrec = count;
    if (xl_out)
        for (int k = 0; k < 3; ++k) xl_out[k] = s->p.xl[k];
    return YK_OK;
}

int yk_get_timing(yk_system *s, yk_timing *out)
{
    if (!s || !out) return fail(YK_ERR_ARG, "handle or out is NULL");
    int rc = timing_collect(s);
    if (rc) return rc;
    unsigned long long sw = 0;
    if ((rc = read_swept(s, &sw))) return rc;
    *out = s->timing;
    out->pairs_swept = (int64_t)(sw - s->swept_base);
    return YK_OK;
}

int yk_reset_timing(yk_system *s)
{
    if (!s) return fail(YK_ERR_ARG, "handle is NULL");
    int rc = timing_collect(s);
    if (rc) return rc;
    if ((rc = read_swept(s, &s->swept_base))) return rc;
    s->timing = yk_timing{};
    return YK_OK;
}

int yk_accel_noalloc(const yk_params *params, int32_t n, const double *x, const double *zii,
                     const double *aii, double *a, int32_t do_measurements, yk_meas *meas,
                     void *stream)
{
    int rc = check_params(params);
    if (rc) return rc;
    if (n < 1 || n > NA_MAX_N)
        return fail(YK_ERR_ARG, "n=%d is outside [1, %d], the no-allocation limit", n, NA_MAX_N);
    if (!x || !a) return fail(YK_ERR_ARG, "x and a are required");
    const int kind = params->sim_type;
    if (kind != YK_PURE_ION && (!zii || !aii))
        return fail(YK_ERR_ARG, "zii and aii are required for sim_type %d", kind);
    const size_t nn = (size_t)n;
    if ((rc = na_check_range(x, sizeof(double) * 3 * nn, "x")) ||
        (rc = na_check_range(a, sizeof(double) * 3 * nn, "a")) ||
        (zii && (rc = na_check_range(zii, sizeof(double) * nn, "zii"))) ||
        (aii && (rc = na_check_range(aii, sizeof(double) * nn, "aii"))) ||
        (meas && (rc = na_check_range(meas, sizeof(yk_meas), "meas"))))
        return rc;
    unsigned char *base = nullptr;
    if ((rc = na_scratch(&base))) return rc;
    double4 *pos = reinterpret_cast<double4 *>(base + NA.pos);
    double *den = reinterpret_cast<double *>(base + NA.den);
    int *orig = reinterpret_cast<int *>(base + NA.orig);
    TileBox *box = reinterpret_cast<TileBox *>(base + NA.box);
    double *part = reinterpret_cast<double *>(base + NA.part);
    double *blockpart = reinterpret_cast<double *>(base + NA.blockpart);
    double *acc = reinterpret_cast<double *>(base + NA.acc);
    double *etab = reinterpret_cast<double *>(base + NA.etab);
    auto *swept = reinterpret_cast<unsigned long long *>(base + NA.swept);
    // the tiled all-pairs decomposition of choose_decomposition, at most
    // NA_MAX_CHUNKS source chunks
    const int nblocks = (n + BLOCK - 1) / BLOCK, ntiles = (n + TILE - 1) / TILE;
    const int nstages = (n + STAGE - 1) / STAGE;
    int want = (2048 + nblocks - 1) / nblocks;
    want = std::max(1, std::min(want, std::min(nstages, NA_MAX_CHUNKS)));
    const int spc = (nstages + want - 1) / want;
    const int nchunks = (nstages + spc - 1) / spc;
    const bool m = do_measurements != 0;
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_na_stage, dim3(nblocks), dim3(BLOCK), 0, st, x, zii, aii, n, kind,
                       params->zi, params->ai, params->xmass, pos, den, orig);
    hipLaunchKernelGGL(k_tile_boxes, dim3(nblocks), dim3(BLOCK), 0, st, pos, n, ntiles, box);
    const bool mixed = params->precision == YK_MIXED;
    if (n < 2 || kind == YK_BFIELD_TEST) {   // no pairs: zero partial sums
        const size_t cnt = (size_t)nchunks * 11 * nn;
        hipLaunchKernelGGL(k_na_fill, dim3((unsigned)((cnt + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0,
                           st, part, cnt);
    } else {
        if (!mixed)
            hipLaunchKernelGGL(k_na_exp_table, dim3(EXP_TN / BLOCK), dim3(BLOCK), 0, st, etab);
        const ForceConsts c = force_consts_of(*params, n, ntiles, etab);
        switch (kind) {
        case YK_PURE_ION:
            mixed ? na_launch_pairs<PURE, 1>(st, pos, box, c, nblocks, nchunks, spc, part, swept, m)
                  : na_launch_pairs<PURE, 0>(st, pos, box, c, nblocks, nchunks, spc, part, swept, m);
            break;
        case YK_ION_MIXTURE:
            mixed ? na_launch_pairs<MIX, 1>(st, pos, box, c, nblocks, nchunks, spc, part, swept, m)
                  : na_launch_pairs<MIX, 0>(st, pos, box, c, nblocks, nchunks, spc, part, swept, m);
            break;
        default:
            mixed ? na_launch_pairs<NN, 1>(st, pos, box, c, nblocks, nchunks, spc, part, swept, m)
                  : na_launch_pairs<NN, 0>(st, pos, box, c, nblocks, nchunks, spc, part, swept, m);
            break;
        }
    }
    const StepConsts sc = step_consts_of(*params, n, 0.0);
    switch (kind) {
    case YK_PURE_ION:
        na_launch_finalize<PURE>(st, nblocks, part, nchunks, pos, den, orig, acc, sc, blockpart, m);
        break;
    case YK_ION_MIXTURE:
        na_launch_finalize<MIX>(st, nblocks, part, nchunks, pos, den, orig, acc, sc, blockpart, m);
        break;
    default:
        na_launch_finalize<NN>(st, nblocks, part, nchunks, pos, den, orig, acc, sc, blockpart, m);
        break;
    }
    if (meas) {
        if (m)
            hipLaunchKernelGGL(k_meas_final, dim3(1), dim3(BLOCK), 0, st, blockpart, nblocks,
                               meas_consts_of(*params, n, MEAS_ACCEL),
                               reinterpret_cast<MeasOut *>(meas));
        else
            hipLaunchKernelGGL(k_na_zero_meas, dim3(1), dim3(64), 0, st, meas);
    }
    hipLaunchKernelGGL(k_na_out, dim3(nblocks), dim3(BLOCK), 0, st, acc, n, a);
    YK_HIP(hipGetLastError());
    return YK_OK;
}

}  // extern "C"
