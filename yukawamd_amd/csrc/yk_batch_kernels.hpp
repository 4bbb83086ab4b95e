// yk_batch_kernels.hpp -- kernels of the batched ensemble engine: many
// independent ion replicas (pure or mixture, FP64 or mixed precision, each
// with its own N, box, screening, cutoff and timestep) integrated together.
//
// Arena layout (one per engine): replica r owns slots [off_r, off_r + n_r)
// with off_r a multiple of 256, so every 64-slot tile and every 256-slot
// finalize block belongs to exactly one replica; unused slots carry rep = -1.
// Per slot: pos (double4 {x,y,z,Z}), vel/vold/acc (component-major over the
// arena capacity), aii, den, orig (caller index inside the replica).
//
// Pair forces: one persistent launch.  Work items (replica, target tile,
// range of source tiles of the same replica, or a share of a tile's cluster
// lists) are pulled from a device queue split into QPARTS counters: the
// workgroups with blockIdx.x = p (mod QPARTS) take the queue positions
// p, p + QPARTS, p + 2 QPARTS, ... in order, one atomic per item on their own
// counter (an atomic on one address from every CU is served one at a time
// beyond the XCDs' L2s, about 25 ns each; workgroups go to the 8 XCDs round
// robin, so each counter serves about one XCD).  Each wave owns one item at a
// time, stages source tiles in its own LDS slice, culls them by AABB and
// takes the shift-only minimum image when the two tiles allow it.  Every item writes its own
// partial sums, and the finalize kernel adds a tile's items in item order,
// so a replica's result does not depend on which wave ran which item nor on
// the other replicas in the launch.
#pragma once
#include "yk_batch_plan.hpp"
#include "yk_kernels.hpp"

namespace yk {

struct RepConsts {
    double L[3], invL[3], halfL_m[3];
    double rc2, mu, evcut, cnum, margin, cull2;
    double dt, afac, rho, kT, xmass, vol;
    float muf, mul2ef, mul2e_lo;   // mul2e_lo: as ForceConsts::mul2e_lo
    float three_n, n_f;
    int kind, prec, off, n;
    int tile0, ntiles, blk0, nblk;
    int ref_f1, grid_bits;   // grid_bits: Hilbert cells per axis = 2^grid_bits
    // cluster lists (list_items > 0): the replica's region of the list arena
    // (entries), segment capacity, list radius^2, skin and the FP32 bounds of
    // the candidate test (as ListConsts)
    long long lists_off;
    int list_items, segcap;
    double rl2, skin;
    float Lf[3], invLf[3], halfLf_m[3], rl2f, skinf;
    float rn2f;                    // near field of the mixed form (ListConsts::rn2f)
    double fs, fs2;                       // as ForceConsts
    float rc2f_lo, rc2f_hi, r2_floor_f;   // as ForceConsts
};

struct RepState {            // device-updated per replica
    int steps_left;          // newton steps still to do (0 = idle)
    int rec_count;           // measurement records written
    int fresh;               // 1 until the accel pass has evaluated the admitted state
    int pad0;
};

// The pair launch's work queue: QPARTS counters, QSTRIDE unsigneds apart (one
// 256-byte line each).
constexpr int QPARTS = 8;
constexpr int QSTRIDE = 64;

struct BatchItem {
    // tile item: target tile and source tiles [s0, s1), global indices;
    // list item (s1 < 0): target tile's four clusters, share s0 of -s1
    int rep, tile, s0, s1;
};

// ---------------------------------------------------------------------------
template <int KIND, int PREC, bool MEAS>
__device__ __forceinline__ void batch_item(const RepConsts &rc, const ExpCoeffs &ec,
                                           const double *__restrict__ etab,
                                           const double4 *__restrict__ pos,
                                           const TileBox *__restrict__ box, const BatchItem &it,
                                           double4 *__restrict__ sw, int lane,
                                           double *__restrict__ out)
{
    ForceConsts c{};
    c.etab = etab;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        c.L[k] = rc.L[k];
        c.invL[k] = rc.invL[k];
        c.halfL_m[k] = rc.halfL_m[k];
    }
    c.rc2 = rc.rc2;
    c.mu = rc.mu;
    c.evcut = rc.evcut;
    c.muf = rc.muf;
    c.mul2ef = rc.mul2ef;
    c.mul2e_lo = rc.mul2e_lo;
    c.ec = ec;
    if constexpr (PREC == 0) pin_vgpr(c.ec.shifter);
    const int t = it.tile;
    const int i = t * TILE + lane;
    const int ilocal = i - rc.off;
    const double nanv = __builtin_nan("");
    const double4 pi = (ilocal < rc.n) ? pos[i] : make_double4(nanv, nanv, nanv, 0.0);
    double ic[3], ih[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        ic[k] = uniform(box[t].c[k]);
        ih[k] = uniform(box[t].h[k]);
    }
    Acc<MEAS> acc;
    for (int st = it.s0; st < it.s1; ++st) {
        double gap2 = 0.0, shift[3];
        bool fast = true;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            double cd = ic[k] - box[st].c[k];
            double h = ih[k] + box[st].h[k] + rc.margin;
            double s = c.L[k] * __builtin_rint(cd * c.invL[k]);
            double cm = __builtin_fabs(cd - s);
            double g = cm - h;
            g = g > 0.0 ? g : 0.0;
            gap2 = __builtin_fma(g, g, gap2);
            fast = fast && (cm + h < c.halfL_m[k]);
            shift[k] = s;
        }
        gap2 = uniform(gap2);
        if (gap2 > rc.cull2) continue;
        const int j = st * TILE + lane;
        double4 pjs = (j - rc.off < rc.n) ? pos[j] : make_double4(nanv, nanv, nanv, 0.0);
        pjs.w = stage_w<KIND, PREC>(pjs.w, c.muf);
        sw[lane] = pjs;
        __builtin_amdgcn_wave_barrier();
        const int self_jj = (st == t) ? lane : -1;
        const int jcount = min(TILE, rc.n - (st * TILE - rc.off));   // padded sources skipped
        if (uniform(fast ? 1.0 : 0.0) != 0.0) {
            const double xs = pi.x - uniform(shift[0]);
            const double ys = pi.y - uniform(shift[1]);
            const double zs = pi.z - uniform(shift[2]);
#pragma unroll 4
            for (int jj = 0; jj < jcount; ++jj) {
                const double4 pj = sw[jj];
                interact<KIND, PREC, MEAS>(c, pi.w, pj.w, xs - pj.x, ys - pj.y, zs - pj.z,
                                           jj == self_jj, acc);
            }
        } else {
#pragma unroll 2
            for (int jj = 0; jj < jcount; ++jj) {
                const double4 pj = sw[jj];
                interact<KIND, PREC, MEAS>(c, pi.w, pj.w,
                                           min_image(pi.x - pj.x, c.L[0], c.invL[0]),
                                           min_image(pi.y - pj.y, c.L[1], c.invL[1]),
                                           min_image(pi.z - pj.z, c.L[2], c.invL[2]),
                                           jj == self_jj, acc);
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
    double *o = out + lane;
    o[0] = acc.fx; o[TILE] = acc.fy; o[2 * TILE] = acc.fz;
    if constexpr (MEAS) {
        o[3 * TILE] = acc.e; o[4 * TILE] = acc.w;
        o[5 * TILE] = acc.pxx; o[6 * TILE] = acc.pxy; o[7 * TILE] = acc.pxz;
        o[8 * TILE] = acc.pyy; o[9 * TILE] = acc.pyz; o[10 * TILE] = acc.pzz;
    }
}

// One list item: the four 16-slot clusters of target tile it.tile against
// share m of M of each cluster's list (clist_share, as k_pairs_clist, with the
// replica's constants and its region of the list arena).
template <int KIND, int PREC, bool MEAS>
__device__ __forceinline__ void batch_list_item(const RepConsts &rc, const ExpCoeffs &ec,
                                                const double *__restrict__ etab,
                                                const double4 *__restrict__ pos,
                                                const unsigned *__restrict__ clist,
                                                const int2 *__restrict__ ccounts,
                                                const BatchItem &it, double4 *__restrict__ sw,
                                                float4 *__restrict__ sf, SegList *__restrict__ hdr,
                                                int lane, double *__restrict__ out)
{
    ForceConsts c{};
    c.etab = etab;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        c.L[k] = rc.L[k];
        c.invL[k] = rc.invL[k];
        c.halfL_m[k] = rc.halfL_m[k];
    }
    c.rc2 = rc.rc2;
    c.mu = rc.mu;
    c.evcut = rc.evcut;
    c.muf = rc.muf;
    c.mul2ef = rc.mul2ef;
    c.mul2e_lo = rc.mul2e_lo;
    c.fs = rc.fs;
    c.fs2 = rc.fs2;
    c.rc2f_lo = rc.rc2f_lo;
    c.rc2f_hi = rc.rc2f_hi;
    c.r2_floor_f = rc.r2_floor_f;
    c.ec = ec;
    pin_vgpr(c.ec.shifter);   // both forms have an FP64 pair path
    const int M = -it.s1, m = it.s0;
    const int g = lane / CL;
    const int t = it.tile;
    const int i = t * TILE + lane;
    const bool ivalid = i - rc.off < rc.n;
    const double nanv = __builtin_nan("");
    const double4 pi = ivalid ? pos[i] : make_double4(nanv, nanv, nanv, 0.0);
    double cc[3] = {0.0, 0.0, 0.0};
    if constexpr (PREC == 1) group_centre(pi, ivalid, cc);
    const int clg = t * CPW + g;                    // global cluster (ccounts index)
    const int cll = (t - rc.tile0) * CPW + g;       // cluster within the replica
    const bool cvalid = cll * CL < rc.n;            // clusters past n have no list
    int2 sc[NSEG];
    int total = 0;
#pragma unroll
    for (int sg = 0; sg < NSEG; ++sg) {
        sc[sg] = cvalid ? ccounts[(size_t)clg * NSEG + sg] : make_int2(0, 0);
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
    const unsigned *mine = clist + rc.lists_off + (size_t)cll * NSEG * rc.segcap;
    const SegList sl = seg_list(mine, rc.segcap, sc, hi);
    Acc<MEAS> acc;
    clist_share<KIND, PREC, MEAS>(c, pi, cc, sl, lo, iters, pos + rc.off, sw, sf, hdr, lane, acc);
    double *o = out + lane;
    o[0] = acc.fx; o[TILE] = acc.fy; o[2 * TILE] = acc.fz;
    if constexpr (MEAS) {
        o[3 * TILE] = acc.e; o[4 * TILE] = acc.w;
        o[5 * TILE] = acc.pxx; o[6 * TILE] = acc.pxy; o[7 * TILE] = acc.pxz;
        o[8 * TILE] = acc.pyy; o[9 * TILE] = acc.pyz; o[10 * TILE] = acc.pzz;
    }
}

// Persistent: every wave pulls queue positions from its workgroup's part of
// the queue until that part is empty (qorder: position -> item).
// accel_pass != 0 evaluates only replicas admitted since the last pass
// (fresh); otherwise replicas whose step budget is used up are skipped.
template <bool MEAS>
__global__ __launch_bounds__(512) void k_batch_pairs(const double4 *__restrict__ pos,
                                                     const TileBox *__restrict__ box,
                              const RepConsts *__restrict__ reps,
                              const RepState *__restrict__ state,
                              const BatchItem *__restrict__ items,
                              const int *__restrict__ qorder, int nitems,
                              unsigned *__restrict__ queue, ExpCoeffs ec,
                              const double *__restrict__ etab_g, int accel_pass,
                              const unsigned *__restrict__ clist,
                              const int2 *__restrict__ ccounts, double *__restrict__ part)
{
    // dynamic LDS: per wave 64 double4 sources, then per wave 64 float4 (the
    // mixed list items' FP32 offsets)
    extern __shared__ double4 sj_dyn[];
    __shared__ double etab_s[EXP_TN];   // FP64 replicas' exp table, one per workgroup
    __shared__ SegList hdr_s[512 / TILE][CPW];   // list items' list headers, per wave
    const int lane = threadIdx.x & (TILE - 1);
    SegList *hdr = hdr_s[threadIdx.x / TILE];
    double4 *sw = sj_dyn + (threadIdx.x / TILE) * TILE;
    float4 *sf = reinterpret_cast<float4 *>(sj_dyn + blockDim.x) + (threadIdx.x / TILE) * TILE;
    load_exp_table(etab_s, etab_g);
    const double *etab = etab_s;
    constexpr int NC = MEAS ? 11 : 3;
    // parts actually served: a launch of fewer than QPARTS workgroups uses
    // as many parts as it has workgroups
    const int parts = gridDim.x < (unsigned)QPARTS ? (int)gridDim.x : QPARTS;
    const int part_id = (int)(blockIdx.x % (unsigned)parts);
    unsigned *counter = queue + part_id * QSTRIDE;
    for (;;) {
        unsigned j = 0;
        if (lane == 0) j = atomicAdd(counter, 1u);
        j = (unsigned)__builtin_amdgcn_readfirstlane((int)__shfl(j, 0, 64));
        const long long k = (long long)part_id + (long long)parts * j;
        if (k >= nitems) break;   // every wave reaches this exit
        const int q = qorder[k];             // a permutation of [0, nitems)
        const BatchItem it = items[q];
        double *out = part + (size_t)q * NC * TILE;
        const RepConsts &rc = reps[it.rep];
        if (accel_pass ? state[it.rep].fresh == 0 : state[it.rep].steps_left <= 0) {
            // replica not part of this pass: leave zero partials
            out[lane] = 0.0; out[TILE + lane] = 0.0; out[2 * TILE + lane] = 0.0;
            if constexpr (MEAS)
                for (int q = 3; q < NC; ++q) out[q * TILE + lane] = 0.0;
            continue;
        }
        const int kind = __builtin_amdgcn_readfirstlane(rc.kind);
        const int prec = __builtin_amdgcn_readfirstlane(rc.prec);
        if (it.s1 < 0) {
            if (kind == PURE) {
                if (prec == 0)
                    batch_list_item<PURE, 0, MEAS>(rc, ec, etab, pos, clist, ccounts, it, sw, sf, hdr, lane, out);
                else
                    batch_list_item<PURE, 1, MEAS>(rc, ec, etab, pos, clist, ccounts, it, sw, sf, hdr, lane, out);
            } else {
                if (prec == 0)
                    batch_list_item<MIX, 0, MEAS>(rc, ec, etab, pos, clist, ccounts, it, sw, sf, hdr, lane, out);
                else
                    batch_list_item<MIX, 1, MEAS>(rc, ec, etab, pos, clist, ccounts, it, sw, sf, hdr, lane, out);
            }
            continue;
        }
        if (kind == PURE) {
            if (prec == 0) batch_item<PURE, 0, MEAS>(rc, ec, etab, pos, box, it, sw, lane, out);
            else batch_item<PURE, 1, MEAS>(rc, ec, etab, pos, box, it, sw, lane, out);
        } else {
            if (prec == 0) batch_item<MIX, 0, MEAS>(rc, ec, etab, pos, box, it, sw, lane, out);
            else batch_item<MIX, 1, MEAS>(rc, ec, etab, pos, box, it, sw, lane, out);
        }
    }
}

// ---------------------------------------------------------------------------
// Per-slot drift (vel_verlet_2 order) + tile AABBs.  Positions stay
// continuous on the device (no wrap into the box): forces see only minimum-
// image differences, and a tile whose particle crosses a box face keeps a
// compact AABB instead of one spanning the box.  Positions are wrapped into
// the box when they are downloaded.
// MOVE = false only recomputes the tile AABBs (before the initial accel pass).
template <bool MOVE>
__global__ __launch_bounds__(BLOCK) void k_batch_drift(double4 *__restrict__ pos,
                                                       double *__restrict__ vel,
                                                       double *__restrict__ vold,
                                                       const double *__restrict__ acc,
                                                       const int *__restrict__ rep,
                                                       const RepConsts *__restrict__ reps,
                                                       const RepState *__restrict__ state,
                                                       int nslots, int cap,
                                                       TileBox *__restrict__ box,
                                                       const int *__restrict__ tile_rep,
                                                       const double4 *__restrict__ xref,
                                                       unsigned long long *__restrict__ maxdisp)
{
    // slots [0, nslots) are the used extent of the arena; cap is the stride of
    // the component-major arrays
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    const size_t nn = (size_t)cap;
    const int r = (i < nslots) ? rep[i] : -1;
    const bool valid = (r >= 0);
    double x[3] = {0, 0, 0};
    if (valid) {
        const RepConsts &rc = reps[r];
        double4 p = pos[i];
        x[0] = p.x; x[1] = p.y; x[2] = p.z;
        if (MOVE && state[r].steps_left > 0) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                double v = vel[k * nn + i], a = acc[k * nn + i];
                vold[k * nn + i] = v;
                vel[k * nn + i] = v + rc.dt * a;
                x[k] = x[k] + v * rc.dt + rc.afac * a;   // src/vel_verlet_2.f:38
            }
            pos[i] = make_double4(x[0], x[1], x[2], p.w);
        }
    }
    const int tile = i / TILE;
    unsigned long long *md = nullptr;   // wave-uniform
    double d2 = 0.0;
    if (tile < nslots / TILE) {
        tile_box_store(x, valid, tile, box);
        // largest displacement since the replica's list build (list replicas)
        const int tr = tile_rep[tile];   // one replica per tile: wave-uniform
        if (MOVE && tr >= 0 && reps[tr].list_items > 0) {
            md = maxdisp + 2 * tr;
            if (valid) {
                const double4 q = xref[i];
                const double dx = x[0] - q.x, dy = x[1] - q.y, dz = x[2] - q.z;
                d2 = __builtin_fma(dx, dx, __builtin_fma(dy, dy, dz * dz));
            }
        }
    }
    if (MOVE) record_disp(md, d2);   // every thread of the workgroup
}

// Cluster lists of the list replicas (grid: arena tiles x NSEG segments),
// rebuilt for a replica whose lists are stale (lists_stale over its two
// largest displacements since its last build; +inf at admission and after a
// re-sort).  Same
// scheme as k_clist_build with the replica's slot range, constants and list
// region; entries hold slots relative to the replica's first slot.
__global__ __launch_bounds__(BLOCK) void k_batch_clist_build(
    const double4 *__restrict__ pos, const TileBox *__restrict__ box,
    const int *__restrict__ list_tiles, const int *__restrict__ tile_rep,
    const RepConsts *__restrict__ reps, const unsigned long long *__restrict__ maxdisp,
    unsigned *__restrict__ clist, int2 *__restrict__ ccounts, double4 *__restrict__ xref)
{
    const int t = list_tiles[blockIdx.x];   // tiles of list replicas
    const int seg = blockIdx.y;
    const int r = tile_rep[t];
    if (r < 0) return;
    const RepConsts &rc = reps[r];
    if (rc.list_items <= 0 || !lists_stale(maxdisp + 2 * r, rc.skin)) return;   // uniform
    constexpr int W = BLOCK / TILE;
    __shared__ float4 stage[2][W][TILE];
    __shared__ int ctiles[CLIST_MAX_N / TILE / NSEG + 1];
    const int lane = threadIdx.x & (TILE - 1), wv = threadIdx.x / TILE;
    const int cll = (t - rc.tile0) * CPW + wv;   // cluster within the replica
    const bool cvalid = cll * CL < rc.n;
    double tc[3], th[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) { tc[k] = box[t].c[k]; th[k] = box[t].h[k]; }
    double c[3], h[3];
    {
        const int il = cll * CL + lane;
        const bool v = lane < CL && il < rc.n;
        const double4 p = v ? pos[rc.off + il] : make_double4(0, 0, 0, 0);
        if (v && seg == 0) xref[rc.off + il] = p;
        const double xi[3] = {p.x, p.y, p.z};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            double lo = v ? xi[k] : __builtin_inf(), hi = v ? xi[k] : -__builtin_inf();
            wave_minmax(lo, hi);
            c[k] = 0.5 * (lo + hi);
            h[k] = 0.5 * (hi - lo);
        }
    }
    const bool live = cvalid && h[0] >= 0.0;
    const ClusterView cv = cluster_view(c, h, live, tc, rc.Lf, rc.halfLf_m, rc.skinf, rc.rl2f,
                                        rc.rl2f, true);   // ions: every source to the one radius
    int ncand = 0;
    {
        const double rl = __builtin_sqrt(rc.rl2);
        const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
        for (int tb = 0; tb < rc.ntiles; tb += TILE * NSEG) {
            const int stl = tb + lane * NSEG + seg;   // source tile within the replica
            bool tcand = false;
            if (stl < rc.ntiles) {
                const int st = rc.tile0 + stl;
                double gap2 = 0.0;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    double d = tc[k] - box[st].c[k];
                    d = __builtin_fma(-rc.L[k], __builtin_rint(d * rc.invL[k]), d);
                    double gk = __builtin_fabs(d) - th[k] - box[st].h[k];
                    gk = gk > 0.0 ? gk : 0.0;
                    gap2 = __builtin_fma(gk, gk, gap2);
                }
                tcand = gap2 <= rl * rl;
            }
            const unsigned long long tm = __ballot(tcand);
            if (wv == 0 && tcand) ctiles[ncand + __popcll(tm & lt)] = stl;
            ncand += __popcll(tm);
        }
    }
    __syncthreads();
    int nf, ng;
    clist_scan<W>(pos + rc.off, rc.n, ctiles, ncand, cv, live, tc, rc.Lf, rc.invLf,
                  rc.rl2f, rc.rn2f, stage, clist,
                  rc.lists_off + ((size_t)(cvalid ? cll : 0) * NSEG + seg) * rc.segcap, rc.segcap,
                  lane, wv, nf, ng);
    if (cvalid && lane == 0) ccounts[(t * CPW + wv) * NSEG + seg] = make_int2(nf, ng);
}

// Gather of every per-slot array by the sorted permutation over the used
// extent [0, nslots) of the arena (component-major arrays have stride cap).
__global__ __launch_bounds__(BLOCK) void k_batch_permute(const int *__restrict__ perm, int nslots,
                                                         int cap, PermArrays a)
{
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= nslots) return;
    const size_t nn = (size_t)cap;
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

// Keys of an arena re-sort (a stable radix sort of 64-bit keys over every
// slot).  A slot of a replica in `due` gets (replica offset << 32 | Hilbert
// cell of its position), and the position is wrapped into the box (a
// particle that crossed a face joins a tile on the side it is now on, so
// tiles stay compact); every other slot, unused ones included, gets
// (slot << 32), which keeps it where it is.  So the sort reorders each due
// replica inside its own slot range (equal cells keep their slot order) and
// leaves the rest of the arena alone.
__global__ __launch_bounds__(BLOCK) void k_batch_sort_keys(double4 *__restrict__ pos,
                                                           const int *__restrict__ rep,
                                                           const RepConsts *__restrict__ reps,
                                                           const int *__restrict__ due, int cap,
                                                           unsigned long long *__restrict__ key,
                                                           int *__restrict__ val,
                                                           unsigned long long *__restrict__ maxdisp)
{
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= cap) return;
    const int r = rep[i];
    unsigned long long k = (unsigned long long)i << 32;
    if (r >= 0 && due[r]) {
        const RepConsts &rc = reps[r];
        if (i == rc.off) maxdisp[2 * r] = as_bits(__builtin_inf());   // slots move: rebuild lists
        const int g = rc.grid_bits;
        const int cmax = (1 << g) - 1;
        const double4 p = pos[i];
        double xs[3] = {p.x, p.y, p.z};
        unsigned c[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            xs[d] = xs[d] - rc.L[d] * __builtin_floor(xs[d] * rc.invL[d]);
            int ci = (int)(xs[d] * rc.invL[d] * (double)(1 << g));
            ci = ci < 0 ? 0 : (ci > cmax ? cmax : ci);
            c[d] = (unsigned)ci;
        }
        pos[i] = make_double4(xs[0], xs[1], xs[2], p.w);
        k = ((unsigned long long)rc.off << 32) | (unsigned long long)hilbert_key3(c[0], c[1], c[2], g);
    }
    key[i] = k;
    val[i] = i;
}

// Per-slot finalize: sum the tile's item partials in item order, scale,
// corrector (MODE 1) and per-block measurement partials.
template <bool MEAS, int MODE>
__global__ __launch_bounds__(BLOCK) void k_batch_finalize(
    const double *__restrict__ part, const int2 *__restrict__ tile_items,
    const double4 *__restrict__ pos, const double *__restrict__ den,
    const double *__restrict__ aii, const int *__restrict__ orig, const int *__restrict__ rep,
    const RepConsts *__restrict__ reps, const RepState *__restrict__ state, int nslots, int cap,
    double *__restrict__ acc, double *__restrict__ vel, const double *__restrict__ vold,
    double *__restrict__ blockpart)
{
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    const size_t nn = (size_t)cap;
    constexpr int NC = MEAS ? 11 : 3;
    double q[NQ];
#pragma unroll
    for (int k = 0; k < NQ; ++k) q[k] = 0.0;
    const int r = (i < nslots) ? rep[i] : -1;
    if (r >= 0 && (MODE == 0 ? state[r].fresh != 0 : state[r].steps_left > 0)) {
        const RepConsts &rc = reps[r];
        const int t = i / TILE, lane = i % TILE;
        const int2 ti = tile_items[t];
        double f[NC];
#pragma unroll
        for (int k = 0; k < NC; ++k) f[k] = 0.0;
        for (int m = 0; m < ti.y; ++m) {
            const double *p = part + (size_t)(ti.x + m) * NC * TILE + lane;
#pragma unroll
            for (int k = 0; k < NC; ++k) f[k] += p[k * TILE];
        }
        const double zi = pos[i].w;
        const bool mix = (rc.kind == MIX);
        const double fz_i = (mix && !(rc.ref_f1 && orig[i] == rc.n - 1)) ? zi : 1.0;
        const double d = den[i];
        double a[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) a[k] = (rc.cnum * (fz_i * f[k])) / d;
#pragma unroll
        for (int k = 0; k < 3; ++k) acc[k * nn + i] = a[k];
        double v[3] = {0, 0, 0};
        if constexpr (MODE == 1) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                double vp = vold[k * nn + i] + rc.dt * a[k];
                v[k] = 0.5 * (vel[k * nn + i] + vp);
                vel[k * nn + i] = v[k];
            }
        }
        if constexpr (MEAS) {
            const double zs = mix ? zi : 1.0;
            q[0] = 0.5 * zs * f[3];
            q[1] = 0.5 * zs * f[4];
#pragma unroll
            for (int k = 0; k < 6; ++k) q[2 + k] = 0.5 * zs * f[5 + k];
            if constexpr (MODE != 0) {
                const double ai = aii[i];
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

// One workgroup per replica: optional measurement record (fixed-order sum of
// the replica's finalize blocks), step countdown, and the reset of the pair
// launch's queue counters for the next step (block 0).
// MODE 0 (accel pass of freshly admitted replicas) writes their record to
// init_records[r] and clears fresh; MODE 1 appends to records[r][rec_count]
// and counts a step.
template <bool MEAS, int MODE>
__global__ __launch_bounds__(BLOCK) void k_batch_replica(const double *__restrict__ blockpart,
                                                         const RepConsts *__restrict__ reps,
                                                         RepState *__restrict__ state,
                                                         int nrep, MeasOut *__restrict__ records,
                                                         int rec_cap,
                                                         MeasOut *__restrict__ init_records,
                                                         unsigned *__restrict__ queue,
                                                         unsigned long long *__restrict__ maxdisp)
{
    const int r = blockIdx.x;
    if (r == 0 && threadIdx.x < QPARTS) queue[threadIdx.x * QSTRIDE] = 0u;
    if (r >= nrep) return;
    const RepConsts &rc = reps[r];
    // the list build of this step has consumed the displacement flag
    if (threadIdx.x == 0 && rc.list_items > 0 && lists_stale(maxdisp + 2 * r, rc.skin) &&
        (MODE == 0 ? state[r].fresh != 0 : state[r].steps_left > 0)) {
        maxdisp[2 * r] = 0ull;
        maxdisp[2 * r + 1] = 0ull;
    }
    const bool live = rc.n > 0 && (MODE == 0 ? state[r].fresh != 0 : state[r].steps_left > 0);
    if constexpr (MEAS) {
        double q[NQ];
#pragma unroll
        for (int k = 0; k < NQ; ++k) q[k] = 0.0;
        if (live)
            for (int b = threadIdx.x; b < rc.nblk; b += BLOCK) {
#pragma unroll
                for (int k = 0; k < NQ; ++k) q[k] += blockpart[(size_t)(rc.blk0 + b) * NQ + k];
            }
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
        if (threadIdx.x == 0 && live) {
            MeasOut m;
            const double ev = rc.cnum * red[0][0], vir = rc.cnum * red[1][0];
            m.ev = ev / (double)rc.n_f;
            m.vir = vir;
            m.px = rc.rho * (rc.kT + vir / (double)rc.three_n);
            for (int k = 0; k < 6; ++k) m.pp[k] = rc.cnum * red[2 + k][0];
            m.ek = 0.0;
            if (MODE != 0) {
                m.ek = (0.5 * rc.xmass) * (red[8][0] + red[11][0] + red[13][0]) / (double)rc.n_f;
                for (int k = 0; k < 6; ++k) m.pp[k] = (m.pp[k] + rc.xmass * red[8 + k][0]) / rc.vol;
            }
            if (MODE == 0) {
                init_records[r] = m;
            } else {
                const int slot = state[r].rec_count;
                if (slot < rec_cap) records[(size_t)r * rec_cap + slot] = m;
                state[r].rec_count = slot + 1;
            }
        }
    }
    __syncthreads();   // every thread has read state[r] before it changes
    if (threadIdx.x == 0) {
        if (MODE == 0) state[r].fresh = 0;
        else if (state[r].steps_left > 0) state[r].steps_left -= 1;
    }
}

}  // namespace yk
