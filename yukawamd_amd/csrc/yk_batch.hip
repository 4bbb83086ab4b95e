// yk_batch.hip -- the batched ensemble engine behind yk_batch_* (include/yk.h).
//
// Many independent ion systems ("replicas": pure or mixture, FP64 or mixed,
// each with its own N, box, screening length, cutoff, timestep and step
// budget) live in one device arena and advance together: per step one drift
// launch, one persistent pair launch pulling (replica, tile, source range or
// list share) items from a device queue, one finalize launch and one
// per-replica launch (records, step countdown, queue reset).  The step sequence is captured
// once as a hipGraph and replayed.  Replicas can be added and removed
// between runs (continuous batching); a replica's arithmetic does not depend
// on what else is in the arena.  Only the vel_verlet_2 integrator without a
// B field is offered here (the single-system API has the rest).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <numeric>
#include <string>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "../../include/yk.h"
#include "yk_batch_kernels.hpp"
#include "yk_batch_plan.hpp"

using namespace yk;
using plan::LIST_UNIT;

namespace {

thread_local std::string b_err;

int bfail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    b_err = buf;
    return code;
}

// A failed allocation is returned as YK_ERR_NOMEM, which callers may handle
// (the serving loop waits for a retirement); the runtime's last-error state
// is cleared of it, so that the next launch check does not report it again.
#define YB_HIP(call)                                                                   \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess) {                                                        \
            if (e_ == hipErrorOutOfMemory) (void)hipGetLastError();                    \
            return bfail(e_ == hipErrorOutOfMemory ? YK_ERR_NOMEM : YK_ERR_HIP,         \
                         "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, \
                         __LINE__);                                                    \
        }                                                                              \
    } while (0)

struct HostReplica {
    bool active = false;
    int n = 0, off = 0, nblocks = 0;   // slots [off, off + n), 256-slot blocks used
    yk_params p{};
    double dt = 0.0;
    // host mirror of the device step countdown (the device skips a replica
    // whose budget is used up; the host knows when, without reading back)
    int steps_left = 0, steps_done = 0, sorted_at = 0;
    // cluster lists (list_items > 0 work items per tile): region of the list
    // arena, fixed for the replica's lifetime
    int list_items = 0, segcap = 0;
    int64_t lists_off = 0, lists_len = 0;
    double near_r = -1.0;   // near-field radius of the mixed form (fm), set at admission
};

}  // namespace

struct yk_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    int cap = 0;             // slots (multiple of 256)
    int max_rep = 0, rec_cap = 0;
    int cus = 256;
    // device arena
    double4 *pos = nullptr;
    double *vel = nullptr, *vold = nullptr, *acc = nullptr, *aii = nullptr, *den = nullptr;
    int *orig = nullptr, *rep = nullptr, *tile_rep = nullptr;
    TileBox *box = nullptr;
    RepConsts *reps = nullptr;
    RepState *state = nullptr;
    MeasOut *records = nullptr, *init_records = nullptr;
    BatchItem *items = nullptr;
    int *qorder = nullptr;    // queue position -> item (list shares 0 first)
    unsigned *queue = nullptr;   // [QPARTS * QSTRIDE] queue counters of the pair launch
    int2 *tile_items = nullptr;
    double *part = nullptr;
    size_t part_items_cap = 0;
    double *blockpart = nullptr;
    // arena re-sort: keys, slot values, radix-sort scratch, gather targets
    unsigned long long *keys = nullptr, *keys2 = nullptr;
    int *vals = nullptr, *vals2 = nullptr, *due = nullptr;
    void *sort_tmp = nullptr;
    size_t sort_tmp_bytes = 0;
    int sort_end_bit = 64;
    double4 *pos2 = nullptr;
    double *vel2 = nullptr, *vold2 = nullptr, *acc2 = nullptr, *aii2 = nullptr, *den2 = nullptr;
    int *orig2 = nullptr;
    int resort_every = 10;   // a replica is re-sorted every this many of its own steps
    // host
    std::vector<HostReplica> table;
    std::vector<char> used;  // per 256-slot block
    bool dirty = true;       // items / constants must be rebuilt
    bool need_accel = true;  // accelerations of new replicas are missing
    int nitems = 0;
    plan::Tuning tune;       // work shape per N bucket
    // cluster-list arena (entries) with first-fit unit bookkeeping
    unsigned *clist = nullptr;
    int64_t clist_cap = 0;
    std::vector<char> list_used;   // per LIST_UNIT
    int2 *ccounts = nullptr;       // [cap/16][NSEG]
    double *etab_g = nullptr;      // [EXP_TN] 2^(j/N) table of the FP64 exp
    double4 *xref = nullptr;       // positions at each replica's last list build
    unsigned long long *rep_maxdisp = nullptr;   // [2 max_rep] two largest displacements^2 (bits)
    int nlist = 0;                 // active list replicas
    int span = 0;                  // used extent of the arena (slots)
    int *list_tiles = nullptr;     // tiles of list replicas (build launch)
    int n_list_tiles = 0;
    size_t list_tiles_cap = 0;
    unsigned long long md_stage[2] = {0, 0};   // host source of an async maxdisp upload
    int waves = 4;           // waves per workgroup of the persistent launch
    // captured step graph (valid while !dirty)
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    int g_steps = 0, g_nind = -1;
    // timing of the last run
    hipEvent_t e0 = nullptr, e1 = nullptr;
    double last_ms = 0.0;
    int64_t last_steps = 0;
    int64_t last_pairs = 0;
};

namespace {

void drop_graph(yk_batch *b)
{
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->gexec = nullptr;
    b->graph = nullptr;
    b->g_steps = 0;
    b->g_nind = -1;
}

// device bytes per arena slot (pos, vel/vold/acc, aii, den, orig, rep), per
// 64-slot tile (box, tile_rep, tile_items), per 256-slot block (measurement
// partials) and per work item (partial sums, item record)
constexpr int64_t SLOT_BYTES = 2 * sizeof(double4) + 11 * sizeof(double) + 2 * sizeof(int) +
                               NSEG * sizeof(int2) / CL;
constexpr int64_t TILE_BYTES = sizeof(TileBox) + sizeof(int) + sizeof(int2);
constexpr int64_t BLOCK_BYTES = NQ * sizeof(double);
constexpr int64_t ITEM_BYTES = 11 * TILE * sizeof(double) + sizeof(BatchItem);

RepConsts rep_consts(const HostReplica &h)
{
    const yk_params &p = h.p;
    RepConsts c{};
    double lmax = std::max(p.xl[0], std::max(p.xl[1], p.xl[2]));
    c.margin = 1e-12 * lmax;
    for (int k = 0; k < 3; ++k) {
        c.L[k] = p.xl[k];
        c.invL[k] = 1.0 / p.xl[k];
        c.halfL_m[k] = 0.5 * p.xl[k] - 4.0 * c.margin;
    }
    c.cull2 = p.rccut * p.rccut * (1.0 + 1e-12) + c.margin;
    // the cutoff the pair test uses (exp_cutoff2: unchanged for every pair
    // that contributes)
    c.rc2 = exp_cutoff2(p.rccut * p.rccut, p.xmuc, true);
    c.mu = p.xmuc;
    double hd = std::sqrt(p.xl[0] * p.xl[0] + p.xl[1] * p.xl[1] + p.xl[2] * p.xl[2]) / 2.0;
    c.evcut = (p.rccut <= hd) ? std::exp(-(p.xmuc * p.rccut)) / p.rccut : 0.0;
    c.cnum = (p.sim_type == YK_PURE_ION) ? p.frp * p.zi * p.zi * p.vc : p.frp * p.vc;
    c.dt = h.dt;
    c.afac = 0.5 * (h.dt * h.dt);   // src/vel_verlet_2.f:33
    c.rho = p.rho;
    c.kT = p.kT;
    c.xmass = p.xmass;
    c.vol = p.xl[0] * p.xl[1] * p.xl[2];
    c.muf = (float)p.xmuc;
    c.mul2ef = (float)(p.xmuc * 1.4426950408889634074);
    c.mul2e_lo = (float)(p.xmuc * 1.4426950408889634074 - (double)c.mul2ef);
    c.three_n = 3.0f * (float)h.n;
    c.n_f = (float)h.n;
    c.kind = (p.sim_type == YK_PURE_ION) ? PURE : MIX;
    // the mixed list items work in lengths scaled by mu log2 e: an unscreened
    // (mu = 0) list replica is computed in FP64
    c.prec = (p.precision == YK_MIXED && (p.xmuc > 0.0 || h.list_items == 0)) ? 1 : 0;
    c.off = h.off;
    c.n = h.n;
    c.tile0 = h.off / TILE;
    c.ntiles = (h.n + TILE - 1) / TILE;
    c.blk0 = h.off / BLOCK;
    c.nblk = (h.n + BLOCK - 1) / BLOCK;
    c.ref_f1 = (p.flags & YK_FLAG_REF_F1) ? 1 : 0;
    int g = 1;   // about 2 particles per Hilbert cell, as at admission
    while (g < 10 && (double)(1 << (3 * g)) < 0.5 * h.n) ++g;
    c.grid_bits = g;
    c.list_items = h.list_items;
    if (h.list_items > 0) {   // as the single-system lists: skin 5 % of the cutoff
        c.lists_off = h.lists_off;
        c.segcap = h.segcap;
        c.skin = 0.05 * p.rccut;
        const double rl = p.rccut + c.skin + c.margin;
        c.rl2 = rl * rl;
        for (int k = 0; k < 3; ++k) {
            c.Lf[k] = (float)c.L[k];
            c.invLf[k] = (float)c.invL[k];
            c.halfLf_m[k] = (float)(c.halfL_m[k] - 1e-5 * p.xl[k]);
        }
        const double lm = std::max(p.xl[0], std::max(p.xl[1], p.xl[2]));
        const double rlf = rl * (1.0 + 5e-6) + 1e-6 * lm;   // as the single-system bound
        c.rl2f = (float)(rlf * rlf);
        c.skinf = (float)c.skin;
        c.rn2f = h.near_r < 0.0 ? -1.0f : (float)(h.near_r * h.near_r);
    }
    c.fs = p.xmuc * 1.4426950408889634074;
    c.fs2 = c.fs * c.fs;
    c.rc2f_lo = (float)(c.rc2 * c.fs2 * (1.0 - 4e-6));
    c.rc2f_hi = (float)(c.rc2 * c.fs2 * (1.0 + 4e-6));
    c.r2_floor_f = 1e-20f;
    return c;
}

// First-fit region of the list arena for `len` entries; grows the arena
// (copying its contents, so existing regions keep their offsets) if needed.
int list_alloc(yk_batch *b, int64_t len, int64_t *off)
{
    const int64_t units = (len + LIST_UNIT - 1) / LIST_UNIT;
    int64_t start = plan::first_fit(b->list_used, units);
    if (start < 0) {
        const plan::Growth g = plan::arena_growth(b->list_used, units);
        unsigned *nc = nullptr;
        YB_HIP(hipMalloc(&nc, sizeof(unsigned) * (size_t)(g.new_units * LIST_UNIT)));
        if (b->clist) {
            YB_HIP(hipMemcpyAsync(nc, b->clist, sizeof(unsigned) * (size_t)b->clist_cap,
                                  hipMemcpyDeviceToDevice, b->stream));
            YB_HIP(hipStreamSynchronize(b->stream));
            YB_HIP(hipFree(b->clist));
        }
        b->clist = nc;
        b->clist_cap = g.new_units * LIST_UNIT;
        b->list_used.resize((size_t)g.new_units, 0);
        b->dirty = true;   // kernels and the captured graph take the new pointer
        start = g.start;
    }
    for (int64_t u = start; u < start + units; ++u) b->list_used[(size_t)u] = 1;
    *off = start * LIST_UNIT;
    return YK_OK;
}

void list_free(yk_batch *b, int64_t off, int64_t len)
{
    const int64_t u0 = off / LIST_UNIT, units = (len + LIST_UNIT - 1) / LIST_UNIT;
    for (int64_t u = u0; u < u0 + units; ++u) b->list_used[(size_t)u] = 0;
}

__global__ void k_fill_rep(int *rep, const int *tile_rep, const RepConsts *reps, int cap)
{
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= cap) return;
    const int r = tile_rep[i / TILE];
    rep[i] = (r >= 0 && i - reps[r].off < reps[r].n) ? r : -1;
}

// Items, tile item ranges, per-slot replica index and replica constants for
// the current membership.
// Room for nitems work items: partial sums, item records and queue order,
// allocated 25 % ahead.  A reallocation drops the captured graph (it holds
// the old pointers) and marks the layout for a rebuild; a failed one leaves
// no buffers (capacity 0) and returns YK_ERR_NOMEM.
int ensure_items(yk_batch *b, size_t nitems)
{
    if (nitems <= b->part_items_cap) return YK_OK;
    if (nitems > (size_t)(INT32_MAX / 2))   // item indices are 32-bit
        return bfail(YK_ERR_NOMEM, "%zu work items exceed the engine's limit", nitems);
    drop_graph(b);
    b->dirty = true;
    if (b->part) YB_HIP(hipFree(b->part));
    if (b->items) YB_HIP(hipFree(b->items));
    if (b->qorder) YB_HIP(hipFree(b->qorder));
    b->part = nullptr;
    b->items = nullptr;
    b->qorder = nullptr;
    b->part_items_cap = 0;
    const size_t want = plan::item_capacity((int)nitems);
    YB_HIP(hipMalloc(&b->part, sizeof(double) * 11 * TILE * want));
    YB_HIP(hipMalloc(&b->items, sizeof(BatchItem) * want));
    YB_HIP(hipMalloc(&b->qorder, sizeof(int) * want));
    b->part_items_cap = want;
    return YK_OK;
}

int rebuild(yk_batch *b)
{
    static_assert(sizeof(plan::Item) == sizeof(BatchItem), "BatchItem is plan::Item");
    static_assert(sizeof(int2) == 2 * sizeof(int), "tile_items are (first, count) pairs");
    drop_graph(b);
    const int ntiles_cap = b->cap / TILE;
    std::vector<RepConsts> rc(b->max_rep);
    std::vector<plan::Member> members(b->max_rep);
    for (int r = 0; r < b->max_rep; ++r) {
        const HostReplica &h = b->table[r];
        rc[r] = h.active ? rep_consts(h) : RepConsts{};
        members[r] = plan::Member{h.active, h.n, h.off, h.nblocks, h.list_items};
    }
    const plan::Layout L = plan::layout(b->tune, members, b->cap);
    b->nitems = (int)L.items.size();
    b->nlist = L.nlist;
    b->span = L.span;
    b->waves = L.waves;
    b->n_list_tiles = (int)L.list_tiles.size();
    if (L.list_tiles.size() > b->list_tiles_cap) {
        if (b->list_tiles) YB_HIP(hipFree(b->list_tiles));
        b->list_tiles = nullptr;
        b->list_tiles_cap = L.list_tiles.size() + L.list_tiles.size() / 4 + 64;
        YB_HIP(hipMalloc(&b->list_tiles, sizeof(int) * b->list_tiles_cap));
    }
    if (!L.list_tiles.empty())
        YB_HIP(hipMemcpyAsync(b->list_tiles, L.list_tiles.data(), sizeof(int) * L.list_tiles.size(),
                              hipMemcpyHostToDevice, b->stream));
    if (const int e = ensure_items(b, (size_t)b->nitems)) return e;
    YB_HIP(hipMemcpyAsync(b->reps, rc.data(), sizeof(RepConsts) * b->max_rep,
                          hipMemcpyHostToDevice, b->stream));
    YB_HIP(hipMemcpyAsync(b->tile_rep, L.tile_rep.data(), sizeof(int) * ntiles_cap,
                          hipMemcpyHostToDevice, b->stream));
    YB_HIP(hipMemcpyAsync(b->tile_items, L.tile_items.data(), sizeof(int2) * ntiles_cap,
                          hipMemcpyHostToDevice, b->stream));
    if (b->nitems) {
        YB_HIP(hipMemcpyAsync(b->items, L.items.data(), sizeof(BatchItem) * b->nitems,
                              hipMemcpyHostToDevice, b->stream));
        YB_HIP(hipMemcpyAsync(b->qorder, L.qorder.data(), sizeof(int) * b->nitems,
                              hipMemcpyHostToDevice, b->stream));
    }
    hipLaunchKernelGGL(k_fill_rep, dim3((b->cap + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, b->stream,
                       b->rep, b->tile_rep, b->reps, b->cap);
    YB_HIP(hipGetLastError());
    YB_HIP(hipStreamSynchronize(b->stream));
    b->dirty = false;
    return YK_OK;
}

int persistent_blocks(const yk_batch *b) { return plan::persistent_blocks(b->cus, b->waves, b->nitems); }

// One step (MODE 1) or the accel pass (MODE 0), stream-ordered, graph-safe:
// no memsets, the queue counters are reset by the per-replica kernel.
template <bool MEAS, int MODE>
int enqueue_step(yk_batch *b)
{
    const dim3 slots(std::max(1, (b->span + BLOCK - 1) / BLOCK));
    const ExpCoeffs ec = exp_coeffs();
    hipLaunchKernelGGL((k_batch_drift<MODE == 1>), slots, dim3(BLOCK), 0, b->stream, b->pos, b->vel,
                           b->vold, b->acc, b->rep, b->reps, b->state, b->span, b->cap,
                           b->box, b->tile_rep, b->xref, b->rep_maxdisp);
    if (b->n_list_tiles > 0)
        hipLaunchKernelGGL(k_batch_clist_build, dim3(b->n_list_tiles, NSEG), dim3(BLOCK), 0,
                           b->stream, b->pos, b->box, b->list_tiles, b->tile_rep, b->reps,
                           b->rep_maxdisp, b->clist, b->ccounts, b->xref);
    const size_t lds = (sizeof(double4) + sizeof(float4)) * TILE * b->waves;
    hipLaunchKernelGGL((k_batch_pairs<MEAS>), dim3(persistent_blocks(b)), dim3(TILE * b->waves),
                       lds, b->stream, b->pos, b->box, b->reps, b->state, b->items, b->qorder,
                       b->nitems, b->queue, ec, b->etab_g, MODE == 0 ? 1 : 0, b->clist,
                       b->ccounts, b->part);
    hipLaunchKernelGGL((k_batch_finalize<MEAS, MODE>), slots, dim3(BLOCK), 0, b->stream, b->part,
                       b->tile_items, b->pos, b->den, b->aii, b->orig, b->rep, b->reps,
                       b->state, b->span, b->cap, b->acc, b->vel, b->vold, b->blockpart);
    hipLaunchKernelGGL((k_batch_replica<MEAS, MODE>), dim3(b->max_rep), dim3(BLOCK), 0,
                       b->stream, b->blockpart, b->reps, b->state, b->max_rep, b->records,
                       b->rec_cap, b->init_records, b->queue, b->rep_maxdisp);
    YB_HIP(hipGetLastError());
    return YK_OK;
}

// Accelerations (and the initial measurement) of the replicas admitted since
// the last pass; the others keep theirs.
int accel_pass(yk_batch *b)
{
    // the queue counters are zero here: they start zeroed and every step's
    // per-replica kernel resets them for the next launch
    int rc = enqueue_step<true, 0>(b);
    if (rc) return rc;
    YB_HIP(hipStreamSynchronize(b->stream));
    b->need_accel = false;
    return YK_OK;
}

// Capture `steps` steps (measuring every nind-th) as one graph.
int capture(yk_batch *b, int steps, int nind)
{
    drop_graph(b);
    YB_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int rc = YK_OK;
    for (int s = 1; s <= steps && rc == YK_OK; ++s) {
        const bool meas = nind > 0 && (s % nind == 0);
        rc = meas ? enqueue_step<true, 1>(b) : enqueue_step<false, 1>(b);
    }
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return bfail(YK_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
    b->graph = g;
    YB_HIP(hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0));
    b->g_steps = steps;
    b->g_nind = nind;
    return YK_OK;
}

// Re-sort the replicas whose own step count reached a multiple of
// resort_every (and that still have steps to do) back into Hilbert order:
// keys, one stable radix sort over the arena, gather into the second buffer
// set, copy back (the captured step graph keeps its buffer addresses).  A
// replica's re-sort points depend only on its own step count, so its
// trajectory is the same whatever else is in the arena and whenever it was
// admitted.
int resort_due(yk_batch *b)
{
    if (b->resort_every <= 0) return YK_OK;
    std::vector<int> due(b->max_rep, 0);
    bool any = false;
    for (int r = 0; r < b->max_rep; ++r) {
        HostReplica &h = b->table[r];
        if (h.active && h.steps_left > 0 && h.steps_done > 0 &&
            h.steps_done % b->resort_every == 0 && h.sorted_at != h.steps_done) {
            due[r] = 1;
            h.sorted_at = h.steps_done;
            any = true;
        }
    }
    if (!any) return YK_OK;
    // the used extent of the arena only; component-major arrays have stride cap
    const size_t nn = (size_t)b->cap, ns = (size_t)b->span;
    const dim3 slots(std::max(1, (b->span + BLOCK - 1) / BLOCK));
    YB_HIP(hipMemcpyAsync(b->due, due.data(), sizeof(int) * b->max_rep, hipMemcpyHostToDevice,
                          b->stream));
    hipLaunchKernelGGL(k_batch_sort_keys, slots, dim3(BLOCK), 0, b->stream, b->pos, b->rep,
                       b->reps, b->due, b->span, b->keys, b->vals, b->rep_maxdisp);
    YB_HIP(hipGetLastError());
    size_t bytes = b->sort_tmp_bytes;
    YB_HIP(hipcub::DeviceRadixSort::SortPairs(b->sort_tmp, bytes, b->keys, b->keys2, b->vals,
                                              b->vals2, b->span, 0, b->sort_end_bit, b->stream));
    PermArrays a{b->pos, b->vel, b->vold, b->acc, b->aii, b->den, b->orig,
                 b->pos2, b->vel2, b->vold2, b->acc2, b->aii2, b->den2, b->orig2};
    hipLaunchKernelGGL(k_batch_permute, slots, dim3(BLOCK), 0, b->stream, b->vals2, b->span,
                       b->cap, a);
    YB_HIP(hipGetLastError());
    YB_HIP(hipMemcpyAsync(b->pos, b->pos2, sizeof(double4) * ns, hipMemcpyDeviceToDevice,
                          b->stream));
    for (auto pr : {std::make_pair(b->vel, b->vel2), std::make_pair(b->vold, b->vold2),
                    std::make_pair(b->acc, b->acc2)})
        for (int c = 0; c < 3; ++c)
            YB_HIP(hipMemcpyAsync(pr.first + c * nn, pr.second + c * nn, sizeof(double) * ns,
                                  hipMemcpyDeviceToDevice, b->stream));
    YB_HIP(hipMemcpyAsync(b->aii, b->aii2, sizeof(double) * ns, hipMemcpyDeviceToDevice,
                          b->stream));
    YB_HIP(hipMemcpyAsync(b->den, b->den2, sizeof(double) * ns, hipMemcpyDeviceToDevice,
                          b->stream));
    YB_HIP(hipMemcpyAsync(b->orig, b->orig2, sizeof(int) * ns, hipMemcpyDeviceToDevice,
                          b->stream));
    return YK_OK;
}

// Run-relative steps [from, to): whole graph blocks where aligned, single
// steps otherwise; step j (1-based in the run) measures when j % nind == 0.
int run_steps(yk_batch *b, int from, int to, int nind, int block)
{
    int j = from;
    while (j < to) {
        if (j % block == 0 && j + block <= to) {
            YB_HIP(hipGraphLaunch(b->gexec, b->stream));
            j += block;
        } else {
            const bool meas = nind > 0 && (j + 1) % nind == 0;
            const int rc = meas ? enqueue_step<true, 1>(b) : enqueue_step<false, 1>(b);
            if (rc) return rc;
            j += 1;
        }
    }
    for (HostReplica &h : b->table)
        if (h.active) {
            const int k = std::min(to - from, h.steps_left);
            h.steps_left -= k;
            h.steps_done += k;
        }
    return YK_OK;
}

// Box wrap of a downloaded position: the reference's one-shift rule
// (src/vel_verlet_2.f:39-40), then a floor for positions several boxes out.
double wrap_into_box(double x, double L)
{
    if (x < 0.0) x = x + L;
    if (x > L) x = x - L;
    if (x < 0.0 || x > L) x = x - L * std::floor(x / L);
    return x;
}

// Hilbert order of one replica's particles (host; done once per admission).
std::vector<int> sfc_order(const double *x, int n, const double *L)
{
    int g = 1;
    while (g < 10 && (double)(1 << (3 * g)) < 0.5 * n) ++g;
    const double G = (double)(1 << g);
    std::vector<unsigned> key(n);
    for (int i = 0; i < n; ++i) {
        unsigned c[3];
        for (int k = 0; k < 3; ++k) {
            double xk = x[3 * (size_t)i + k];
            xk -= L[k] * std::floor(xk / L[k]);
            int ci = (int)(xk * G / L[k]);
            ci = ci < 0 ? 0 : (ci > (1 << g) - 1 ? (1 << g) - 1 : ci);
            c[k] = (unsigned)ci;
        }
        key[i] = hilbert_key3(c[0], c[1], c[2], g);
    }
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return key[a] < key[c]; });
    return order;
}

int check_replica_params(const yk_params *p, int n)
{
    if (!p) return bfail(YK_ERR_ARG, "params is NULL");
    if (p->sim_type != YK_PURE_ION && p->sim_type != YK_ION_MIXTURE)
        return bfail(YK_ERR_UNSUPPORTED, "batched replicas are ion systems (pure or mixture)");
    if (p->precision != YK_FP64 && p->precision != YK_MIXED)
        return bfail(YK_ERR_ARG, "precision=%d", p->precision);
    if (p->bfield != 0.0 || (p->flags & (YK_FLAG_KDK | YK_FLAG_SW)))
        return bfail(YK_ERR_UNSUPPORTED, "batched replicas use vel_verlet_2 without a B field");
    for (int k = 0; k < 3; ++k)
        if (!(p->xl[k] > 0.0)) return bfail(YK_ERR_ARG, "xl(%d)=%g", k + 1, p->xl[k]);
    if (!(p->rccut > 0.0) || !(p->xmass > 0.0) || p->xmuc < 0.0)
        return bfail(YK_ERR_ARG, "rccut, xmass must be > 0 and xmuc >= 0");
    if (p->sim_type == YK_PURE_ION && !(p->ai > 0.0))
        return bfail(YK_ERR_ARG, "pure-ion needs ai > 0");
    if (n < 1) return bfail(YK_ERR_ARG, "n=%d must be >= 1", n);
    return YK_OK;
}

}  // namespace

// ===========================================================================
extern "C" {

const char *yk_batch_last_error(void) { return b_err.c_str(); }

int yk_batch_create(int64_t capacity, int32_t max_replicas, int32_t max_records,
                    yk_batch **out)
{
    if (!out) return bfail(YK_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (capacity < 1 || capacity > ((int64_t)1 << 28))
        return bfail(YK_ERR_ARG, "capacity=%lld slots out of range", (long long)capacity);
    if (max_replicas < 1 || max_records < 0)
        return bfail(YK_ERR_ARG, "max_replicas must be >= 1 and max_records >= 0");
    yk_batch *b = new (std::nothrow) yk_batch();
    if (!b) return bfail(YK_ERR_NOMEM, "host allocation failed");
    b->cap = (int)((capacity + BLOCK - 1) / BLOCK * BLOCK);
    b->max_rep = max_replicas;
    b->rec_cap = std::max(1, (int)max_records);
    b->table.resize(max_replicas);
    b->used.assign(b->cap / BLOCK, 0);
    auto bail = [&](int code) {
        yk_batch_destroy(b);
        return code;
    };
#define YB_HIPB(call)                                                                  \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess) {                                                        \
            if (e_ == hipErrorOutOfMemory) (void)hipGetLastError();                    \
            return bail(bfail(e_ == hipErrorOutOfMemory ? YK_ERR_NOMEM : YK_ERR_HIP,    \
                              "%s failed: %s", #call, hipGetErrorString(e_)));         \
        }                                                                              \
    } while (0)
    YB_HIPB(hipGetDevice(&b->device));
    hipDeviceProp_t prop;
    YB_HIPB(hipGetDeviceProperties(&prop, b->device));
    b->cus = prop.multiProcessorCount;
    YB_HIPB(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    const size_t nn = (size_t)b->cap;
    YB_HIPB(hipMalloc(&b->pos, sizeof(double4) * nn));
    for (double **pp : {&b->vel, &b->vold, &b->acc})
        YB_HIPB(hipMalloc(pp, sizeof(double) * 3 * nn));
    YB_HIPB(hipMalloc(&b->aii, sizeof(double) * nn));
    YB_HIPB(hipMalloc(&b->den, sizeof(double) * nn));
    YB_HIPB(hipMalloc(&b->orig, sizeof(int) * nn));
    YB_HIPB(hipMalloc(&b->rep, sizeof(int) * nn));
    YB_HIPB(hipMalloc(&b->tile_rep, sizeof(int) * (nn / TILE)));
    YB_HIPB(hipMalloc(&b->box, sizeof(TileBox) * (nn / TILE)));
    YB_HIPB(hipMalloc(&b->tile_items, sizeof(int2) * (nn / TILE)));
    YB_HIPB(hipMalloc(&b->blockpart, sizeof(double) * NQ * (nn / BLOCK)));
    YB_HIPB(hipMalloc(&b->reps, sizeof(RepConsts) * (size_t)max_replicas));
    YB_HIPB(hipMalloc(&b->state, sizeof(RepState) * (size_t)max_replicas));
    YB_HIPB(hipMalloc(&b->records, sizeof(MeasOut) * (size_t)max_replicas * b->rec_cap));
    YB_HIPB(hipMalloc(&b->init_records, sizeof(MeasOut) * (size_t)max_replicas));
    YB_HIPB(hipMalloc(&b->queue, sizeof(unsigned) * QPARTS * QSTRIDE));
    YB_HIPB(hipMalloc(&b->xref, sizeof(double4) * nn));
    YB_HIPB(hipMalloc(&b->ccounts, sizeof(int2) * (nn / CL) * NSEG));
    {
        std::vector<double> t(EXP_TN);
        exp_table(t.data());
        YB_HIPB(hipMalloc(&b->etab_g, sizeof(double) * EXP_TN));
        YB_HIPB(hipMemcpy(b->etab_g, t.data(), sizeof(double) * EXP_TN, hipMemcpyHostToDevice));
    }
    YB_HIPB(hipMalloc(&b->rep_maxdisp, 2 * sizeof(unsigned long long) * (size_t)max_replicas));
    YB_HIPB(hipMalloc(&b->keys, sizeof(unsigned long long) * nn));
    YB_HIPB(hipMalloc(&b->keys2, sizeof(unsigned long long) * nn));
    YB_HIPB(hipMalloc(&b->vals, sizeof(int) * nn));
    YB_HIPB(hipMalloc(&b->vals2, sizeof(int) * nn));
    YB_HIPB(hipMalloc(&b->due, sizeof(int) * (size_t)max_replicas));
    YB_HIPB(hipMalloc(&b->pos2, sizeof(double4) * nn));
    for (double **pp : {&b->vel2, &b->vold2, &b->acc2})
        YB_HIPB(hipMalloc(pp, sizeof(double) * 3 * nn));
    YB_HIPB(hipMalloc(&b->aii2, sizeof(double) * nn));
    YB_HIPB(hipMalloc(&b->den2, sizeof(double) * nn));
    YB_HIPB(hipMalloc(&b->orig2, sizeof(int) * nn));
    b->sort_end_bit = 33;
    while (b->sort_end_bit < 64 && ((unsigned long long)1 << (b->sort_end_bit - 32)) < nn)
        ++b->sort_end_bit;
    YB_HIPB(hipcub::DeviceRadixSort::SortPairs(nullptr, b->sort_tmp_bytes, b->keys, b->keys2,
                                               b->vals, b->vals2, b->cap, 0, b->sort_end_bit,
                                               b->stream));
    YB_HIPB(hipMalloc(&b->sort_tmp, b->sort_tmp_bytes > 0 ? b->sort_tmp_bytes : 16));
    YB_HIPB(hipEventCreate(&b->e0));
    YB_HIPB(hipEventCreate(&b->e1));
    // empty arena: every slot unused, every replica idle
    YB_HIPB(hipMemsetAsync(b->pos, 0, sizeof(double4) * nn, b->stream));
    YB_HIPB(hipMemsetAsync(b->vel, 0, sizeof(double) * 3 * nn, b->stream));
    YB_HIPB(hipMemsetAsync(b->vold, 0, sizeof(double) * 3 * nn, b->stream));
    YB_HIPB(hipMemsetAsync(b->acc, 0, sizeof(double) * 3 * nn, b->stream));
    YB_HIPB(hipMemsetAsync(b->state, 0, sizeof(RepState) * (size_t)max_replicas, b->stream));
    YB_HIPB(hipMemsetAsync(b->queue, 0, sizeof(unsigned) * QPARTS * QSTRIDE, b->stream));
    std::vector<double> one(nn, 1.0);
    YB_HIPB(hipMemcpyAsync(b->aii, one.data(), sizeof(double) * nn, hipMemcpyHostToDevice,
                           b->stream));
    YB_HIPB(hipMemcpyAsync(b->den, one.data(), sizeof(double) * nn, hipMemcpyHostToDevice,
                           b->stream));
    YB_HIPB(hipStreamSynchronize(b->stream));
    int rc = rebuild(b);
    if (rc) return bail(rc);
#undef YB_HIPB
    *out = b;
    return YK_OK;
}

void yk_batch_destroy(yk_batch *b)
{
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graph(b);
    for (void *p : {(void *)b->pos, (void *)b->vel, (void *)b->vold, (void *)b->acc,
                    (void *)b->aii, (void *)b->den, (void *)b->orig, (void *)b->rep,
                    (void *)b->tile_rep, (void *)b->box, (void *)b->tile_items,
                    (void *)b->blockpart, (void *)b->reps, (void *)b->state,
                    (void *)b->records, (void *)b->init_records, (void *)b->queue,
                    (void *)b->part, (void *)b->items, (void *)b->qorder, (void *)b->keys, (void *)b->keys2,
                    (void *)b->vals, (void *)b->vals2, (void *)b->due, b->sort_tmp,
                    (void *)b->pos2, (void *)b->vel2, (void *)b->vold2, (void *)b->acc2,
                    (void *)b->aii2, (void *)b->den2, (void *)b->orig2, (void *)b->xref,
                    (void *)b->ccounts, (void *)b->rep_maxdisp, (void *)b->clist,
                    (void *)b->list_tiles, (void *)b->etab_g})
        if (p) (void)hipFree(p);
    if (b->e0) (void)hipEventDestroy(b->e0);
    if (b->e1) (void)hipEventDestroy(b->e1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

int yk_batch_footprint(const yk_batch *b, int32_t n, int64_t *slots_out, int64_t *bytes_out)
{
    if (!b || n < 1) return bfail(YK_ERR_ARG, "batch is NULL or n < 1");
    const int64_t slots = ((int64_t)n + BLOCK - 1) / BLOCK * BLOCK;
    if (slots_out) *slots_out = slots;
    if (bytes_out)
        *bytes_out = slots * SLOT_BYTES + slots / TILE * TILE_BYTES + slots / BLOCK * BLOCK_BYTES +
                     plan::items_of(b->tune, n) * ITEM_BYTES +
                     (b->tune.list_items_for(n) > 0
                          ? (plan::list_entries_of(n) + LIST_UNIT - 1) / LIST_UNIT * LIST_UNIT *
                                (int64_t)sizeof(unsigned)
                          : 0);
    return YK_OK;
}

int yk_batch_add(yk_batch *b, const yk_params *params, int32_t n, const double *zii,
                 const double *aii, const double *x, const double *v, double dt,
                 int32_t steps, int32_t *id_out)
{
    if (!b || !id_out || !x) return bfail(YK_ERR_ARG, "batch, x or id_out is NULL");
    int rc = check_replica_params(params, n);
    if (rc) return rc;
    if (params->sim_type == YK_ION_MIXTURE && (!zii || !aii))
        return bfail(YK_ERR_ARG, "zii and aii are required for a mixture");
    if (!std::isfinite(dt) || steps < 0) return bfail(YK_ERR_ARG, "dt or steps invalid");
    int id = -1;
    for (int r = 0; r < b->max_rep; ++r)
        if (!b->table[r].active) {
            id = r;
            break;
        }
    if (id < 0) return bfail(YK_ERR_CAPACITY, "all %d replica slots are in use", b->max_rep);
    const int nblk = (n + BLOCK - 1) / BLOCK;
    const int start = (int)plan::first_fit(b->used, nblk);
    if (start < 0)
        return bfail(YK_ERR_CAPACITY, "no room for %d particles in the arena", n);
    for (size_t k = 0; k < (size_t)3 * n; ++k)
        if (!std::isfinite(x[k]) || (v && !std::isfinite(v[k])))
            return bfail(YK_ERR_NONFINITE, "positions or velocities contain NaN/Inf");
    // charges and masses, checked before anything is allocated for the replica
    auto z_of = [&](int i) {
        return (params->sim_type == YK_PURE_ION) ? (zii ? zii[i] : params->zi) : zii[i];
    };
    auto a_of = [&](int i) {
        return (params->sim_type == YK_PURE_ION) ? (aii ? aii[i] : params->ai) : aii[i];
    };
    for (int i = 0; i < n; ++i)
        if (!(a_of(i) > 0.0) || !std::isfinite(z_of(i)))
            return bfail(YK_ERR_ARG, "particle %d: A=%g Z=%g", i, a_of(i), z_of(i));
    HostReplica &h = b->table[id];
    h = HostReplica{};
    h.p = *params;
    h.n = n;
    h.off = start * BLOCK;
    h.nblocks = nblk;
    h.dt = dt;
    h.steps_left = steps;
    h.list_items = b->tune.list_items_for(n);
    if (h.list_items > 0) {   // skin 5 % of the cutoff, as rep_consts
        h.near_r = near_radius(1, params->precision == YK_MIXED, params->rho, params->xmuc,
                               params->rccut, 0.05 * params->rccut);
        h.segcap = plan::segcap_of(n);
        h.lists_len = plan::list_entries_of(n);
        if ((rc = list_alloc(b, h.lists_len, &h.lists_off))) {
            h = HostReplica{};
            return rc;
        }
    }
    {   // the membership's work items with this replica: device memory for
        // them is taken now, so that a shortage is reported here, not by a run
        int64_t total = plan::member_items(b->tune, plan::Member{true, n, h.off, nblk,
                                                                 h.list_items});
        for (int r = 0; r < b->max_rep; ++r) {
            const HostReplica &o = b->table[r];
            if (o.active)
                total += plan::member_items(b->tune, plan::Member{true, o.n, o.off, o.nblocks,
                                                                  o.list_items});
        }
        if ((rc = ensure_items(b, (size_t)total))) {
            if (h.list_items > 0) list_free(b, h.lists_off, h.lists_len);
            h = HostReplica{};
            return rc;
        }
    }
    const std::vector<int> order = sfc_order(x, n, params->xl);
    // sorted host staging
    const size_t nb = (size_t)nblk * BLOCK;
    std::vector<double4> p4(nb, make_double4(0, 0, 0, 0));
    std::vector<double> vv(3 * nb, 0.0), aa(nb, 1.0), dd(nb, 1.0);
    std::vector<int> oo(nb, -1);
    for (int k = 0; k < n; ++k) {
        const int i = order[k];
        const double zi = z_of(i), ai = a_of(i);
        double xs[3];
        for (int c = 0; c < 3; ++c) {
            double xc = x[3 * (size_t)i + c];
            xs[c] = xc - params->xl[c] * std::floor(xc / params->xl[c]);
        }
        p4[k] = make_double4(xs[0], xs[1], xs[2], zi);
        for (int c = 0; c < 3; ++c) vv[c * nb + k] = v ? v[3 * (size_t)i + c] : 0.0;
        aa[k] = ai;
        dd[k] = (params->sim_type == YK_PURE_ION) ? params->ai * params->xmass : ai * params->xmass;
        oo[k] = i;
    }
    const size_t o = (size_t)h.off;
    const size_t cap = (size_t)b->cap;
    YB_HIP(hipMemcpyAsync(b->pos + o, p4.data(), sizeof(double4) * nb, hipMemcpyHostToDevice,
                          b->stream));
    for (int c = 0; c < 3; ++c)
        YB_HIP(hipMemcpyAsync(b->vel + c * cap + o, vv.data() + c * nb, sizeof(double) * nb,
                              hipMemcpyHostToDevice, b->stream));
    YB_HIP(hipMemcpyAsync(b->aii + o, aa.data(), sizeof(double) * nb, hipMemcpyHostToDevice,
                          b->stream));
    YB_HIP(hipMemcpyAsync(b->den + o, dd.data(), sizeof(double) * nb, hipMemcpyHostToDevice,
                          b->stream));
    YB_HIP(hipMemcpyAsync(b->orig + o, oo.data(), sizeof(int) * nb, hipMemcpyHostToDevice,
                          b->stream));
    RepState st{steps, 0, 1, 0};
    YB_HIP(hipMemcpyAsync(b->state + id, &st, sizeof st, hipMemcpyHostToDevice, b->stream));
    const double inf = __builtin_inf();   // lists to be built at the accel pass
    unsigned long long infbits;
    std::memcpy(&infbits, &inf, sizeof infbits);
    b->md_stage[0] = infbits;
    b->md_stage[1] = 0;
    YB_HIP(hipMemcpyAsync(b->rep_maxdisp + 2 * id, b->md_stage, sizeof b->md_stage,
                          hipMemcpyHostToDevice, b->stream));
    YB_HIP(hipStreamSynchronize(b->stream));
    for (int s = start; s < start + nblk; ++s) b->used[s] = 1;
    h.active = true;
    b->dirty = true;
    b->need_accel = true;
    *id_out = id;
    return YK_OK;
}

int yk_batch_remove(yk_batch *b, int32_t id)
{
    if (!b || id < 0 || id >= b->max_rep || !b->table[id].active)
        return bfail(YK_ERR_ARG, "no active replica %d", id);
    HostReplica &h = b->table[id];
    YB_HIP(hipStreamSynchronize(b->stream));
    RepState st{0, 0, 0, 0};
    YB_HIP(hipMemcpy(b->state + id, &st, sizeof st, hipMemcpyHostToDevice));
    for (int s = h.off / BLOCK; s < h.off / BLOCK + h.nblocks; ++s) b->used[s] = 0;
    if (h.list_items > 0) list_free(b, h.lists_off, h.lists_len);
    h = HostReplica{};
    b->dirty = true;
    return YK_OK;
}

int yk_batch_set_tuning(yk_batch *b, int32_t nbuckets, const int32_t *n_upper,
                        const int32_t *chunk_tiles, const int32_t *waves_per_block,
                        const int32_t *list_items)
{
    if (!b || nbuckets < 0 ||
        (nbuckets > 0 && (!n_upper || !chunk_tiles || !waves_per_block || !list_items)))
        return bfail(YK_ERR_ARG, "invalid tuning table");
    for (int k = 0; k < nbuckets; ++k) {
        const int w = waves_per_block[k];
        if (chunk_tiles[k] < 1) return bfail(YK_ERR_ARG, "chunk_tiles must be >= 1");
        if (list_items[k] < 0 || list_items[k] > 1024)
            return bfail(YK_ERR_ARG, "list_items must be in [0, 1024]");
        if (w != 1 && w != 2 && w != 4 && w != 8)
            return bfail(YK_ERR_ARG, "waves_per_block must be 1, 2, 4 or 8");
        if (k > 0 && n_upper[k] <= n_upper[k - 1])
            return bfail(YK_ERR_ARG, "n_upper must increase");
    }
    b->tune.n_upper.assign(n_upper, n_upper + nbuckets);
    b->tune.chunk_tiles.assign(chunk_tiles, chunk_tiles + nbuckets);
    b->tune.waves.assign(waves_per_block, waves_per_block + nbuckets);
    b->tune.list_items.assign(list_items, list_items + nbuckets);
    b->dirty = true;
    return YK_OK;
}

int yk_batch_set_resort(yk_batch *b, int32_t every)
{
    if (!b || every < 0) return bfail(YK_ERR_ARG, "resort interval must be >= 0");
    b->resort_every = every;
    return YK_OK;
}

int yk_batch_run(yk_batch *b, int32_t nsteps, int32_t nind)
{
    if (!b || nsteps < 0) return bfail(YK_ERR_ARG, "invalid arguments");
    int rc;
    if (b->dirty && (rc = rebuild(b))) return rc;
    if (b->need_accel && (rc = accel_pass(b))) return rc;
    b->last_steps = 0;
    b->last_ms = 0.0;
    if (nsteps == 0 || b->nitems == 0) return YK_OK;
    // one graph holds a block of steps: nind steps when measuring, else up to 16
    const int block = (nind > 0) ? nind : std::min(nsteps, 16);
    if (!b->gexec || b->g_steps != block || b->g_nind != nind)
        if ((rc = capture(b, block, nind))) return rc;
    YB_HIP(hipEventRecord(b->e0, b->stream));
    // segments end where some replica's own step count reaches a re-sort point
    int done = 0;
    for (;;) {
        if ((rc = resort_due(b))) return rc;
        if (done >= nsteps) break;
        int seg = nsteps - done;
        if (b->resort_every > 0)
            for (const HostReplica &h : b->table)
                if (h.active && h.steps_left > 0) {
                    const int d = b->resort_every - h.steps_done % b->resort_every;
                    if (d <= h.steps_left) seg = std::min(seg, d);
                }
        if ((rc = run_steps(b, done, done + seg, nind, block))) return rc;
        done += seg;
    }
    YB_HIP(hipEventRecord(b->e1, b->stream));
    YB_HIP(hipEventSynchronize(b->e1));
    float ms = 0.f;
    YB_HIP(hipEventElapsedTime(&ms, b->e0, b->e1));
    b->last_ms = ms;
    b->last_steps = nsteps;
    return YK_OK;
}

int yk_batch_status(yk_batch *b, int32_t id, int32_t *steps_left, int32_t *nrec)
{
    if (!b || id < 0 || id >= b->max_rep || !b->table[id].active)
        return bfail(YK_ERR_ARG, "no active replica %d", id);
    RepState st;
    YB_HIP(hipStreamSynchronize(b->stream));
    YB_HIP(hipMemcpy(&st, b->state + id, sizeof st, hipMemcpyDeviceToHost));
    if (steps_left) *steps_left = st.steps_left;
    if (nrec) *nrec = std::min(st.rec_count, b->rec_cap);
    return YK_OK;
}

int yk_batch_get(yk_batch *b, int32_t id, double *x, double *v, double *a, yk_meas *init,
                 yk_meas *records, int32_t rec_cap, int32_t *nrec)
{
    if (!b || id < 0 || id >= b->max_rep || !b->table[id].active)
        return bfail(YK_ERR_ARG, "no active replica %d", id);
    const HostReplica &h = b->table[id];
    YB_HIP(hipStreamSynchronize(b->stream));
    const size_t o = (size_t)h.off, n = (size_t)h.n, cap = (size_t)b->cap;
    std::vector<int> orig(n);   // slot -> caller index (re-sorts move slots)
    YB_HIP(hipMemcpy(orig.data(), b->orig + o, sizeof(int) * n, hipMemcpyDeviceToHost));
    if (x) {
        std::vector<double4> p4(n);
        YB_HIP(hipMemcpy(p4.data(), b->pos + o, sizeof(double4) * n, hipMemcpyDeviceToHost));
        for (size_t k = 0; k < n; ++k) {
            double *xi = x + 3 * (size_t)orig[k];
            xi[0] = wrap_into_box(p4[k].x, h.p.xl[0]);
            xi[1] = wrap_into_box(p4[k].y, h.p.xl[1]);
            xi[2] = wrap_into_box(p4[k].z, h.p.xl[2]);
        }
    }
    std::vector<double> t(n);
    for (int which = 0; which < 2; ++which) {
        double *dst = which == 0 ? v : a;
        const double *src = which == 0 ? b->vel : b->acc;
        if (!dst) continue;
        for (int c = 0; c < 3; ++c) {
            YB_HIP(hipMemcpy(t.data(), src + c * cap + o, sizeof(double) * n,
                             hipMemcpyDeviceToHost));
            for (size_t k = 0; k < n; ++k) dst[3 * (size_t)orig[k] + c] = t[k];
        }
    }
    auto conv = [](const MeasOut &m, yk_meas *d) {
        d->ev = m.ev;
        d->ek = m.ek;
        d->px = m.px;
        d->vir = m.vir;
        for (int k = 0; k < 6; ++k) d->pp[k] = m.pp[k];
    };
    if (init) {
        MeasOut m;
        YB_HIP(hipMemcpy(&m, b->init_records + id, sizeof m, hipMemcpyDeviceToHost));
        conv(m, init);
    }
    RepState st;
    YB_HIP(hipMemcpy(&st, b->state + id, sizeof st, hipMemcpyDeviceToHost));
    const int have = std::min(st.rec_count, b->rec_cap);
    if (nrec) *nrec = have;
    if (records && rec_cap > 0 && have > 0) {
        const int k = std::min(have, (int)rec_cap);
        std::vector<MeasOut> m(k);
        YB_HIP(hipMemcpy(m.data(), b->records + (size_t)id * b->rec_cap,
                         sizeof(MeasOut) * k, hipMemcpyDeviceToHost));
        for (int q = 0; q < k; ++q) conv(m[q], &records[q]);
    }
    return YK_OK;
}

int yk_batch_last_run(yk_batch *b, double *ms, int64_t *steps, int64_t *items)
{
    if (!b) return bfail(YK_ERR_ARG, "batch is NULL");
    if (ms) *ms = b->last_ms;
    if (steps) *steps = b->last_steps;
    if (items) *items = b->nitems;
    return YK_OK;
}

}  // extern "C"
