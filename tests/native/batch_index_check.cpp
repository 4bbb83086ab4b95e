// batch_index_check.cpp -- CPU check of the batched engine's device indexing.
//
// Plays an engine's life on the host (create, set_tuning, add/remove
// replicas, run: rebuild, the accel pass, steps, re-sorts) with the engine's
// own planner (yukawamd_amd/csrc/yk_batch_plan.hpp) and its own list walker
// (SegList, yk_layout.hpp), and follows every kernel of the accel pass and of
// a step thread by thread, checking each global-memory index against the
// size of the buffer the engine allocates for it, list entries against the
// replica's own slots and list region, partial sums against the item that
// writes them, and the LDS extents against the launch's dynamic and static
// LDS.  No GPU is used; the arithmetic of the kernels (forces, positions) is
// not evaluated, only the indices they form.
//
// The cluster-list build is restated in FP64 (the device test is FP32 with
// widened bounds): its counts differ from the device's by a few borderline
// sources, which moves no index outside the bounds checked here (a segment
// holds every slot of the source tiles it covers).  The build keeps only the
// counts per cluster segment (entries are slots of the replica's own tiles,
// checked as they are written), and every list read of the pair kernel is
// checked to land on an entry the build wrote: offset within the cluster's
// region, fast entries below the segment's fast count, back entries within
// its back count from the segment's end.
//
// Built as a shared library by the CPU tests (g++, no HIP):
//   g++ -O2 -std=c++17 -shared -fPIC -I yukawamd_amd/csrc batch_index_check.cpp
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "yk_batch_plan.hpp"
#include "yk_layout.hpp"

using namespace yk;

extern "C" {

// One replica: its size, kind and geometry, positions (3, n) AoS.
struct ykc_replica {
    int32_t n;
    int32_t steps;          // step budget
    double L[3];
    double rccut;           // cutoff (fm)
    double near_r;          // near-field radius of the mixed form, < 0: none
    const double *x;
};

struct ykc_engine {
    int64_t capacity;
    int32_t max_replicas, max_records, cus, resort_every;
    int32_t nbuckets;
    const int32_t *n_upper, *chunk_tiles, *waves, *list_items;
    // a mistake planted in this restatement, to show that the checks catch
    // it: 0 none; 1 the partial-sum buffer one item short; 2 a list share one
    // entry longer; 3 a segment count one above the segment; 4 a tile item's
    // source range one tile longer; 5 a list region one cluster off; 6 an
    // admitted replica's lists not marked for building
    int32_t plant;
};

// ops: {ADD, k} admits replica k; {REMOVE, k} removes the replica admitted as
// k; {RUN, s} runs s steps (measuring every nind-th, nind = 10).
enum { YKC_ADD = 1, YKC_REMOVE = 2, YKC_RUN = 3 };

struct ykc_report {
    int64_t accesses;        // indices checked
    int64_t violations;      // indices outside their bounds
    int64_t launches;        // kernels followed
    int64_t items_max;       // largest item count of a rebuild
    int64_t list_entries_read;
    int64_t clist_cap;       // list arena (entries) at the end
    int64_t part_items_cap;
    int32_t span, waves, blocks, status;   // status: 0 ok, else why the life stopped
    int32_t refused;         // admissions refused for room (the caller retries them later)
    char first[512];         // the first violation
};

// exhaustive != 0: every list entry a pair kernel reads is checked; else the
// two ends of every linear piece of the list walk (SegList::offset is linear
// in v inside a segment's back or fast part).
int ykc_check(const ykc_engine *cfg, const ykc_replica *reps, int32_t nrep, const int32_t *ops,
              int32_t nops, int32_t exhaustive, ykc_report *out);

// Forgets the list geometry kept across calls (keyed by the position arrays'
// addresses: call it before passing arrays that may reuse freed addresses).
void ykc_clear_cache(void);
}

namespace {

constexpr int NCM = 11;   // partial-sum components of a measuring launch
constexpr int EXP_TN = 1024;
constexpr size_t LDS_LIMIT = 160 * 1024;

struct Checker {
    ykc_report *r;
    std::string where;
    void fail(const char *buf, int64_t idx, int64_t size)
    {
        if (r->violations++ == 0)
            snprintf(r->first, sizeof r->first, "%s: %s[%lld] outside [0, %lld)", where.c_str(),
                     buf, (long long)idx, (long long)size);
    }
    // one access of buf[idx] with buf holding size elements
    bool at(const char *buf, int64_t idx, int64_t size)
    {
        ++r->accesses;
        if (idx < 0 || idx >= size) {
            fail(buf, idx, size);
            return false;
        }
        return true;
    }
    // an access that must fall inside [lo, hi) (a sub-range of a buffer)
    bool in(const char *what, int64_t idx, int64_t lo, int64_t hi)
    {
        ++r->accesses;
        if (idx < lo || idx >= hi) {
            if (r->violations++ == 0)
                snprintf(r->first, sizeof r->first, "%s: %s %lld outside [%lld, %lld)",
                         where.c_str(), what, (long long)idx, (long long)lo, (long long)hi);
            return false;
        }
        return true;
    }
};

struct Rep {                // the engine's HostReplica and the RepConsts fields used for indexing
    bool active = false;
    int src = -1;           // index into the caller's replicas
    int n = 0, off = 0, nblocks = 0, list_items = 0, segcap = 0;
    int64_t lists_off = 0, lists_len = 0;
    int steps_left = 0, steps_done = 0, sorted_at = 0;
    bool fresh = false;
    bool stale = false;     // lists to be rebuilt (device maxdisp)
    int rec_count = 0;
    int64_t serial = 0;     // admission number (what a list build writes is stamped with it)
};

struct Count2 {
    int x, y;
};

// Morton order of a replica's particles (the engine's admission sorts them
// into Hilbert order; any compact order gives the same index bounds).
std::vector<int> compact_order(const ykc_replica &w)
{
    int g = 1;
    while (g < 10 && (double)(1 << (3 * g)) < 0.5 * w.n) ++g;
    std::vector<uint64_t> key((size_t)w.n);
    for (int i = 0; i < w.n; ++i) {
        uint64_t k = 0;
        unsigned c[3];
        for (int d = 0; d < 3; ++d) {
            double x = w.x[3 * (size_t)i + d];
            x -= w.L[d] * std::floor(x / w.L[d]);
            int ci = (int)(x / w.L[d] * (1 << g));
            c[d] = (unsigned)std::min(std::max(ci, 0), (1 << g) - 1);
        }
        for (int b = g - 1; b >= 0; --b)
            for (int d = 0; d < 3; ++d) k = (k << 1) | ((c[d] >> b) & 1u);
        key[(size_t)i] = k;
    }
    std::vector<int> o((size_t)w.n);
    for (int i = 0; i < w.n; ++i) o[(size_t)i] = i;
    std::stable_sort(o.begin(), o.end(), [&](int a, int b) { return key[(size_t)a] < key[(size_t)b]; });
    return o;
}

struct Engine {
    // sizes of the engine's buffers (elements)
    int cap = 0, max_rep = 0, rec_cap = 0, cus = 256, resort_every = 10;
    plan::Tuning tune;
    std::vector<Rep> table;
    std::vector<char> used, list_used;
    int64_t clist_cap = 0;
    size_t part_items_cap = 0, list_tiles_cap = 0;
    plan::Layout L;
    bool dirty = true, need_accel = true;
    // device contents that decide indices
    std::vector<int> rep;                 // [cap]
    std::vector<Count2> ccounts;          // [cap / CL * NSEG]
    // written-before-read stamps: the admission serial of the replica whose
    // build wrote a ccounts entry; the launch that wrote an item's partial
    // sums and a tile's box
    std::vector<int64_t> ccounts_by, part_by, box_by;
    int64_t serials = 0, launch = 0;
    ykc_report *rpt = nullptr;
    const ykc_replica *reps = nullptr;
    bool exhaustive = false;
    int plant = 0;
};

// ---------------------------------------------------------------------------
// The engine's buffers and their element counts (yk_batch_create and the
// growth paths of yk_batch_add and rebuild).
struct Sizes {
    int64_t pos, comp3, slot, tiles, tile_items2, blockpart, reps, state, records, init_records,
        xref, ccounts, maxdisp, keys, due, part, items, qorder, list_tiles, clist, etab;
};

Sizes sizes_of(const Engine &e)
{
    Sizes s{};
    const int64_t nn = e.cap;
    s.pos = nn;
    s.comp3 = 3 * nn;
    s.slot = nn;
    s.tiles = nn / TILE;
    s.tile_items2 = 2 * (nn / TILE);
    s.blockpart = NQ * (nn / BLOCK);
    s.reps = e.max_rep;
    s.state = e.max_rep;
    s.records = (int64_t)e.max_rep * e.rec_cap;
    s.init_records = e.max_rep;
    s.xref = nn;
    s.ccounts = (nn / CL) * NSEG;
    s.maxdisp = 2 * (int64_t)e.max_rep;
    s.keys = nn;
    s.due = e.max_rep;
    s.part = (int64_t)e.part_items_cap * NCM * TILE;
    s.items = (int64_t)e.part_items_cap;
    s.qorder = (int64_t)e.part_items_cap;
    s.list_tiles = (int64_t)e.list_tiles_cap;
    s.clist = e.clist_cap;
    s.etab = EXP_TN;
    return s;
}

// ---------------------------------------------------------------------------
int add(Engine &e, int k, int &id_out)
{
    const ykc_replica &w = e.reps[k];
    int id = -1;
    for (int r = 0; r < e.max_rep; ++r)
        if (!e.table[r].active) {
            id = r;
            break;
        }
    if (id < 0) return 7;   // YK_ERR_CAPACITY
    const int n = w.n;
    const int nblk = (n + BLOCK - 1) / BLOCK;
    const int start = (int)plan::first_fit(e.used, nblk);
    if (start < 0) return 7;
    Rep &h = e.table[id];
    h = Rep{};
    h.src = k;
    h.n = n;
    h.off = start * BLOCK;
    h.nblocks = nblk;
    h.steps_left = w.steps;
    h.list_items = e.tune.list_items_for(n);
    if (h.list_items > 0) {
        h.segcap = plan::segcap_of(n);
        h.lists_len = plan::list_entries_of(n);
        const int64_t units = (h.lists_len + plan::LIST_UNIT - 1) / plan::LIST_UNIT;
        int64_t s0 = plan::first_fit(e.list_used, units);
        if (s0 < 0) {
            const plan::Growth g = plan::arena_growth(e.list_used, units);
            e.clist_cap = g.new_units * plan::LIST_UNIT;
            e.list_used.resize((size_t)g.new_units, 0);
            s0 = g.start;
        }
        for (int64_t u = s0; u < s0 + units; ++u) e.list_used[(size_t)u] = 1;
        h.lists_off = s0 * plan::LIST_UNIT;
    }
    {   // the engine takes the membership's work-item buffers at admission
        int64_t total = plan::member_items(e.tune, plan::Member{true, n, h.off, nblk,
                                                               h.list_items});
        for (int r = 0; r < e.max_rep; ++r) {
            const Rep &o = e.table[r];
            if (o.active)
                total += plan::member_items(e.tune, plan::Member{true, o.n, o.off, o.nblocks,
                                                                 o.list_items});
        }
        if ((size_t)total > e.part_items_cap) {
            e.part_items_cap = plan::item_capacity((int)total);
            e.part_by.assign(e.part_items_cap, -1);   // a new allocation: nothing written yet
        }
    }
    for (int s = start; s < start + nblk; ++s) e.used[(size_t)s] = 1;
    h.active = true;
    h.fresh = true;
    h.stale = e.plant != 6;
    h.serial = ++e.serials;
    e.dirty = true;
    e.need_accel = true;
    id_out = id;
    return 0;
}

void remove(Engine &e, int id)
{
    Rep &h = e.table[id];
    for (int s = h.off / BLOCK; s < h.off / BLOCK + h.nblocks; ++s) e.used[(size_t)s] = 0;
    if (h.list_items > 0) {
        const int64_t u0 = h.lists_off / plan::LIST_UNIT;
        const int64_t units = (h.lists_len + plan::LIST_UNIT - 1) / plan::LIST_UNIT;
        for (int64_t u = u0; u < u0 + units; ++u) e.list_used[(size_t)u] = 0;
    }
    h = Rep{};
    e.dirty = true;
}

// rebuild(): layout, buffer capacities, k_fill_rep
void rebuild(Engine &e, Checker &ck)
{
    std::vector<plan::Member> members((size_t)e.max_rep);
    for (int r = 0; r < e.max_rep; ++r) {
        const Rep &h = e.table[r];
        members[(size_t)r] = plan::Member{h.active, h.n, h.off, h.nblocks, h.list_items};
    }
    e.L = plan::layout(e.tune, members, e.cap);
    const int nitems = (int)e.L.items.size();
    {   // the items reserved at admission (member_items) are the layout's
        int64_t reserved = 0;
        for (const plan::Member &m : members)
            if (m.active) reserved += plan::member_items(e.tune, m);
        ck.in("work items of the layout against member_items", nitems, reserved, reserved + 1);
    }
    e.rpt->items_max = std::max<int64_t>(e.rpt->items_max, nitems);
    if (e.L.list_tiles.size() > e.list_tiles_cap)
        e.list_tiles_cap = e.L.list_tiles.size() + e.L.list_tiles.size() / 4 + 64;
    if ((size_t)nitems > e.part_items_cap) {
        e.part_items_cap = plan::item_capacity(nitems);
        e.part_by.assign(e.part_items_cap, -1);   // a new allocation: nothing written yet
    }
    if (e.plant == 1 && nitems > 0) e.part_items_cap = (size_t)nitems - 1;
    const Sizes S = sizes_of(e);
    // host copies into device buffers
    ck.where = "rebuild";
    if (!e.L.list_tiles.empty())
        ck.at("list_tiles", (int64_t)e.L.list_tiles.size() - 1, S.list_tiles);
    if (nitems) {
        ck.at("items", nitems - 1, S.items);
        ck.at("qorder", nitems - 1, S.qorder);
    }
    // qorder is a permutation; items of a tile are its tile_items range
    std::vector<char> seen((size_t)nitems, 0);
    for (int k = 0; k < nitems; ++k) {
        const int q = e.L.qorder[(size_t)k];
        if (ck.at("items (qorder)", q, nitems)) {
            if (seen[(size_t)q]) ck.fail("qorder duplicate", q, nitems);
            seen[(size_t)q] = 1;
        }
    }
    for (int t = 0; t < e.cap / TILE; ++t) {
        const int first = e.L.tile_items[2 * (size_t)t], cnt = e.L.tile_items[2 * (size_t)t + 1];
        for (int m = 0; m < cnt; ++m)
            if (ck.at("items (tile_items)", first + m, nitems))
                ck.in("item tile", e.L.items[(size_t)(first + m)].tile, t, t + 1);
    }
    // k_fill_rep: grid cap / BLOCK
    ck.where = "k_fill_rep";
    ++e.rpt->launches;
    e.rep.assign((size_t)e.cap, -1);
    for (int i = 0; i < e.cap; ++i) {
        if (!ck.at("tile_rep", i / TILE, S.tiles)) continue;
        const int r = e.L.tile_rep[(size_t)(i / TILE)];
        if (r >= 0 && ck.at("reps", r, S.reps)) {
            const Rep &h = e.table[r];
            if (i - h.off < h.n) e.rep[(size_t)i] = r;
        }
        ck.at("rep", i, S.slot);
    }
    e.dirty = false;
}

// ---------------------------------------------------------------------------
// The kernels of one step (MODE 1) or of the accel pass (MODE 0).

void k_batch_drift(Engine &e, Checker &ck, bool move)
{
    const Sizes S = sizes_of(e);
    const int nslots = e.L.span, cap = e.cap;
    const int blocks = std::max(1, (nslots + BLOCK - 1) / BLOCK);
    ck.where = move ? "k_batch_drift<move>" : "k_batch_drift<boxes>";
    ++e.rpt->launches;
    for (int i = 0; i < blocks * BLOCK; ++i) {
        int r = -1;
        if (i < nslots && ck.at("rep", i, S.slot)) r = e.rep[(size_t)i];
        const bool valid = r >= 0;
        if (valid && ck.at("reps", r, S.reps)) {
            ck.at("pos", i, S.pos);
            if (move && e.table[r].steps_left > 0) {
                ck.at("state", r, S.state);
                for (int k = 0; k < 3; ++k) {
                    ck.at("vel", (int64_t)k * cap + i, S.comp3);
                    ck.at("acc", (int64_t)k * cap + i, S.comp3);
                    ck.at("vold", (int64_t)k * cap + i, S.comp3);
                }
            }
        }
        const int tile = i / TILE;
        if (tile < nslots / TILE) {
            if (i % TILE == 0 && ck.at("box", tile, S.tiles)) e.box_by[(size_t)tile] = e.launch;
            if (ck.at("tile_rep", tile, S.tiles)) {
                const int tr = e.L.tile_rep[(size_t)tile];
                if (move && tr >= 0 && ck.at("reps", tr, S.reps) && e.table[tr].list_items > 0) {
                    ck.at("maxdisp", 2 * (int64_t)tr + 1, S.maxdisp);
                    if (valid) ck.at("xref", i, S.xref);
                }
            }
        }
    }
}

// Minimum-image distance^2 from point p to the AABB (c, h).
double aabb_dist2(const double p[3], const double c[3], const double h[3], const double L[3])
{
    double d2 = 0.0;
    for (int k = 0; k < 3; ++k) {
        double d = p[k] - c[k];
        d -= L[k] * std::rint(d / L[k]);
        const double g = std::max(std::fabs(d) - h[k], 0.0);
        d2 += g * g;
    }
    return d2;
}

struct Box {
    double c[3], h[3];
};

// AABBs of consecutive groups of `size` slots of a replica in `order`.
std::vector<Box> group_boxes(const ykc_replica &w, const std::vector<int> &order, int size)
{
    const int ng = (w.n + size - 1) / size;
    std::vector<Box> out((size_t)ng);
    for (int g = 0; g < ng; ++g) {
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (int s = g * size; s < std::min(w.n, (g + 1) * size); ++s)
            for (int k = 0; k < 3; ++k) {
                const double x = w.x[3 * (size_t)order[(size_t)s] + k];
                lo[k] = std::min(lo[k], x);
                hi[k] = std::max(hi[k], x);
            }
        for (int k = 0; k < 3; ++k) {
            out[(size_t)g].c[k] = 0.5 * (lo[k] + hi[k]);
            out[(size_t)g].h[k] = 0.5 * (hi[k] - lo[k]);
        }
    }
    return out;
}

// A replica's list geometry for one particle order: tile boxes, and the
// build's counts per (cluster, segment) -- the same for every engine life
// that admits it, so kept across calls.
struct ListGeometry {
    std::vector<Box> tiles;
    std::vector<Count2> counts;        // [clusters * NSEG]
    std::vector<std::vector<int>> ctiles;   // [tiles * NSEG] candidate source tiles
};

std::vector<std::pair<const double *, ListGeometry>> &geometry_cache()
{
    static std::vector<std::pair<const double *, ListGeometry>> cache;
    return cache;
}

ListGeometry &geometry(const ykc_replica &w, const std::vector<int> &order)
{
    auto &cache = geometry_cache();
    for (auto &c : cache)
        if (c.first == w.x && (int)c.second.counts.size() == (w.n + CL - 1) / CL * NSEG)
            return c.second;
    if (cache.size() > 512) cache.clear();
    cache.emplace_back(w.x, ListGeometry{});
    ListGeometry &G = cache.back().second;
    const int ntiles = (w.n + TILE - 1) / TILE, ncl = (w.n + CL - 1) / CL;
    const double rl = w.rccut * 1.05;
    G.tiles = group_boxes(w, order, TILE);
    const std::vector<Box> cl = group_boxes(w, order, CL);
    G.counts.assign((size_t)ncl * NSEG, Count2{0, 0});
    G.ctiles.assign((size_t)ntiles * NSEG, {});
    for (int t = 0; t < ntiles; ++t)
        for (int seg = 0; seg < NSEG; ++seg) {
            std::vector<int> &ct = G.ctiles[(size_t)t * NSEG + seg];
            // the device's scan order: tb + lane * NSEG + seg
            for (int tb = 0; tb < ntiles; tb += TILE * NSEG)
                for (int lane = 0; lane < TILE; ++lane) {
                    const int stl = tb + lane * NSEG + seg;
                    if (stl >= ntiles) continue;
                    double gap2 = 0.0;
                    for (int k = 0; k < 3; ++k) {
                        double d = G.tiles[(size_t)t].c[k] - G.tiles[(size_t)stl].c[k];
                        d -= w.L[k] * std::rint(d / w.L[k]);
                        const double g = std::max(
                            std::fabs(d) - G.tiles[(size_t)t].h[k] - G.tiles[(size_t)stl].h[k], 0.0);
                        gap2 += g * g;
                    }
                    if (gap2 <= rl * rl) ct.push_back(stl);
                }
            for (int wv = 0; wv < CPW; ++wv) {
                const int c = t * CPW + wv;
                if (c >= ncl) continue;
                int nf = 0, ng = 0;
                for (int stl : ct)
                    for (int s = stl * TILE; s < std::min(w.n, (stl + 1) * TILE); ++s) {
                        double p[3];
                        for (int k = 0; k < 3; ++k) p[k] = w.x[3 * (size_t)order[(size_t)s] + k];
                        const double d2 = aabb_dist2(p, cl[(size_t)c].c, cl[(size_t)c].h, w.L);
                        if (d2 > rl * rl) continue;
                        const bool back =
                            w.near_r >= 0.0 ? d2 <= w.near_r * w.near_r : (s % 7) == 0;
                        if (back) ++ng;
                        else ++nf;
                    }
                G.counts[(size_t)c * NSEG + seg] = Count2{nf, ng};
            }
        }
    return G;
}

// k_batch_clist_build over the list replicas' tiles (grid n_list_tiles x
// NSEG), for replicas whose lists are stale: the ctiles, pos, xref, box,
// clist and ccounts indices of every wave, and the counts it leaves.
void k_batch_clist_build(Engine &e, Checker &ck, std::vector<std::vector<int>> &order)
{
    const Sizes S = sizes_of(e);
    ck.where = "k_batch_clist_build";
    ++e.rpt->launches;
    constexpr int W = BLOCK / TILE;
    constexpr int CTILES = CLIST_MAX_N / TILE / NSEG + 1;
    const int nlt = (int)e.L.list_tiles.size();
    for (int bx = 0; bx < nlt; ++bx) {
        if (!ck.at("list_tiles", bx, S.list_tiles)) continue;
        const int t = e.L.list_tiles[(size_t)bx];
        if (!ck.at("tile_rep", t, S.tiles)) continue;
        const int r = e.L.tile_rep[(size_t)t];
        if (r < 0) continue;
        if (!ck.at("reps", r, S.reps)) continue;
        const Rep &h = e.table[r];
        ck.at("maxdisp", 2 * (int64_t)r + 1, S.maxdisp);
        if (h.list_items <= 0 || !h.stale) continue;
        ListGeometry &G = geometry(e.reps[h.src], order[(size_t)r]);
        const int tile0 = h.off / TILE, ntiles = (h.n + TILE - 1) / TILE;
        const int tl = t - tile0;
        ck.in("list tile", tl, 0, ntiles);
        if (ck.at("box", t, S.tiles))
            ck.in("box written by this step's drift", e.box_by[(size_t)t], e.launch, e.launch + 1);
        for (int seg = 0; seg < NSEG; ++seg) {
            for (int tb = 0; tb < ntiles; tb += TILE * NSEG)
                for (int lane = 0; lane < TILE; ++lane) {
                    const int stl = tb + lane * NSEG + seg;
                    if (stl < ntiles && ck.at("box", tile0 + stl, S.tiles))
                        ck.in("box written by this step's drift", e.box_by[(size_t)(tile0 + stl)],
                              e.launch, e.launch + 1);
                }
            const std::vector<int> &ct = G.ctiles[(size_t)tl * NSEG + seg];
            const int ncand = (int)ct.size();
            if (ncand > 0) ck.at("ctiles", ncand - 1, CTILES);
            for (int wv = 0; wv < W; ++wv) {
                const int cll = tl * CPW + wv;
                const bool cvalid = cll * CL < h.n;
                for (int lane = 0; lane < CL; ++lane) {
                    const int il = cll * CL + lane;
                    if (il >= h.n) continue;
                    ck.at("pos", (int64_t)h.off + il, S.pos);
                    if (seg == 0) ck.at("xref", (int64_t)h.off + il, S.xref);
                }
                // staged sources: slots ctiles[idx] * 64 + lane below n
                for (int idx = 0; idx < ncand; ++idx) {
                    const int base = ct[(size_t)idx] * TILE;
                    ck.in("source tile", ct[(size_t)idx], 0, ntiles);
                    const int last = std::min(h.n, base + TILE) - 1;
                    if (last >= base) {
                        ck.at("pos", (int64_t)h.off + base, S.pos);
                        ck.at("pos", (int64_t)h.off + last, S.pos);
                    }
                }
                if (!cvalid) continue;
                Count2 c = G.counts[(size_t)cll * NSEG + seg];
                if (e.plant == 3 && cll == 0 && seg == 0) c.x = h.segcap + 1 - c.y;
                const int64_t seg_off = h.lists_off + ((int64_t)cll * NSEG + seg) * h.segcap;
                // the entries written: fast at [0, nf), back at [segcap - ng,
                // segcap) of the segment; each is a slot of the candidate tiles
                ck.in("segment count", (int64_t)c.x + c.y, 0, (int64_t)h.segcap + 1);
                ck.in("segment count (candidate slots)", (int64_t)c.x + c.y, 0,
                      (int64_t)ncand * TILE + 1);
                if (c.x > 0) {
                    ck.at("clist", seg_off, S.clist);
                    ck.at("clist", seg_off + c.x - 1, S.clist);
                }
                if (c.y > 0) {
                    ck.at("clist", seg_off + h.segcap - c.y, S.clist);
                    ck.at("clist", seg_off + h.segcap - 1, S.clist);
                }
                ck.in("list segment (replica region)", seg_off + h.segcap - 1, h.lists_off,
                      h.lists_off + h.lists_len);
                const int64_t ci = ((int64_t)t * CPW + wv) * NSEG + seg;
                if (ck.at("ccounts", ci, S.ccounts)) {
                    e.ccounts[(size_t)ci] = c;
                    e.ccounts_by[(size_t)ci] = h.serial;
                }
            }
        }
    }
}

void batch_tile_item(Engine &e, Checker &ck, const plan::Item &it, const Sizes &S)
{
    const Rep &h = e.table[it.rep];
    const int t = it.tile;
    const int tile0 = h.off / TILE, ntiles = (h.n + TILE - 1) / TILE;
    ck.in("tile item target tile", t, tile0, tile0 + ntiles);
    if (ck.at("box", t, S.tiles))
        ck.in("box written by this step's drift", e.box_by[(size_t)t], e.launch, e.launch + 1);
    for (int lane = 0; lane < TILE; ++lane) {
        const int i = t * TILE + lane;
        if (i - h.off < h.n) ck.at("pos", i, S.pos);
    }
    for (int st = it.s0; st < it.s1 + (e.plant == 4 ? 1 : 0); ++st) {
        ck.in("tile item source tile", st, tile0, tile0 + ntiles);
        if (ck.at("box", st, S.tiles))
            ck.in("box written by this step's drift", e.box_by[(size_t)st], e.launch, e.launch + 1);
        for (int lane = 0; lane < TILE; ++lane) {
            const int j = st * TILE + lane;
            if (j - h.off < h.n) ck.at("pos", j, S.pos);
        }
        const int jcount = std::min(TILE, h.n - (st * TILE - h.off));
        if (jcount > 0) ck.in("LDS source (tile item)", jcount - 1, 0, TILE);
    }
}

void batch_list_item(Engine &e, Checker &ck, const plan::Item &it, const Sizes &S)
{
    const Rep &h = e.table[it.rep];
    const int M = -it.s1, m = it.s0;
    const int t = it.tile;
    const int tile0 = h.off / TILE;
    ck.in("list item share", m, 0, M);
    int lo[CPW], hi[CPW], iters = 0;
    SegList sl[CPW];
    Count2 cnt[CPW][NSEG];
    for (int g = 0; g < CPW; ++g) {
        const int clg = t * CPW + g, cll = (t - tile0) * CPW + g;
        const bool cvalid = cll * CL < h.n;
        Count2 sc[NSEG];
        int total = 0;
        for (int sg = 0; sg < NSEG; ++sg) {
            sc[sg] = Count2{0, 0};
            if (cvalid) {
                const int64_t ci = (int64_t)clg * NSEG + sg;
                if (ck.at("ccounts", ci, S.ccounts)) {
                    sc[sg] = e.ccounts[(size_t)ci];
                    ck.in("ccounts written by this replica's build (serial)",
                          e.ccounts_by[(size_t)ci], h.serial, h.serial + 1);
                }
            }
            total += sc[sg].x + sc[sg].y;
            cnt[g][sg] = sc[sg];
        }
        list_share(total, M, m, lo[g], hi[g]);
        if (e.plant == 2 && m == M - 1 && total > 0) hi[g] += 1;
        iters = std::max(iters, hi[g] - lo[g]);
        sl[g] = seg_list(nullptr, h.segcap, sc, hi[g]);
        for (int gl = 0; gl < CL; ++gl) {
            const int i = t * TILE + g * CL + gl;
            if (i - h.off < h.n) ck.at("pos", i, S.pos);
        }
    }
    // entries read: lo + gl and lo + CL + gl before the loop, then
    // lo + b0 + 2 CL + gl per batch; SegList::entry reads only v < hi
    for (int g = 0; g < CPW; ++g) {
        const int cll = (t - tile0) * CPW + g;
        const int64_t mine =
            h.lists_off + (int64_t)(cll + (e.plant == 5 ? 1 : 0)) * NSEG * h.segcap;
        auto read = [&](int v) {
            if (v >= sl[g].hi) return;   // entry 0 (slot 0 of the replica) without a read
            bool back = false;
            const long long o = sl[g].offset(v, back);
            ck.in("list entry (cluster region)", o, 0, (int64_t)NSEG * h.segcap);
            ck.in("list entry (replica region)", mine + o, h.lists_off, h.lists_off + h.lists_len);
            ck.at("clist", mine + o, S.clist);
            // a written entry of its segment
            const int sg = (int)(o / h.segcap), within = (int)(o % h.segcap);
            if (ck.in("list entry segment", sg, 0, NSEG)) {
                if (back) ck.in("back entry (written)", within, h.segcap - cnt[g][sg].y, h.segcap);
                else ck.in("fast entry (written)", within, 0, cnt[g][sg].x);
            }
            ++e.rpt->list_entries_read;
        };
        // the walk reads v = lo + gl and lo + CL + gl, then lo + b0 + 2 CL + gl
        // for b0 < iters: every v of [lo, lo + iters + 2 CL) below hi
        const int vend = std::min(sl[g].hi, lo[g] + iters + 2 * CL);
        if (e.exhaustive) {
            for (int v = lo[g]; v < vend; ++v) read(v);
            // the kernel reaches these entries through a cursor per lane
            for (int gl = 0; gl < CL; ++gl) {
                SegList::Cursor cu = sl[g].cursor(lo[g] + gl, CL);
                for (int v = lo[g] + gl; v < vend; v += CL) {
                    bool back = false;
                    const long long want = sl[g].offset(v, back);
                    ck.in("list cursor offset", (long long)cu.off - want, 0, 1);
                    sl[g].advance(cu, v + CL, CL);
                }
            }
        } else if (lo[g] < vend) {
            const SegList &q = sl[g];
            const int cuts[] = {q.g1, q.g2, q.g3, q.nback, q.nback + q.f1, q.nback + q.f2,
                                q.nback + q.f3};
            read(lo[g]);
            read(vend - 1);
            for (int c : cuts) {
                if (c - 1 >= lo[g] && c - 1 < vend) read(c - 1);
                if (c >= lo[g] && c < vend) read(c);
            }
        }
        if (lo[g] < vend) ck.at("pos", h.off, S.pos);
    }
}

void k_batch_pairs(Engine &e, Checker &ck, bool meas, bool accel)
{
    const Sizes S = sizes_of(e);
    const int nitems = (int)e.L.items.size();
    const int NC = meas ? NCM : 3;
    ck.where = meas ? "k_batch_pairs<measuring>" : "k_batch_pairs<force-only>";
    ++e.rpt->launches;
    const int waves = e.L.waves;
    const int blocks = plan::persistent_blocks(e.cus, waves, nitems);
    e.rpt->blocks = blocks;
    e.rpt->waves = waves;
    // LDS: dynamic (double4 + float4) x 64 per wave, static exp table
    const size_t lds_dyn = (32 + 16) * (size_t)TILE * waves, lds_static = 8 * EXP_TN;
    ck.in("LDS bytes", (int64_t)(lds_dyn + lds_static), 0, (int64_t)LDS_LIMIT + 1);
    ck.in("workgroup threads", TILE * waves, 1, 512 + 1);
    for (int thr = 0; thr < TILE * waves; ++thr) {
        const int wave = thr / TILE, lane = thr & (TILE - 1);
        ck.in("LDS sw", (int64_t)(wave * TILE + lane) * 32 + 31, 0, (int64_t)lds_dyn);
        ck.in("LDS sf", (int64_t)TILE * waves * 32 + (int64_t)(wave * TILE + lane) * 16 + 15, 0,
              (int64_t)lds_dyn);
    }
    for (int k = 0; k < nitems; ++k) {
        if (!ck.at("qorder", k, S.qorder)) continue;
        const int q = e.L.qorder[(size_t)k];
        if (!ck.at("items", q, S.items)) continue;
        const plan::Item it = e.L.items[(size_t)q];
        if (ck.at("part", ((int64_t)q * NC + NC - 1) * TILE + TILE - 1, S.part))
            e.part_by[(size_t)q] = e.launch;
        if (!ck.at("reps", it.rep, S.reps) || !ck.at("state", it.rep, S.state)) continue;
        const Rep &h = e.table[it.rep];
        if (!h.active) {
            ck.fail("item of an inactive replica", it.rep, e.max_rep);
            continue;
        }
        if (accel ? !h.fresh : h.steps_left <= 0) continue;
        if (it.s1 < 0) batch_list_item(e, ck, it, S);
        else batch_tile_item(e, ck, it, S);
    }
}

void k_batch_finalize(Engine &e, Checker &ck, bool meas, bool accel)
{
    const Sizes S = sizes_of(e);
    const int nslots = e.L.span, cap = e.cap;
    const int blocks = std::max(1, (nslots + BLOCK - 1) / BLOCK);
    const int NC = meas ? NCM : 3;
    ck.where = "k_batch_finalize";
    ++e.rpt->launches;
    for (int i = 0; i < blocks * BLOCK; ++i) {
        int r = -1;
        if (i < nslots && ck.at("rep", i, S.slot)) r = e.rep[(size_t)i];
        if (r >= 0 && ck.at("state", r, S.state) &&
            (accel ? e.table[r].fresh : e.table[r].steps_left > 0)) {
            ck.at("reps", r, S.reps);
            const int t = i / TILE, lane = i % TILE;
            if (!ck.at("tile_items", 2 * (int64_t)t + 1, S.tile_items2)) continue;
            const int first = e.L.tile_items[2 * (size_t)t], cnt = e.L.tile_items[2 * (size_t)t + 1];
            for (int m = 0; m < cnt; ++m)
                for (int k = 0; k < NC; ++k)
                    ck.at("part", ((int64_t)(first + m) * NC + k) * TILE + lane, S.part);
            for (int m = 0; m < cnt; ++m)
                if (ck.at("part (stamp)", first + m, (int64_t)e.part_by.size()))
                    ck.in("partial sums written by this step's pair launch",
                          e.part_by[(size_t)(first + m)], e.launch, e.launch + 1);
            ck.at("pos", i, S.pos);
            ck.at("orig", i, S.slot);
            ck.at("den", i, S.slot);
            ck.at("aii", i, S.slot);
            for (int k = 0; k < 3; ++k) {
                ck.at("acc", (int64_t)k * cap + i, S.comp3);
                if (!accel) {
                    ck.at("vold", (int64_t)k * cap + i, S.comp3);
                    ck.at("vel", (int64_t)k * cap + i, S.comp3);
                }
            }
        }
        if (meas && i % BLOCK == 0) ck.at("blockpart", (int64_t)(i / BLOCK) * NQ + NQ - 1, S.blockpart);
    }
}

void k_batch_replica(Engine &e, Checker &ck, bool meas, bool accel)
{
    const Sizes S = sizes_of(e);
    ck.where = "k_batch_replica";
    ++e.rpt->launches;
    for (int r = 0; r < e.max_rep; ++r) {
        ck.at("reps", r, S.reps);
        Rep &h = e.table[r];
        const bool live = h.active && h.n > 0 && (accel ? h.fresh : h.steps_left > 0);
        if (h.list_items > 0 && h.stale && live) {
            ck.at("maxdisp", 2 * (int64_t)r + 1, S.maxdisp);
            h.stale = false;
        }
        if (meas && live) {
            const int blk0 = h.off / BLOCK, nblk = (h.n + BLOCK - 1) / BLOCK;
            for (int b = 0; b < nblk; ++b)
                ck.at("blockpart", (int64_t)(blk0 + b) * NQ + NQ - 1, S.blockpart);
            if (accel) {
                ck.at("init_records", r, S.init_records);
            } else {
                if (h.rec_count < e.rec_cap) ck.at("records", (int64_t)r * e.rec_cap + h.rec_count, S.records);
                ++h.rec_count;
            }
        }
        ck.at("state", r, S.state);
        if (accel) {
            if (h.active) h.fresh = false;
        } else if (h.active && h.steps_left > 0) {
            h.steps_left -= 1;
        }
    }
}

void step(Engine &e, Checker &ck, std::vector<std::vector<int>> &order, bool meas, bool accel)
{
    ++e.launch;
    std::vector<char> live((size_t)e.max_rep, 0);
    for (int r = 0; r < e.max_rep; ++r) live[(size_t)r] = e.table[r].active && e.table[r].steps_left > 0;
    k_batch_drift(e, ck, !accel);
    if (!e.L.list_tiles.empty()) k_batch_clist_build(e, ck, order);
    k_batch_pairs(e, ck, meas, accel);
    k_batch_finalize(e, ck, meas, accel);
    k_batch_replica(e, ck, meas, accel);
    if (accel) return;
    for (int r = 0; r < e.max_rep; ++r) {
        Rep &h = e.table[r];
        if (!live[(size_t)r]) continue;
        ++h.steps_done;
        // lists go stale again: from a random start the device rebuilds
        // about every second step
        if (h.list_items > 0 && h.steps_done % 2 == 1) h.stale = true;
    }
}

// resort_due: k_batch_sort_keys, the radix sort, k_batch_permute, copies.
void resort_due(Engine &e, Checker &ck, std::vector<std::vector<int>> &order)
{
    if (e.resort_every <= 0) return;
    std::vector<int> due((size_t)e.max_rep, 0);
    bool any = false;
    for (int r = 0; r < e.max_rep; ++r) {
        Rep &h = e.table[r];
        if (h.active && h.steps_left > 0 && h.steps_done > 0 && h.steps_done % e.resort_every == 0 &&
            h.sorted_at != h.steps_done) {
            due[(size_t)r] = 1;
            h.sorted_at = h.steps_done;
            any = true;
        }
    }
    if (!any) return;
    const Sizes S = sizes_of(e);
    const int span = e.L.span;
    ck.where = "k_batch_sort_keys";
    ++e.rpt->launches;
    for (int i = 0; i < span; ++i) {
        const int r = e.rep[(size_t)i];
        if (r >= 0 && ck.at("due", r, S.due) && due[(size_t)r]) {
            ck.at("reps", r, S.reps);
            ck.at("maxdisp", 2 * (int64_t)r, S.maxdisp);
            ck.at("pos", i, S.pos);
            e.table[r].stale = true;
        }
        ck.at("keys", i, S.keys);
        ck.at("vals", i, S.keys);
    }
    ck.where = "radix sort / k_batch_permute";
    ck.in("sorted items", span, 0, S.keys + 1);
    // the permutation keeps every slot inside its own replica (keys
    // (offset << 32 | cell) for due replicas, (slot << 32) for the rest); the
    // slots keep their order here (the particles do not move), so the list
    // counts of the last build stay those of the same clusters
    (void)order;
    for (int i = 0; i < span; ++i) {
        ck.at("pos2", i, S.pos);
        for (int k = 0; k < 3; ++k) ck.at("vel2", (int64_t)k * e.cap + i, S.comp3);
    }
}

}  // namespace

extern "C" void ykc_clear_cache(void) { geometry_cache().clear(); }

// The pair kernels' list walk (clist_share): lane gl of a 16-lane group reads
// entries lo + gl, lo + gl + stride, ... through a SegList::Cursor, first at
// lo + gl, then advanced by stride per batch, for as many batches as the
// wave runs (iters rounded up, plus the one-batch prefetch).  counts: the
// cluster's segments as (fast, back) pairs.  Every entry below hi is compared
// with SegList::offset and checked to lie in its segment's written part;
// visits: how often each v of [0, hi) is read (each v of [lo, hi) is read
// once).  plant 1: the cursor never recomputes at a run's end; plant 2: it
// steps fast runs the wrong way.  Returns the number of mismatches.
extern "C" int64_t ykc_cursor_walk(const int32_t *counts, int32_t segcap, int32_t lo,
                                   int32_t hi, int32_t iters, int32_t stride, int32_t plant,
                                   int32_t *visits)
{
    Count2 sc[NSEG];
    for (int k = 0; k < NSEG; ++k) sc[k] = Count2{counts[2 * k], counts[2 * k + 1]};
    const SegList sl = seg_list<Count2>(nullptr, segcap, sc, hi);
    int64_t bad = 0;
    for (int gl = 0; gl < stride; ++gl) {
        int v = lo + gl;
        SegList::Cursor cu = sl.cursor(v, stride);
        for (int b0 = -stride; b0 < iters; b0 += stride, v += stride) {
            if (v < hi) {
                bool back = false;
                const long long want = sl.offset(v, back);
                if ((long long)cu.off != want || (cu.step < 0) != back) ++bad;
                const int sg = (int)(cu.off / (unsigned)segcap);
                const int within = (int)(cu.off % (unsigned)segcap);
                if (sg < 0 || sg >= NSEG) {
                    ++bad;
                } else if (cu.step < 0 ? within < segcap - sc[sg].y : within >= sc[sg].x) {
                    ++bad;
                }
                ++visits[v];
            }
            const int vn = v + stride;
            if (plant == 1) {
                cu.off += (unsigned)cu.step;
            } else if (plant == 2 && cu.step > 0 && vn < cu.vend) {
                cu.off -= (unsigned)cu.step;
            } else {
                sl.advance(cu, vn, stride);
            }
        }
    }
    return bad;
}

extern "C" int ykc_check(const ykc_engine *cfg, const ykc_replica *reps, int32_t nrep,
                         const int32_t *ops, int32_t nops, int32_t exhaustive, ykc_report *out)
{
    std::memset(out, 0, sizeof *out);
    Engine e;
    e.rpt = out;
    e.reps = reps;
    e.exhaustive = exhaustive != 0;
    e.plant = cfg->plant;
    Checker ck{out, "create"};
    // yk_batch_create
    if (cfg->capacity < 1 || cfg->capacity > ((int64_t)1 << 28) || cfg->max_replicas < 1) {
        out->status = 1;
        return 1;
    }
    e.cap = (int)((cfg->capacity + BLOCK - 1) / BLOCK * BLOCK);
    e.max_rep = cfg->max_replicas;
    e.rec_cap = std::max(1, (int)cfg->max_records);
    e.cus = cfg->cus;
    e.resort_every = cfg->resort_every;
    e.table.assign((size_t)e.max_rep, Rep{});
    e.used.assign((size_t)(e.cap / BLOCK), 0);
    e.ccounts.assign((size_t)(e.cap / CL) * NSEG, Count2{0, 0});
    e.ccounts_by.assign(e.ccounts.size(), -1);
    e.box_by.assign((size_t)(e.cap / TILE), -1);
    rebuild(e, ck);
    // yk_batch_set_tuning (its refusals are the engine's: checked by the caller)
    for (int k = 0; k < cfg->nbuckets; ++k) {
        e.tune.n_upper.push_back(cfg->n_upper[k]);
        e.tune.chunk_tiles.push_back(cfg->chunk_tiles[k]);
        e.tune.waves.push_back(cfg->waves[k]);
        e.tune.list_items.push_back(cfg->list_items[k]);
    }
    e.dirty = true;
    std::vector<int> id_of((size_t)nrep, -1);
    std::vector<std::vector<int>> order((size_t)e.max_rep);   // slot -> caller particle
    for (int op = 0; op < nops; ++op) {
        const int code = ops[2 * op], arg = ops[2 * op + 1];
        if (code == YKC_ADD) {
            int id = -1;
            if (add(e, arg, id)) {   // no replica slot or arena room: YK_ERR_CAPACITY
                ++out->refused;
                continue;
            }
            id_of[(size_t)arg] = id;
            order[(size_t)id] = compact_order(reps[arg]);
        } else if (code == YKC_REMOVE) {
            const int id = id_of[(size_t)arg];
            if (id >= 0) remove(e, id);
            id_of[(size_t)arg] = -1;
        } else if (code == YKC_RUN) {
            if (e.dirty) rebuild(e, ck);
            if (e.need_accel) {
                step(e, ck, order, true, true);
                e.need_accel = false;
            }
            const int nind = 10;
            for (int s = 1; s <= arg; ++s) {
                resort_due(e, ck, order);
                step(e, ck, order, s % nind == 0, false);
            }
            resort_due(e, ck, order);
        }
    }
    out->clist_cap = e.clist_cap;
    out->part_items_cap = (int64_t)e.part_items_cap;
    out->span = e.L.span;
    return out->violations ? 2 : 0;
}
