// yk_batch_plan.hpp -- host-side planning of the batched ensemble engine:
// the work shape per N bucket, first-fit placement of replicas in the slot
// arena and of their cluster lists in the list arena, and the work items,
// tile item ranges and queue order of a membership.  Plain C++ (no HIP): the
// engine (yk_batch.hip) plans with it, and a CPU check of the engine's device
// indexing runs the same plans without a GPU.
#pragma once
#include <algorithm>
#include <cstdint>
#include <numeric>
#include <vector>

#include "yk_layout.hpp"

namespace yk {
namespace plan {

constexpr int64_t LIST_UNIT = 1 << 20;   // list-arena allocation unit (entries)

// Work shape per N bucket (replicas of n <= n_upper[k] use bucket k).
struct Tuning {
    std::vector<int> n_upper, chunk_tiles, waves, list_items;

    int bucket_of(int n) const
    {
        for (size_t k = 0; k < n_upper.size(); ++k)
            if (n <= n_upper[k]) return (int)k;
        return -1;
    }
    // tile items: source tiles per item
    int chunk_tiles_for(int n) const
    {
        const int nt = (n + TILE - 1) / TILE;
        const int k = bucket_of(n);
        if (k >= 0) return std::max(1, std::min(nt, chunk_tiles[k]));
        return std::max(1, (nt + 3) / 4);   // default: four source ranges per target tile
    }
    // work items per tile of a list replica (0: tile items); without a
    // bucket: 16 up to 65,536 particles (1 GiB of lists at most)
    int list_items_for(int n) const
    {
        if (n > CLIST_MAX_N) return 0;
        const int k = bucket_of(n);
        if (k >= 0) return list_items[k];
        return n <= 65536 ? 16 : 0;
    }
};

// Entries per list segment of a replica of n particles: every slot of the
// source tiles the segment covers, so a build never overflows it.
constexpr int segcap_of(int n) { return ((n + TILE - 1) / TILE + NSEG - 1) / NSEG * TILE; }

inline int64_t list_entries_of(int n)
{
    return (int64_t)((n + CL - 1) / CL) * NSEG * segcap_of(n);
}

// Work items of a replica of n particles.
inline int64_t items_of(const Tuning &t, int n)
{
    const int64_t nt = (n + TILE - 1) / TILE;
    const int M = t.list_items_for(n);
    if (M > 0) return nt * M;
    const int64_t C = t.chunk_tiles_for(n);
    return nt * ((nt + C - 1) / C);
}

// First run of `units` free units in `used`, or -1.
inline int64_t first_fit(const std::vector<char> &used, int64_t units)
{
    for (int64_t u = 0, run = 0; u < (int64_t)used.size(); ++u) {
        run = used[(size_t)u] ? 0 : run + 1;
        if (run == units) return u - units + 1;
    }
    return -1;
}

// Growth of the list arena when no run of `units` free units exists: the used
// prefix stays where it is (its regions keep their offsets), the arena at
// least doubles, and the new region starts after the last used unit.
struct Growth {
    int64_t new_units, start;
};

inline Growth arena_growth(const std::vector<char> &used, int64_t units)
{
    int64_t tail = 0;
    while (tail < (int64_t)used.size() && !used[used.size() - 1 - (size_t)tail]) ++tail;
    const int64_t old_units = (int64_t)used.size();
    return Growth{std::max(old_units + units - tail, 2 * old_units), old_units - tail};
}

// What the layout needs of a replica slot of the engine's table.
struct Member {
    bool active = false;
    int n = 0, off = 0, nblocks = 0;   // slots [off, off + n), 256-slot blocks held
    int list_items = 0;                // > 0: cluster lists, this many items per tile
};

// Work items of a member as layout() makes them: per target tile, M list
// shares or one item per chunk of source tiles.
inline int64_t member_items(const Tuning &t, const Member &h)
{
    const int64_t nt = (h.n + TILE - 1) / TILE;
    if (h.list_items > 0) return nt * h.list_items;
    const int64_t C = t.chunk_tiles_for(h.n);
    return nt * ((nt + C - 1) / C);
}

// A work item as the pair kernel reads it (BatchItem's layout): tile item =
// target tile and source tiles [s0, s1), global indices; list item (s1 < 0) =
// the target tile's four clusters, share s0 of -s1.
struct Item {
    int rep, tile, s0, s1;
};

struct Layout {
    std::vector<Item> items;
    std::vector<int> tile_rep;      // [cap / TILE] replica of a tile, -1 if none
    std::vector<int> tile_items;    // [2 cap / TILE] (first item, item count) per tile
    std::vector<int> qorder;        // queue position -> item
    std::vector<int> list_tiles;    // tiles of list replicas (the build's grid)
    int span = 0;                   // used extent of the arena (slots)
    int nlist = 0;                  // list replicas
    int waves = 4;                  // waves per workgroup of the persistent launch
};

// Items, tile item ranges, queue order and launch shape of a membership in an
// arena of cap slots.
inline Layout layout(const Tuning &tune, const std::vector<Member> &table, int cap)
{
    Layout L;
    const int ntiles_cap = cap / TILE;
    L.tile_rep.assign((size_t)ntiles_cap, -1);
    L.tile_items.assign(2 * (size_t)ntiles_cap, 0);
    std::vector<int> order;
    for (int r = 0; r < (int)table.size(); ++r)
        if (table[r].active) order.push_back(r);
    // largest replicas first: their items start early and the queue's tail
    // is made of small items
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int c) { return table[a].n > table[c].n; });
    for (int r : order) {
        const Member &h = table[r];
        const int t0 = h.off / TILE, nt = (h.n + TILE - 1) / TILE;
        const int C = tune.chunk_tiles_for(h.n);
        const int M = h.list_items;
        L.nlist += (M > 0) ? 1 : 0;
        for (int t = t0; t < t0 + nt; ++t) {
            L.tile_rep[(size_t)t] = r;
            const int first = (int)L.items.size();
            if (M > 0) {
                for (int m = 0; m < M; ++m) L.items.push_back(Item{r, t, m, -M});
            } else {
                for (int s = t0; s < t0 + nt; s += C)
                    L.items.push_back(Item{r, t, s, std::min(s + C, t0 + nt)});
            }
            L.tile_items[2 * (size_t)t] = first;
            L.tile_items[2 * (size_t)t + 1] = (int)L.items.size() - first;
        }
    }
    for (int r : order) {
        const Member &h = table[r];
        L.span = std::max(L.span, h.off + h.nblocks * BLOCK);
        if (h.list_items > 0)
            for (int t = h.off / TILE; t < h.off / TILE + (h.n + TILE - 1) / TILE; ++t)
                L.list_tiles.push_back(t);
    }
    // launch shape: the waves per workgroup of the bucket carrying the most
    // pair work
    {
        std::vector<double> work(tune.n_upper.size(), 0.0);
        for (int r : order) {
            const int k = tune.bucket_of(table[r].n);
            if (k >= 0) work[(size_t)k] += (double)table[r].n * table[r].n;
        }
        double best = 0.0;
        for (size_t k = 0; k < work.size(); ++k)
            if (work[k] > best) {
                best = work[k];
                L.waves = tune.waves[k];
            }
    }
    // queue order: a list item's share m leads with the list's back entries
    // (FP64 in the mixed form), so shares go out by m -- every list's share 0
    // first, tile items with them -- and the longest items start first;
    // otherwise the build order (largest replica first)
    L.qorder.resize(L.items.size());
    std::iota(L.qorder.begin(), L.qorder.end(), 0);
    std::stable_sort(L.qorder.begin(), L.qorder.end(), [&](int a, int c) {
        const int ma = L.items[(size_t)a].s1 < 0 ? L.items[(size_t)a].s0 : 0;
        const int mc = L.items[(size_t)c].s1 < 0 ? L.items[(size_t)c].s0 : 0;
        return ma < mc;
    });
    return L;
}

// Item capacity of the partial-sum, item and queue-order buffers for nitems
// items (a quarter of headroom, so small changes of membership reuse them).
inline size_t item_capacity(int nitems)
{
    return std::max<size_t>((size_t)nitems + (size_t)nitems / 4, 1024);
}

// Workgroups of the persistent pair launch: every CU filled with 32 waves
// (8 per SIMD), never more workgroups than items need.
inline int persistent_blocks(int cus, int waves, int nitems)
{
    const int per_cu = std::max(1, 32 / waves);
    const int blocks = cus * per_cu;
    const int need = (nitems + waves - 1) / waves;
    return std::max(1, std::min(blocks, need));
}


}  // namespace plan
}  // namespace yk
