// yk_layout.hpp -- sizes that fix the device data layout: workgroup, tile,
// cluster and list-segment granularity.  Plain C++ (no HIP), so host-only
// code (the batched engine's planner and its CPU checks) sees the same values
// as the kernels.
#pragma once

namespace yk {

constexpr int BLOCK = 256;   // threads per workgroup (4 wave64)
constexpr int TILE = 64;     // particles per tile (one wave)
constexpr int NQ = 14;       // reduced scalars: E, W, PP[6], KE[6]

// A cluster is 16 consecutive sorted slots; a wave of the cluster-list pair
// kernel takes four.
constexpr int CL = 16;              // slots per cluster
constexpr int CPW = TILE / CL;      // clusters per wave in the pair kernel
// Cluster lists hold up to n entries per cluster (n^2/4 bytes in all): used
// up to CLIST_MAX_N particles (16 GiB of lists at the top); larger single
// systems take the 64-slot tile lists (n^2/16 bytes), larger batched replicas
// tile items.
constexpr int CLIST_MAX_N = 262144;
// A cluster's list is stored as NSEG segments: segment s holds the entries
// whose source tile index is s mod NSEG (each built by its own workgroup, for
// parallelism), each with its fast entries from the front and general ones
// from the back of a region of segcap entries; the list is the segments in
// order.
constexpr int NSEG = 4;

// List entries: the source slot in the low 26 bits, a per-axis image code in
// bits 26-31.
constexpr unsigned SLOT_MASK = (1u << 26) - 1u;

// A work item's share m of M of one cluster's list of `total` entries:
// [lo, hi).  constexpr, so device and host code share it.
constexpr void list_share(int total, int M, int m, int &lo, int &hi)
{
    const int per = (total + M - 1) / M;
    lo = m * per < total ? m * per : total;
    hi = lo + per < total ? lo + per : total;
}

// A cluster's list as the pair kernels walk it: entry v of [0, hi), the back
// (general) entries of the four segments first, then the fast ones.  Segment
// sg occupies base[sg * segcap .. (sg + 1) * segcap): fast entries from the
// front, back entries from the end.  constexpr, so device and host code (the
// kernels and their CPU checks) share it.
struct SegList {
    const unsigned *base;
    int segcap, hi;
    int nback;           // back entries of all segments
    int g1, g2, g3;      // back entries of segments below 1, 2, 3
    int f1, f2, f3;      // fast entries of segments below 1, 2, 3

    // A walk through the list with a fixed stride: the offset of the current
    // entry v, the offset's step to entry v + stride while that stays in v's
    // run (+stride in a segment's fast entries, -stride in its back entries),
    // and the first v past the run (INT_MAX at or past hi, where nothing is
    // read).  Offsets are below NSEG * segcap, which fits 32 bits.
    struct Cursor {
        unsigned off;
        int step, vend;
    };

    [[gnu::always_inline]] constexpr Cursor cursor(int v, int stride) const
    {
        const bool back = v < nback;
        // arithmetic selection of the segment (selects of addresses would put
        // the fields on the stack)
        const int w = back ? v : v - nback;
        const int c1 = back ? g1 : f1, c2 = back ? g2 : f2, c3 = back ? g3 : f3;
        const int s1 = w >= c1, s2 = w >= c2, s3 = w >= c3;
        const int sg = s1 + s2 + s3;
        const int start = s1 * (c1 + s2 * (c2 - c1 + s3 * (c3 - c2)));
        const int o = w - start;
        // the run's end: the next segment's first entry of the same kind; the
        // back entries of segment 3 end at nback, the fast ones of segment 3
        // at the list's end
        const int wend = s3 ? 0x7fffffff : s2 ? c3 : s1 ? c2 : c1;
        Cursor cu{};
        cu.off = (unsigned)(sg * segcap + (back ? segcap - 1 - o : o));
        cu.step = back ? -stride : stride;
        cu.vend = v >= hi ? 0x7fffffff : back ? (s3 ? nback : wend) : (s3 ? wend : nback + wend);
        return cu;
    }
    // The cursor of v + stride from that of v.
    [[gnu::always_inline]] constexpr void advance(Cursor &cu, int vnew, int stride) const
    {
        if (vnew >= cu.vend) cu = cursor(vnew, stride);
        else cu.off += (unsigned)cu.step;
    }
    // The entry at the cursor of v, or 0 (slot 0, a valid source) for
    // v >= hi; back: whether it is a back entry (false for v >= hi).
    [[gnu::always_inline]] constexpr unsigned fetch(const Cursor &cu, int v, bool &back) const
    {
        back = false;
        if (v >= hi) return 0u;
        back = cu.step < 0;
        return base[cu.off];
    }

    // The entry's offset from base (v < hi); back: whether it is a back entry.
    [[gnu::always_inline]] constexpr long long offset(int v, bool &back) const
    {
        back = v < nback;
        return (long long)cursor(v, 1).off;
    }
    // Entry v, or 0 (slot 0, a valid source) for v >= hi.
    [[gnu::always_inline]] constexpr unsigned entry(int v, bool &back) const
    {
        back = false;
        if (v >= hi) return 0u;
        return base[offset(v, back)];
    }
};

// A cluster's segment counts (sc[s].x fast, sc[s].y back) as a SegList
// limited to [0, hi).
template <class Count2>
[[gnu::always_inline]] constexpr SegList seg_list(const unsigned *base, int segcap,
                                                  const Count2 (&sc)[NSEG], int hi)
{
    static_assert(NSEG == 4, "SegList holds four segments");
    SegList s{};
    s.base = base;
    s.segcap = segcap;
    s.hi = hi;
    s.g1 = sc[0].y;
    s.g2 = s.g1 + sc[1].y;
    s.g3 = s.g2 + sc[2].y;
    s.nback = s.g3 + sc[3].y;
    s.f1 = sc[0].x;
    s.f2 = s.f1 + sc[1].x;
    s.f3 = s.f2 + sc[2].x;
    return s;
}

}  // namespace yk
