// cycles.hip — conflicted 3-, 4- and 5-cycle search and triangulation
// (src/conflicted_cycles_cuda.cu:13-343, include/rama_utils.h:62-84,296-363).
//
// A conflicted cycle is one repulsive edge (v1, v2) closed by a path of
// attractive edges.  Each stage enumerates "expansions" and intersects two
// sorted attractive neighbour lists per expansion:
//   3-cycles: expansion = rep edge r;                 N(v1) ∩ N(v2)
//   4-cycles: expansion = (r, n1 ∈ N(v1));            N(n1) ∩ N(v2)
//   5-cycles: expansion = (r, n1 ∈ N(v1), n2 ∈ N(v2)); N(n1) ∩ N(n2), n1≠n2
// and writes 1, 2 or 3 triangles per common neighbour `mid`.
//
// Deterministic output order (the reference's atomic slot counter is
// replaced): a count pass stores each expansion's number of mids (uint8,
// saturating) and exact per-tile totals; a scan of the tile totals gives
// every tile its write offset; an emit pass revisits only expansions with
// mids and writes at enumeration position (rep edge, expansion, mid).  The
// triangle budget therefore keeps exactly the first `budget` writes.
#include <cstdlib>
#include <cstring>

#include "rama_internal.hpp"
#include "tuning.hpp"

namespace rama {

constexpr int kCycTile = kBlock;  // expansions per block

struct CycArgs {
    const int* off;      // attractive CSR offsets (n+1)
    const int* adj;      // attractive CSR neighbours
    const int* rv1;      // repulsive edge rows (v1 < v2)
    const int* rv2;
    const i64* xoff;     // per-rep exclusive expansion offsets (R+1)
    const int* tile_rep; // first rep edge of each tile
    i64 R;
    i64 X;               // number of expansions
    // Several instances packed in one graph (disjoint node ranges): nslot
    // maps a node to its instance; the count kernels add each instance's
    // attempted writes to slot_total, and the emit kernel places instance q's
    // writes at slot_out[q] + slot_base[q] + (global position - slot_first[q]),
    // keeping those below slot_budget[q].  All null for a single instance.
    const int* nslot = nullptr;
    i64* slot_total = nullptr;
    const i64* slot_first = nullptr;
    const i64* slot_base = nullptr;
    const i64* slot_budget = nullptr;
    const i64* slot_out = nullptr;
    // Counting in chunks of blocks (single instance): this launch's first
    // block, and where the blocks add their attempted writes ([0]) and the
    // end of the expansions they covered ([1], a maximum).
    unsigned blk0 = 0;
    i64* counted = nullptr;
    // diagnostic (RAMA_CYCLE_STATS): threads of the 5-cycle count, and those
    // that counted through a 2-hop table
    unsigned long long* stats = nullptr;
    // 5-cycle count: node ids fit the 2-hop tables' keys (below kT5MaxNode)
    int tables = 0;
    // 4-cycle count: wave-wide counting of long lists allowed
    int wide4 = 1;
};

// Diagnostic switches (tests force each path on the same inputs; every
// path gives the same counts and triangles):
//   RAMA_CYCLE_TABLES=0   no 2-hop tables (5-cycles) and no wave-wide sets
//                         (4-cycles): every expansion counted by its thread
//   RAMA_EMIT_FORM=dense | queue   the emit form, instead of the choice by
//                         the share of expansions with mids
static bool cycle_tables_on() {
    const char* e = std::getenv("RAMA_CYCLE_TABLES");
    return !(e && e[0] == '0' && e[1] == 0);
}
static int emit_form_forced() {  // 0: choose, 1: dense, 2: queue
    const char* e = std::getenv("RAMA_EMIT_FORM");
    if (!e) return 0;
    if (std::strcmp(e, "dense") == 0) return 1;
    if (std::strcmp(e, "queue") == 0) return 2;
    return 0;
}

// Adds the block's attempted writes to its counter line's [0] and raises
// its [1] to the end of its expansions.  The counters are spread over
// kCountedLines lines of 16 i64 (128 bytes): device atomics from every block
// to one address serialise; the host sums the [0]s and takes the largest
// [1].  Block-uniform call.
constexpr int kCountedLines = 16;
__device__ __forceinline__ void block_counted(i64* counted, i64 mine, i64 xend) {
    __shared__ i64 ws[kWaves];
    const i64 s = block_reduce(mine, OpSum(), ws);
    const i64 m = block_reduce(xend, OpMax(), ws);
    if (threadIdx.x == 0) {
        i64* line = counted + (blockIdx.x % kCountedLines) * 16;
        if (s) atomicAdd(reinterpret_cast<unsigned long long*>(line), (unsigned long long)s);
        atomicMax(reinterpret_cast<unsigned long long*>(line + 1), (unsigned long long)m);
    }
}

// Adds v to base[key] for every lane with `on`, one atomic per distinct key
// in the wave (lanes of a wave mostly share their instance).  Wave-uniform
// call.
__device__ __forceinline__ void wave_add_keyed(i64* base, int key, i64 v, bool on) {
    const int lane = lane_id();
    unsigned long long active = __ballot(on && v != 0);
    while (active) {
        const int leader = __builtin_ctzll(active);
        const int kl = __builtin_amdgcn_readlane(key, leader);
        const unsigned long long m = __ballot(key == kl) & active;
        i64 x = ((m >> lane) & 1ull) ? v : 0;
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) x += shfl_down64(x, d);
        if (lane == 0) atomicAdd(reinterpret_cast<unsigned long long*>(base + kl), (unsigned long long)x);
        active &= ~m;
    }
}

// Locate the rep edge owning expansion x: xoff[r] <= x < xoff[r+1], searching
// from the tile's first rep edge.
__device__ __forceinline__ i64 find_rep(const CycArgs& a, i64 x, i64 r0) {
    i64 lo = r0, hi = a.R - 1;
    // galloping upper bound then binary search
    i64 step = 1;
    while (lo + step <= hi && a.xoff[lo + step] <= x) {
        lo += step;
        step <<= 1;
    }
    hi = lo + step < a.R - 1 ? lo + step : a.R - 1;
    while (lo < hi) {
        const i64 mid = (lo + hi + 1) >> 1;
        if (a.xoff[mid] <= x) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Resolve expansion x to the two vertices whose neighbour lists intersect and
// the triangle writers' (v1, v2, n1, n2).  Returns false if the expansion is
// skipped (5-cycles with n1 == n2).
template <int S>
__device__ __forceinline__ bool resolve(const CycArgs& a, i64 x, i64 r0, int& v1, int& v2, int& n1, int& n2,
                                        int& qa, int& qb) {
    // 3-cycles: expansion x is rep edge x (no offsets, no tile reps)
    const i64 r = S == 3 ? x : find_rep(a, x, r0);
    v1 = a.rv1[r];
    v2 = a.rv2[r];
    const i64 loc = S == 3 ? 0 : x - a.xoff[r];
    if (S == 3) {
        qa = v1;
        qb = v2;
        n1 = n2 = -1;
        return true;
    } else if (S == 4) {
        n1 = a.adj[a.off[v1] + loc];
        n2 = -1;
        qa = n1;
        qb = v2;
        return true;
    } else {
        const int d1 = a.off[v1 + 1] - a.off[v1];
        const int l1 = (int)(loc % d1), l2 = (int)(loc / d1);
        n1 = a.adj[a.off[v1] + l1];
        n2 = a.adj[a.off[v2] + l2];
        qa = n1;
        qb = n2;
        return n1 != n2;
    }
}

// Number of common neighbours of qa and qb (merge walk over sorted lists).
__device__ __forceinline__ int count_common(const int* __restrict__ off, const int* __restrict__ adj, int qa, int qb) {
    int i = off[qa], ie = off[qa + 1], j = off[qb], je = off[qb + 1];
    int c = 0;
    while (i < ie && j < je) {
        const int x = adj[i], y = adj[j];
        c += (x == y);
        i += (x <= y);
        j += (y <= x);
    }
    return c;
}

// Pairs whose two lists hold more than kHeavy entries together are
// intersected by the whole wave instead of by their own lane: after
// contraction a supernode can have thousands of neighbours, and one lane
// walking two such lists would hold up its wave, its block and the launch.
constexpr int kHeavy = 96;

__device__ __forceinline__ int lower_in(const int* __restrict__ adj, int lo, int hi, int x) {
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (adj[mid] < x) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// Whether entry i of the sorted list [sb, se) (value x) has a partner in the
// sorted list [lb, le): its rank among the equal values of its own list is
// below their count in the other, so a value held a and b times matches
// min(a, b) times — the merge walk's rule.
__device__ __forceinline__ bool has_partner(const int* __restrict__ adj, int sb, int i, int x, int lb, int le) {
    const int rank = (i > sb && adj[i - 1] == x) ? i - lower_in(adj, sb, i, x) : 0;
    const int p = lower_in(adj, lb, le, x);
    if (p >= le || adj[p] != x) return false;
    const int cnt = (p + 1 < le && adj[p + 1] == x) ? lower_in(adj, p, le, x + 1) - p : 1;
    return rank < cnt;
}

// The shorter of the lists of qa and qb as [sb, se), the other as [lb, le).
__device__ __forceinline__ void short_long(const int* __restrict__ off, int qa, int qb, int& sb, int& se, int& lb,
                                           int& le) {
    const int ab = off[qa], ae = off[qa + 1], bb = off[qb], be = off[qb + 1];
    if (ae - ab <= be - bb) {
        sb = ab, se = ae, lb = bb, le = be;
    } else {
        sb = bb, se = be, lb = ab, le = ae;
    }
}

// Lane h's 64-bit value, as two 32-bit lane reads (h wave-uniform).
__device__ __forceinline__ i64 readlane64(i64 v, int h) {
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(v & 0xffffffffll), h);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)((unsigned long long)v >> 32), h);
    return (i64)(((unsigned long long)hi << 32) | lo);
}

// Lists at most this long are probed entry by entry into the other list by
// their own lane (binary searches) when the pair is heavy; pairs of two
// longer lists go to the whole wave.
constexpr int kProbe = 16;

// count_common by one lane from the shorter list [sb, se) into [lb, le).
__device__ __forceinline__ int probe_count(const int* __restrict__ adj, int sb, int se, int lb, int le) {
    int c = 0;
    for (int i = sb; i < se; ++i) c += has_partner(adj, sb, i, adj[i], lb, le) ? 1 : 0;
    return c;
}

// count_common of the lists [ia, ie) and [ib, ib + nb) by one lane: from
// the shorter list by binary searches when the pair is heavy and that list
// is short, else the merge walk.
__device__ __forceinline__ int lane_count_common(const int* __restrict__ adj, int ia, int ie, int ib, int nb) {
    const int na = ie - ia;
    if (na + nb > kHeavy && min(na, nb) <= kProbe)
        return na <= nb ? probe_count(adj, ia, ie, ib, ib + nb) : probe_count(adj, ib, ib + nb, ia, ie);
    int i = ia, j = ib, je = ib + nb, c = 0;
    while (i < ie && j < je) {
        const int x = adj[i], y = adj[j];
        c += (x == y);
        i += (x <= y);
        j += (y <= x);
    }
    return c;
}

// count_common by the whole wave (qa, qb wave-uniform): each lane takes
// entries of the shorter list and binary-searches the longer one.
__device__ __forceinline__ int wave_count_common(const int* __restrict__ off, const int* __restrict__ adj, int qa,
                                                 int qb) {
    int sb, se, lb, le;
    short_long(off, qa, qb, sb, se, lb, le);
    int c = 0;
    for (int i = sb + lane_id(); i < se; i += 64) c += has_partner(adj, sb, i, adj[i], lb, le) ? 1 : 0;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) c += __shfl_xor(c, d, 64);
    return c;
}

// Register-list intersection: one sorted list (at most kRegList entries)
// is held in registers, padded with -1, and the other is streamed from
// memory with independent loads.  The k-th copy of a value in the streamed
// list counts when the register list holds more than k copies, so a value
// occurring a and b times counts min(a, b) — the merge walk's rule.
constexpr int kRegList = 16;

__device__ __forceinline__ void load_reg_list(const int* __restrict__ adj, int b, int nb, int (&rb)[kRegList]) {
#pragma unroll
    for (int q = 0; q < kRegList; ++q) rb[q] = q < nb ? adj[b + q] : -1;
}

__device__ __forceinline__ int count_vs_regs(const int* __restrict__ adj, int ia, int ie, const int (&rb)[kRegList]) {
    int c = 0, prev = -1, occ = 0;
    for (int k = ia; k < ie; ++k) {
        const int x = adj[k];
        int m = 0;
#pragma unroll
        for (int q = 0; q < kRegList; ++q) m += (rb[q] == x);
        occ = x == prev ? occ + 1 : 0;
        prev = x;
        c += occ < m;
    }
    return c;
}

// count_vs_regs when the register list has no repeated value (rdup false):
// then a streamed value counts on its first copy only, when some register
// entry equals it.  The comparisons per streamed value become lane masks
// combined on the scalar unit, and the walk stops at the first value above
// the register list's largest entry (rmax; both lists are sorted).  Only
// the first KL register entries are compared (the entries past a list's
// length are -1 and match nothing).
template <int KL>
__device__ __forceinline__ int count_regs_masks(const int* __restrict__ adj, int ia, int ie,
                                                const int (&rb)[kRegList], int rmax) {
    const int lane = lane_id();
    int c = 0, prev = -1;
    // four streamed entries per round of loads (past the end reads as
    // above rmax), walked in order up to the first entry above rmax
    for (int k = ia; k < ie; k += 4) {
        int x[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) x[j] = k + j < ie ? adj[k + j] : 0x7fffffff;
        bool stop = false;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (x[j] > rmax) {
                stop = true;
                break;
            }
            unsigned long long hit = 0;
#pragma unroll
            for (int q = 0; q < KL; ++q) hit |= __ballot(rb[q] == x[j]);
            c += (int)((hit >> lane) & 1ull) & (int)(x[j] != prev);
            prev = x[j];
        }
        if (stop) break;
    }
    return c;
}

// Register entries the lanes about to intersect must compare (the longest
// of their lists: 8, 12 or 16), agreed by the active lanes of the wave.
__device__ __forceinline__ int reg_width(int nb) {
    if (__ballot(nb > 8) == 0ull) return 8;
    if (__ballot(nb > 12) == 0ull) return 12;
    return kRegList;
}

// kw: reg_width of the lanes that call it together.  With repeated
// register entries it falls back to count_vs_regs.
__device__ __forceinline__ int count_vs_regs_fast(const int* __restrict__ adj, int ia, int ie,
                                                  const int (&rb)[kRegList], int rmax, bool rdup, int kw) {
    if (rdup) return count_vs_regs(adj, ia, ie, rb);
    if (kw == 8) return count_regs_masks<8>(adj, ia, ie, rb, rmax);
    if (kw == 12) return count_regs_masks<12>(adj, ia, ie, rb, rmax);
    return count_regs_masks<kRegList>(adj, ia, ie, rb, rmax);
}

// Largest entry and whether a value repeats, for a register list of nb
// sorted entries.
__device__ __forceinline__ void reg_list_props(const int (&rb)[kRegList], int nb, int& rmax, bool& rdup) {
    rmax = -1;
    rdup = false;
#pragma unroll
    for (int q = 0; q < kRegList; ++q) {
        if (q < nb) rmax = rb[q];
        if (q > 0 && q < nb) rdup |= rb[q] == rb[q - 1];
    }
}

// Adds `v` to the total of the tile holding expansion x, flushing the
// running sum whenever the thread's expansions cross into the next tile.
struct TileAcc {
    i64 tile = -1;
    i64 sum = 0;
    __device__ __forceinline__ void add(i64* tot, i64 x, i64 v) {
        const i64 t = x / kCycTile;
        if (t != tile) {
            flush(tot);
            tile = t;
        }
        sum += v;
    }
    __device__ __forceinline__ void flush(i64* tot) {
        if (tile >= 0 && sum != 0)
            atomicAdd(reinterpret_cast<unsigned long long*>(tot + tile), (unsigned long long)sum);
        sum = 0;
    }
};

// 3-cycles: one thread per repulsive edge r = expansion r; N(v2) in
// registers, N(v1) streamed.  Tile totals by block reduction.
__global__ __launch_bounds__(kBlock) void k_cycle_count3(CycArgs a, unsigned char* __restrict__ cnt8,
                                                         i64* __restrict__ tile_tot) {
    __shared__ i64 ws[kWaves];
    const unsigned blk = xcd_block();
    const i64 r = (i64)blk * kCycTile + threadIdx.x;
    int c = 0;
    int v1 = 0, v2 = 0;
    bool heavy = false;
    if (r < a.R) {
        v1 = a.rv1[r], v2 = a.rv2[r];
        const int b2 = a.off[v2], n2 = a.off[v2 + 1] - b2;
        const int d1 = a.off[v1 + 1] - a.off[v1];
        if (d1 + n2 > kHeavy) {
            if (min(d1, n2) <= kProbe) {
                int sb, se, lb, le;
                short_long(a.off, v1, v2, sb, se, lb, le);
                c = probe_count(a.adj, sb, se, lb, le);
            } else {
                heavy = true;  // counted by the whole wave below
            }
        } else if (n2 <= kRegList) {
            int rb[kRegList];
            load_reg_list(a.adj, b2, n2, rb);
            int rmax;
            bool rdup;
            reg_list_props(rb, n2, rmax, rdup);
            c = count_vs_regs_fast(a.adj, a.off[v1], a.off[v1 + 1], rb, rmax, rdup, reg_width(n2));
        } else {
            c = count_common(a.off, a.adj, v1, v2);
        }
    }
    for (unsigned long long hm = __ballot(heavy); hm; hm &= hm - 1) {
        const int h = __builtin_ctzll(hm);
        const int s = wave_count_common(a.off, a.adj, __builtin_amdgcn_readlane(v1, h),
                                        __builtin_amdgcn_readlane(v2, h));
        if (lane_id() == h) c = s;
    }
    if (r < a.R) cnt8[r] = (unsigned char)(c < 255 ? c : 255);
    if (a.nslot) wave_add_keyed(a.slot_total, r < a.R ? a.nslot[a.rv1[r]] : 0, (i64)c, r < a.R);
    const i64 tot = block_reduce((i64)c, OpSum(), ws);
    if (threadIdx.x == 0) tile_tot[blk] = tot;
}

// 4-cycles: one thread per repulsive edge; its expansions n1 ∈ N(v1) are
// xoff[r] .. xoff[r+1]-1 in order, each intersecting N(n1) with N(v2)
// (held in registers).  tile_tot must be zero on entry.
// Rep edges whose v1 has at least kWideRep neighbours, or whose v2 has more
// than kRegList, are counted by the whole wave, one rep edge at a time: the
// wave puts N(v2) in an LDS hash set (kSetSlots slots, lists of at most
// kSetMax entries) and lane l takes expansions l, l + 64, ..., streaming
// N(n1) and probing each entry.  After contraction supernodes have dozens
// of neighbours: one lane per rep edge would walk long merge walks serially
// while its wave waits.  A repeated value in either list (the min(a, b)
// rule) or a longer N(v2) keeps the merge walk for that expansion.
constexpr int kWideRep = 48;
constexpr int kSetSlots = 256;
constexpr int kSetMax = 128;

__device__ __forceinline__ int set_hash(int m) { return (int)(((unsigned)m * 2654435761u) >> 24); }

__global__ __launch_bounds__(kBlock) void k_cycle_count4(CycArgs a, unsigned char* __restrict__ cnt8,
                                                         i64* __restrict__ tile_tot) {
    __shared__ int wset[kWaves][4 * kSetSlots];
    const int w = threadIdx.x >> 6;
    const i64 r = (i64)(a.blk0 + xcd_block()) * kBlock + threadIdx.x;
    const bool on = r < a.R;
    i64 mine = 0, xend = 0;
    int v1 = 0, v2 = 0, b1 = 0, d1 = 0, b2 = 0, n2 = 0;
    i64 x0 = 0;
    if (on) {
        v1 = a.rv1[r];
        v2 = a.rv2[r];
        b1 = a.off[v1];
        d1 = a.off[v1 + 1] - b1;
        b2 = a.off[v2];
        n2 = a.off[v2 + 1] - b2;
        x0 = a.xoff[r];
        xend = x0 + d1;
    }
    const bool wide = on && a.wide4 && (d1 >= kWideRep || n2 > kRegList);
    if (on && !wide) {
        int rb[kRegList];
        const bool regs = n2 <= kRegList;
        int rmax = -1, kw = kRegList;
        bool rdup = false;
        if (regs) {
            load_reg_list(a.adj, b2, n2, rb);
            reg_list_props(rb, n2, rmax, rdup);
            kw = reg_width(n2);
        }
        TileAcc acc;
        // the next n1's list bounds are loaded while this one is walked
        int ia = 0, ie = 0;
        if (d1 > 0) {
            const int n1 = a.adj[b1];
            ia = a.off[n1];
            ie = a.off[n1 + 1];
        }
        for (int l = 0; l < d1; ++l) {
            int na = 0, ne = 0;
            if (l + 1 < d1) {
                const int nn = a.adj[b1 + l + 1];
                na = a.off[nn];
                ne = a.off[nn + 1];
            }
            const int c = regs ? count_vs_regs_fast(a.adj, ia, ie, rb, rmax, rdup, kw)
                               : lane_count_common(a.adj, ia, ie, b2, n2);
            cnt8[x0 + l] = (unsigned char)(c < 255 ? c : 255);
            acc.add(tile_tot, x0 + l, 2 * (i64)c);
            mine += 2 * (i64)c;
            ia = na;
            ie = ne;
        }
        acc.flush(tile_tot);
    }
    // wide rep edges: the wave works as four groups of 16 lanes, group g
    // taking the wide rep edges of lanes 16g .. 16g + 15 one after another
    // (four rep edges in flight per wave); the attempted writes are credited
    // to the rep's own lane (its instance, its block total)
    const int lane = lane_id(), g = lane >> 4, gl = lane & 15;
    const unsigned long long wm = __ballot(wide);
    unsigned gm = (unsigned)((wm >> (16 * g)) & 0xffffull);  // this group's wide lanes
    int niter = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) niter = max(niter, __popcll((wm >> (16 * q)) & 0xffffull));
    int* gset = &wset[w][g * kSetSlots];
    const unsigned long long gmask = 0xffffull << (16 * g);
    for (int it = 0; it < niter; ++it) {
        const bool gact = gm != 0u;  // uniform within the group
        const int h = gact ? 16 * g + __builtin_ctz(gm) : lane;
        if (gact) gm &= gm - 1u;
        const int hb1 = __shfl(b1, h, 64), hd1 = gact ? __shfl(d1, h, 64) : 0;
        const int hb2 = __shfl(b2, h, 64), hn2 = gact ? __shfl(n2, h, 64) : 0;
        const i64 hx0 = __shfl(x0, h, 64);
        // N(v2) into the group's set (a repeated value, or too long a list:
        // merge walks instead)
        bool use_set = gact && hn2 <= kSetMax;
        int vmax = -1;
        bool dup = false;
        if (use_set) {
            for (int k = gl; k < kSetSlots; k += 16) gset[k] = -1;
        }
        __builtin_amdgcn_wave_barrier();
        if (use_set) {
            for (int k = gl; k < hn2; k += 16) {
                const int m = a.adj[hb2 + k];
                dup |= k > 0 && a.adj[hb2 + k - 1] == m;
                int hs = set_hash(m);
                while (true) {
                    const int old = atomicCAS(&gset[hs], -1, m);
                    if (old == -1 || old == m) break;
                    hs = (hs + 1) & (kSetSlots - 1);
                }
            }
            vmax = hn2 > 0 ? a.adj[hb2 + hn2 - 1] : -1;
        }
        use_set = use_set && (__ballot(dup) & gmask) == 0ull;
        __builtin_amdgcn_wave_barrier();
        TileAcc acc;
        i64 part = 0;
        for (int l = gl; l < hd1; l += 16) {
            const int n1 = a.adj[hb1 + l];
            const int ia = a.off[n1], ie = a.off[n1 + 1];
            int c = 0;
            bool merge = !use_set;
            if (use_set) {
                // four entries per round of loads (past the end reads as
                // above vmax), walked in order up to the first above vmax
                int prev = -1;
                bool stop = false;
                for (int k = ia; k < ie && !stop; k += 4) {
                    int xs[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) xs[u] = k + u < ie ? a.adj[k + u] : 0x7fffffff;
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int x = xs[u];
                        if (stop || x > vmax) {
                            stop = true;
                            continue;
                        }
                        if (x == prev) {  // N(n1) repeats a value: the merge walk counts it
                            merge = true;
                            stop = true;
                            continue;
                        }
                        prev = x;
                        int hs = set_hash(x);
                        while (true) {
                            const int key = gset[hs];
                            if (key == x) {
                                ++c;
                                break;
                            }
                            if (key == -1) break;
                            hs = (hs + 1) & (kSetSlots - 1);
                        }
                    }
                }
            }
            if (merge) c = lane_count_common(a.adj, ia, ie, hb2, hn2);
            cnt8[hx0 + l] = (unsigned char)(c < 255 ? c : 255);
            acc.add(tile_tot, hx0 + l, 2 * (i64)c);
            part += 2 * (i64)c;
        }
        acc.flush(tile_tot);
#pragma unroll
        for (int d = 8; d >= 1; d >>= 1) part += __shfl_xor(part, d, 64);  // within the group
        if (gact && lane == h) mine += part;
        __builtin_amdgcn_wave_barrier();
    }
    if (a.nslot) wave_add_keyed(a.slot_total, on ? a.nslot[v1] : 0, mine, on);
    if (a.counted) block_counted(a.counted, mine, xend);
}

// 5-cycles: one thread per (repulsive edge r, n2 = l2-th neighbour of v2);
// its expansions (n1 = l1-th neighbour of v1, n2) are xoff[r] + l2*d1 + l1
// for l1 = 0 .. d1-1, each intersecting N(n1) with N(n2); n1 == n2 is
// skipped.  Threads of rep r are yoff[r] .. yoff[r+1]-1; ytile_rep gives
// each block's first rep.  tile_tot must be zero on entry.
//
// Two ways to count one thread's d1 expansions, with identical counts:
//  * 2-hop table (the common case): the block builds, for each distinct v1
//    among its threads, an LDS hash table of v1's 2-hop neighbourhood,
//    mid -> bitmask of the l1 whose list N(n1) holds mid.  The thread then
//    walks N(n2) once, looks every mid up and adds its mask to bit-sliced
//    per-l1 counters, instead of streaming every N(n1) against N(n2).  Valid
//    when no list involved holds a value twice (then |N(n1) ∩ N(n2)| is the
//    merge walk's min(a, b) count), d1 <= 32 and the 2-hop lists fit.
//  * per expansion (otherwise): N(n2) in registers (or the merge walk /
//    probes for long lists), each N(n1) streamed against it.
constexpr int kT5Tables = 16;      // 2-hop tables (distinct v1 runs) per block
constexpr int kT5SlotBits = 11;
constexpr int kT5Slots = 1 << kT5SlotBits;  // slots of the block's shared hash table
constexpr int kT5MaxFill = 1536;   // entries of all tables together at most this (load <= 3/4)
constexpr int kT5MaxD1 = 32;       // d(v1) at most this (32-bit masks)
constexpr int kT5RepWin = 256;     // rep edges a block locates its threads among in LDS
constexpr int kT5MaxNode = 1 << 27;  // node ids below this: (mid, run) fits one key

// The tables of a block share one open-addressing hash table of keys
// mid * kT5Tables + run.
__device__ __forceinline__ int t5_hash(int key) {
    return (int)(((unsigned)key * 2654435761u) >> (32 - kT5SlotBits));
}

// The per-expansion form for one thread (see above).
__device__ __forceinline__ void count5_direct(const CycArgs& a, int b1, int d1, int n2, int bq, int nq, i64 x0,
                                              unsigned char* __restrict__ cnt8, i64* __restrict__ tile_tot,
                                              i64& mine) {
    int rb[kRegList];
    const bool regs = nq <= kRegList;
    int rmax = -1, kw = kRegList;
    bool rdup = false;
    if (regs) {
        load_reg_list(a.adj, bq, nq, rb);
        reg_list_props(rb, nq, rmax, rdup);
        kw = reg_width(nq);
    }
    TileAcc acc;
    // the next n1 and its list bounds are loaded while this one is walked
    int n1 = d1 > 0 ? a.adj[b1] : 0;
    int ia = d1 > 0 ? a.off[n1] : 0, ie = d1 > 0 ? a.off[n1 + 1] : 0;
    for (int l = 0; l < d1; ++l) {
        const int nn = l + 1 < d1 ? a.adj[b1 + l + 1] : 0;
        const int na = l + 1 < d1 ? a.off[nn] : 0, ne = l + 1 < d1 ? a.off[nn + 1] : 0;
        int c = 0;
        if (n1 != n2)
            c = regs ? count_vs_regs_fast(a.adj, ia, ie, rb, rmax, rdup, kw)
                     : lane_count_common(a.adj, ia, ie, bq, nq);
        cnt8[x0 + l] = (unsigned char)(c < 255 ? c : 255);
        acc.add(tile_tot, x0 + l, 3 * (i64)c);
        mine += 3 * (i64)c;
        n1 = nn;
        ia = na;
        ie = ne;
    }
    acc.flush(tile_tot);
}

__global__ __launch_bounds__(kBlock) void k_cycle_count5(CycArgs a, const i64* __restrict__ yoff,
                                                         const int* __restrict__ ytile_rep, i64 Y,
                                                         unsigned char* __restrict__ cnt8,
                                                         i64* __restrict__ tile_tot) {
    __shared__ int tkey[kT5Slots];
    __shared__ unsigned tmask[kT5Slots];
    __shared__ int tv1[kT5Tables];
    __shared__ int tok[kT5Tables];
    __shared__ int tfill[kT5Tables];
    __shared__ int sv1[kBlock];
    __shared__ int wruns[kWaves];
    const unsigned blk = a.blk0 + xcd_block();
    const int tid = threadIdx.x, lane = lane_id(), w = tid >> 6;
    const i64 y = (i64)blk * kCycTile + tid;
    const bool on = y < Y;
    i64 mine = 0, xend = 0;
    int slot = 0;
    int v1 = -1, b1 = 0, d1 = 0, n2 = 0, bq = 0, nq = 0;
    i64 x0 = 0;
    // rep owning each thread: the block's threads belong to reps r0, r0 + 1,
    // ...; their thread offsets are loaded once (coalesced) and every thread
    // marks where its rep starts, so a block max-scan (kept as a running max
    // over the marks) hands each thread its rep.  Blocks whose threads span
    // more than kT5RepWin reps fall back to a galloping search.
    const i64 y0 = (i64)blk * kCycTile;
    const i64 r0 = y0 < Y ? (i64)ytile_rep[blk] : 0;
    sv1[tid] = -1;
    __syncthreads();
    bool wide = false;
    for (int k = tid; k < kT5RepWin; k += kBlock) {
        const i64 r = r0 + k;
        if (r < a.R) {
            const i64 ys = yoff[r];
            if (ys >= y0 && ys < y0 + kCycTile && yoff[r + 1] > ys) sv1[ys - y0] = k;
        }
    }
    if (y0 < Y) {
        const i64 rl = r0 + kT5RepWin - 1;
        wide = rl < a.R && yoff[rl + 1] < (y0 + kCycTile < Y ? y0 + kCycTile : Y);
    }
    __syncthreads();
    int mark = sv1[tid];
    // inclusive max-scan of the marks over the block (first thread: its rep
    // r0 starts before the block, mark 0)
    if (tid == 0 && mark < 0) mark = 0;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int t = __shfl_up(mark, d, 64);
        if (lane >= d) mark = max(mark, t);
    }
    __shared__ int wmax[kWaves];
    if (lane == 63) wmax[w] = mark;
    __syncthreads();
    for (int q = 0; q < w; ++q) mark = max(mark, wmax[q]);
    __syncthreads();
    if (on) {
        i64 r = r0 + mark;
        if (wide) {
            // rep owning thread y (galloping from the block's first rep)
            i64 lo = r0, hi = a.R - 1, step = 1;
            while (lo + step <= hi && yoff[lo + step] <= y) {
                lo += step;
                step <<= 1;
            }
            hi = lo + step < a.R - 1 ? lo + step : a.R - 1;
            while (lo < hi) {
                const i64 mid = (lo + hi + 1) >> 1;
                if (yoff[mid] <= y) lo = mid; else hi = mid - 1;
            }
            r = lo;
        }
        v1 = a.rv1[r];
        const int v2 = a.rv2[r];
        const int l2 = (int)(y - yoff[r]);
        b1 = a.off[v1];
        d1 = a.off[v1 + 1] - b1;
        n2 = a.adj[a.off[v2] + l2];
        bq = a.off[n2];
        nq = a.off[n2 + 1] - bq;
        x0 = a.xoff[r] + (i64)l2 * d1;
        if (a.nslot) slot = a.nslot[v1];
    }
    // runs of equal v1 over the block's threads (rep edges are in row order,
    // so a run is usually all threads of one v1); run k < kT5Tables gets
    // table k
    sv1[tid] = v1;
    for (int k = tid; k < kT5Slots; k += kBlock) {
        tkey[k] = -1;
        tmask[k] = 0u;
    }
    __syncthreads();
    const bool start = on && (tid == 0 || sv1[tid - 1] != v1);
    const unsigned long long sm = __ballot(start);
    if (lane == 0) wruns[w] = __popcll(sm);
    __syncthreads();
    int before = 0, runs = 0;
    for (int q = 0; q < kWaves; ++q) {
        before += q < w ? wruns[q] : 0;
        runs += wruns[q];
    }
    // this thread's run index (its run starts at or before it)
    const int run = before + __popcll(sm & ((2ull << lane) - 1ull)) - 1;
    if (start && run < kT5Tables) {
        tv1[run] = v1;
        tok[run] = 1;
    }
    const int ntab = runs < kT5Tables ? runs : kT5Tables;
    __syncthreads();
    // build, in three steps: each wave sizes its tables (w, w + kWaves, ...:
    // lane l holds the bounds of the list of v1's l-th neighbour n1); the
    // tables that fit the shared capacity in run order are kept; each wave
    // inserts its kept tables' (mid, l) pairs, flattened over its 64 lanes
    constexpr int kPerWave = kT5Tables / kWaves;
    int s0[kPerWave], s1[kPerWave];
#pragma unroll
    for (int i = 0; i < kPerWave; ++i) {
        const int k = w + i * kWaves;
        s0[i] = s1[i] = 0;
        if (k >= ntab) continue;
        const int v = tv1[k];
        const int bv = a.off[v], dv = a.off[v + 1] - bv;
        if (dv <= kT5MaxD1 && lane < dv) {
            const int n1 = a.adj[bv + lane];
            s0[i] = a.off[n1];
            s1[i] = a.off[n1 + 1];
        }
        int fill = s1[i] - s0[i];
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) fill += __shfl_xor(fill, d, 64);
        if (lane == 0) tfill[k] = dv <= kT5MaxD1 && a.tables ? fill : kT5MaxFill + 1;
    }
    __syncthreads();
    if (tid == 0) {
        int used = 0;
        for (int k = 0; k < ntab; ++k) {
            const bool fits = used + tfill[k] <= kT5MaxFill;
            tok[k] = fits ? 1 : 0;
            used += fits ? tfill[k] : 0;
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kPerWave; ++i) {
        const int k = w + i * kWaves;
        if (k >= ntab || !tok[k]) continue;  // wave-uniform
        bool dup = false;
        const int fill = tfill[k];
        const int len = s1[i] - s0[i];
        int excl = len;  // where list l starts among the table's pairs
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int t = __shfl_up(excl, d, 64);
            if (lane >= d) excl += t;
        }
        excl -= len;
        const int dv = a.off[tv1[k] + 1] - a.off[tv1[k]];
        for (int f0 = 0; f0 < fill; f0 += 64) {  // wave-uniform: every lane takes part in the shuffles
            const int f = f0 + lane;
            int l = 0;
            for (int j = 1; j < dv; ++j) l = __builtin_amdgcn_readlane(excl, j) <= f ? j : l;
            const int q = __shfl(s0[i], l, 64) + (f - __shfl(excl, l, 64));
            if (f >= fill) continue;
            const int key = a.adj[q] * kT5Tables + k;
            const unsigned bit = 1u << l;
            int h = t5_hash(key);
            while (true) {
                const int old = atomicCAS(&tkey[h], -1, key);
                if (old == -1 || old == key) {
                    dup |= (atomicOr(&tmask[h], bit) & bit) != 0u;
                    break;
                }
                h = (h + 1) & (kT5Slots - 1);
            }
        }
        // a list that repeats a value disables its table
        if (__ballot(dup) != 0ull && lane == 0) tok[k] = 0;
    }
    __syncthreads();
    if (on) {
        const bool table = run < ntab && tok[run] && nq <= 255;
        bool direct = !table;
        if (table) {
            // per-l1 counts of N(n2)'s mids, bit-sliced over 8 planes
            unsigned p[8] = {0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u};
            int prev = -1;
            // four entries of N(n2) per round of loads
            for (int q0 = bq; q0 < bq + nq && !direct; q0 += 4) {
                int ms[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) ms[u] = q0 + u < bq + nq ? a.adj[q0 + u] : -1;
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    if (direct || q0 + u >= bq + nq) continue;
                    const int m = ms[u];
                    if (m == prev) {  // N(n2) repeats a value: the min(a, b) rule needs the direct form
                        direct = true;
                        continue;
                    }
                    prev = m;
                    const int want = m * kT5Tables + run;
                    int h = t5_hash(want);
                    unsigned mk = 0u;
                    while (true) {
                        const int key = tkey[h];
                        if (key == want) {
                            mk = tmask[h];
                            break;
                        }
                        if (key == -1) break;
                        h = (h + 1) & (kT5Slots - 1);
                    }
#pragma unroll
                    for (int b = 0; b < 8; ++b) {
                        const unsigned carry = p[b] & mk;
                        p[b] ^= mk;
                        mk = carry;
                    }
                }
            }
            if (!direct) {
                unsigned nz = 0u;
#pragma unroll
                for (int b = 0; b < 8; ++b) nz |= p[b];
                TileAcc acc;
                for (int l = 0; l < d1; ++l) {
                    int c = 0;
                    if ((nz >> l) & 1u) {
#pragma unroll
                        for (int b = 0; b < 8; ++b) c |= (int)((p[b] >> l) & 1u) << b;
                        if (a.adj[b1 + l] == n2) c = 0;
                    }
                    cnt8[x0 + l] = (unsigned char)(c < 255 ? c : 255);
                    acc.add(tile_tot, x0 + l, 3 * (i64)c);
                    mine += 3 * (i64)c;
                }
                acc.flush(tile_tot);
            }
        }
        if (direct) count5_direct(a, b1, d1, n2, bq, nq, x0, cnt8, tile_tot, mine);
        xend = x0 + d1;
        if (a.stats) {
            atomicAdd(a.stats, 1ull);
            if (!direct) atomicAdd(a.stats + 1, 1ull);
        }
    }
    if (a.nslot) wave_add_keyed(a.slot_total, slot, mine, on);
    if (a.counted) block_counted(a.counted, mine, xend);
}

__device__ __forceinline__ void put_tri(int* t1, int* t2, int* t3, i64 g, i64 budget, int a, int b, int c) {
    if (g >= budget) return;
    const int lo = min(a, min(b, c));
    const int hi = max(a, max(b, c));
    const int md = max(min(a, b), min(max(a, b), c));
    t1[g] = lo;
    t2[g] = md;
    t3[g] = hi;
}

// One queued expansion per lane (have: the lane holds one), counted c > 0
// with its first write at `pos` in enumeration order: resolve it, move to the
// instance's triangle region on packed graphs, then write its triangles by
// the merge walk, by probing the short list into the long one, or (two long
// lists) with the whole wave.  Wave-uniform call.
template <int S>
__device__ __forceinline__ void emit_item(const CycArgs& a, bool have, i64 x, i64 pos, i64 budget, int* t1, int* t2,
                                          int* t3) {
    int v1 = 0, v2 = 0, n1 = 0, n2 = 0, qa = 0, qb = 0;
    i64 mypos = pos, tofs = 0;  // tofs: start of the instance's triangle region (packed)
    if (have) {
        resolve<S>(a, x, S == 3 ? 0 : a.tile_rep[x / kCycTile], v1, v2, n1, n2, qa, qb);
        if (a.nslot) {
            // instance q's own position, budget and region
            const int q = a.nslot[v1];
            mypos = a.slot_base[q] + (pos - a.slot_first[q]);
            budget = a.slot_budget[q];
            tofs = a.slot_out[q];
        }
    }
    const bool writes = have && mypos < budget;
    const int i0 = writes ? a.off[qa] : 0, ie = writes ? a.off[qa + 1] : 0;
    const int j0 = writes ? a.off[qb] : 0, je = writes ? a.off[qb + 1] : 0;
    const bool big = writes && (ie - i0) + (je - j0) > kHeavy;
    const bool heavy = big && min(ie - i0, je - j0) > kProbe;
    if (big && !heavy) {
        // one short list: probe its entries into the other, in order
        int sb, se, lb, le;
        short_long(a.off, qa, qb, sb, se, lb, le);
        int* u1 = t1 + tofs;
        int* u2 = t2 + tofs;
        int* u3 = t3 + tofs;
        i64 g = mypos;
        for (int i = sb; i < se && g < budget; ++i) {
            const int mid = a.adj[i];
            if (!has_partner(a.adj, sb, i, mid, lb, le)) continue;
            if (S == 3) {
                put_tri(u1, u2, u3, g++, budget, v1, v2, mid);
            } else if (S == 4) {
                put_tri(u1, u2, u3, g++, budget, v1, v2, n1);
                put_tri(u1, u2, u3, g++, budget, v2, n1, mid);
            } else {
                put_tri(u1, u2, u3, g++, budget, v1, v2, n1);
                put_tri(u1, u2, u3, g++, budget, v2, n1, mid);
                put_tri(u1, u2, u3, g++, budget, v2, mid, n2);
            }
        }
    }
    // heavy items: the whole wave walks the shorter list in order, finds each
    // entry's partner by binary search, and places the k-th common neighbour
    // at mypos + k * (S - 2) as the merge walk would
    for (unsigned long long hm = __ballot(heavy); hm; hm &= hm - 1) {
        const int h = __builtin_ctzll(hm);
        const int hqa = __builtin_amdgcn_readlane(qa, h), hqb = __builtin_amdgcn_readlane(qb, h);
        const int hv1 = __builtin_amdgcn_readlane(v1, h), hv2 = __builtin_amdgcn_readlane(v2, h);
        const int hn1 = __builtin_amdgcn_readlane(n1, h), hn2 = __builtin_amdgcn_readlane(n2, h);
        const i64 hpos = readlane64(mypos, h);
        const i64 hbud = readlane64(budget, h);
        const i64 hofs = readlane64(tofs, h);
        int sb, se, lb, le;
        short_long(a.off, hqa, hqb, sb, se, lb, le);
        const unsigned long long below = (1ull << lane_id()) - 1ull;
        i64 done = 0;
        for (int base = sb; base < se && hpos + done * (S - 2) < hbud; base += 64) {
            const int i = base + lane_id();
            int mid = 0;
            bool m = false;
            if (i < se) {
                mid = a.adj[i];
                m = has_partner(a.adj, sb, i, mid, lb, le);
            }
            const unsigned long long bm = __ballot(m);
            if (m) {
                i64 g = hpos + (done + __popcll(bm & below)) * (S - 2);
                int* u1 = t1 + hofs;
                int* u2 = t2 + hofs;
                int* u3 = t3 + hofs;
                if (S == 3) {
                    put_tri(u1, u2, u3, g, hbud, hv1, hv2, mid);
                } else if (S == 4) {
                    put_tri(u1, u2, u3, g, hbud, hv1, hv2, hn1);
                    put_tri(u1, u2, u3, g + 1, hbud, hv2, hn1, mid);
                } else {
                    put_tri(u1, u2, u3, g, hbud, hv1, hv2, hn1);
                    put_tri(u1, u2, u3, g + 1, hbud, hv2, hn1, mid);
                    put_tri(u1, u2, u3, g + 2, hbud, hv2, mid, hn2);
                }
            }
            done += __popcll(bm);
        }
    }
    if (!writes || big) return;
    t1 += tofs;
    t2 += tofs;
    t3 += tofs;
    int i = i0, j = j0;
    i64 g = mypos;
    while (i < ie && j < je && g < budget) {
        const int p = a.adj[i], q = a.adj[j];
        if (p == q) {
            const int mid = p;
            if (S == 3) {
                put_tri(t1, t2, t3, g++, budget, v1, v2, mid);
            } else if (S == 4) {
                put_tri(t1, t2, t3, g++, budget, v1, v2, n1);
                put_tri(t1, t2, t3, g++, budget, v2, n1, mid);
            } else {
                put_tri(t1, t2, t3, g++, budget, v1, v2, n1);
                put_tri(t1, t2, t3, g++, budget, v2, n1, mid);
                put_tri(t1, t2, t3, g++, budget, v2, mid, n2);
            }
        }
        i += (p <= q);
        j += (q <= p);
    }
}

__device__ __forceinline__ i64 wave_excl_scan64(i64 v, i64* total) {
    i64 x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const i64 t = __shfl_up(x, d, 64);
        if (lane_id() >= d) x += t;
    }
    *total = readlane64(x, 63);
    return x - v;
}

// Dense emit: one thread per expansion (block = one count tile), write
// positions by a block scan of the counts.  Used when a large share of the
// expansions have mids (every lane then has one to write).
template <int S>
__global__ __launch_bounds__(kBlock) void k_cycle_emit_dense(CycArgs a, const unsigned char* __restrict__ cnt8,
                                                             const i64* __restrict__ tile_off, i64 base, i64 budget,
                                                             int* __restrict__ t1, int* __restrict__ t2,
                                                             int* __restrict__ t3) {
    __shared__ i64 ws[kWaves];
    // plain block order: tiles past the budget exit at once, and keeping the
    // working tiles spread over all XCDs balances them
    const unsigned blk = blockIdx.x;
    const bool packed = a.nslot != nullptr;
    if (tile_off[blk + 1] == tile_off[blk]) return;  // uniform: the tile writes nothing
    const i64 tstart = (packed ? 0 : base) + tile_off[blk];
    if (!packed && tstart >= budget) return;  // uniform per block
    const i64 x = (i64)blk * kCycTile + threadIdx.x;
    int c = x < a.X ? (int)cnt8[x] : 0;
    {
        // a saturated count (255): count again (by the wave when both lists
        // are long)
        int v1 = 0, v2 = 0, n1 = 0, n2 = 0, qa = 0, qb = 0;
        bool wave = false;
        if (c == 255) {
            resolve<S>(a, x, S == 3 ? 0 : a.tile_rep[blk], v1, v2, n1, n2, qa, qb);
            int sb, se, lb, le;
            short_long(a.off, qa, qb, sb, se, lb, le);
            if (se - sb <= kProbe) c = probe_count(a.adj, sb, se, lb, le);
            else wave = true;
        }
        for (unsigned long long hm = __ballot(wave); hm; hm &= hm - 1) {
            const int h = __builtin_ctzll(hm);
            const int s = wave_count_common(a.off, a.adj, __builtin_amdgcn_readlane(qa, h),
                                            __builtin_amdgcn_readlane(qb, h));
            if (lane_id() == h) c = s;
        }
    }
    i64 tot;
    const i64 pos = tstart + block_excl_scan((i64)c * (S - 2), ws, &tot);
    emit_item<S>(a, c > 0, x, pos, budget, t1, t2, t3);
}

// Sparse emit: one wave per block walks a contiguous range of count tiles in
// order.  Each tile's 256 counts are read four per lane (expansion order is
// lane-major), a wave scan turns them into write positions (tile start +
// the counts before), and the expansions with mids are queued in LDS; every
// 64 queued expansions are written by the 64 lanes at once (emit_item).
// Most expansions have no mid, so lanes stay busy instead of idling beside
// the few that write.  Saturated counts (255) are counted again first.
constexpr int kEmitQueue = 320;  // at most 63 waiting + 256 from one tile
constexpr i64 kEmitTilesPerWave = 8;  // count tiles a queued-emit wave walks (latency-bound loop: keep it short)

template <int S>
__global__ __launch_bounds__(64) void k_cycle_emit(CycArgs a, const unsigned char* __restrict__ cnt8,
                                                   const i64* __restrict__ tile_off, i64 ntiles, i64 per_block,
                                                   i64 base, i64 budget, int* __restrict__ t1, int* __restrict__ t2,
                                                   int* __restrict__ t3) {
    __shared__ i64 qx[kEmitQueue];
    __shared__ i64 qp[kEmitQueue];
    const int lane = lane_id();
    const bool packed = a.nslot != nullptr;
    const i64 tb = (i64)blockIdx.x * per_block;
    const i64 te = tb + per_block < ntiles ? tb + per_block : ntiles;
    int qn = 0;  // queued expansions (wave-uniform)
    for (i64 t = tb; t < te; ++t) {
        const i64 o0 = tile_off[t], o1 = tile_off[t + 1];
        if (o1 == o0) continue;  // the tile writes nothing (uniform)
        const i64 tstart = (packed ? 0 : base) + o0;
        if (!packed && tstart >= budget) break;  // later tiles start later still
        const i64 x0 = t * kCycTile + 4 * lane;
        int c[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) c[k] = x0 + k < a.X ? (int)cnt8[x0 + k] : 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const bool sat = c[k] == 255;
            if (__ballot(sat) == 0ull) continue;
            // the count saturated its byte: count again (by the wave when
            // both lists are long)
            int v1 = 0, v2 = 0, n1 = 0, n2 = 0, qa = 0, qb = 0;
            bool wave = false;
            if (sat) {
                resolve<S>(a, x0 + k, S == 3 ? 0 : a.tile_rep[t], v1, v2, n1, n2, qa, qb);
                int sb, se, lb, le;
                short_long(a.off, qa, qb, sb, se, lb, le);
                if (se - sb <= kProbe) c[k] = probe_count(a.adj, sb, se, lb, le);
                else wave = true;
            }
            for (unsigned long long hm = __ballot(wave); hm; hm &= hm - 1) {
                const int h = __builtin_ctzll(hm);
                const int s = wave_count_common(a.off, a.adj, __builtin_amdgcn_readlane(qa, h),
                                                __builtin_amdgcn_readlane(qb, h));
                if (lane == h) c[k] = s;
            }
        }
        i64 tot;
        i64 p = tstart + wave_excl_scan64((i64)(c[0] + c[1] + c[2] + c[3]) * (S - 2), &tot);
        const int nz = (c[0] > 0) + (c[1] > 0) + (c[2] > 0) + (c[3] > 0);
        int zx = nz;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int u = __shfl_up(zx, d, 64);
            if (lane >= d) zx += u;
        }
        const int ztot = __builtin_amdgcn_readlane(zx, 63);
        int slot = qn + zx - nz;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (c[k] > 0) {
                qx[slot] = x0 + k;
                qp[slot] = p;
                ++slot;
            }
            p += (i64)c[k] * (S - 2);
        }
        qn += ztot;
        __syncthreads();
        while (qn >= 64) {
            qn -= 64;
            const i64 x = qx[qn + lane], pos = qp[qn + lane];
            __syncthreads();
            emit_item<S>(a, true, x, pos, budget, t1, t2, t3);
        }
    }
    __syncthreads();
    const bool have = lane < qn;
    emit_item<S>(a, have, have ? qx[lane] : 0, have ? qp[lane] : 0, budget, t1, t2, t3);
}

// The emit over `ntiles` count tiles with W attempted writes in all: the
// per-expansion form when at least one expansion in kEmitDenseShare has a
// mid (W counts S - 2 writes per mid), else the queued form.
constexpr i64 kEmitDenseShare = 4;
template <int S>
static void launch_emit(Ctx& ctx, const CycArgs& a, const unsigned char* cnt8, const i64* tile_off, i64 ntiles,
                        i64 W, i64 base, i64 budget, int* t1, int* t2, int* t3) {
    const int forced = emit_form_forced();
    if (forced == 1 || (forced == 0 && W * kEmitDenseShare >= (i64)(S - 2) * a.X)) {
        k_cycle_emit_dense<S><<<(unsigned)ntiles, kBlock, 0, ctx.stream>>>(a, cnt8, tile_off, base, budget, t1, t2,
                                                                           t3);
    } else {
        const i64 per = kEmitTilesPerWave;
        k_cycle_emit<S><<<(unsigned)((ntiles + per - 1) / per), 64, 0, ctx.stream>>>(a, cnt8, tile_off, ntiles, per,
                                                                                     base, budget, t1, t2, t3);
    }
    RAMA_LAUNCHED();
}

// tile_rep[t] = rep edge containing the first expansion of tile t.
__global__ void k_tile_rep(const i64* __restrict__ xoff, i64 R, int* __restrict__ tile_rep) {
    const i64 r = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R) return;
    const i64 b = xoff[r], e = xoff[r + 1];
    if (e <= b) return;
    for (i64 t = (b + kCycTile - 1) / kCycTile; t * kCycTile < e; ++t) tile_rep[t] = (int)r;
}

// Work of one count launch for the profiler (only while it is on): the
// expansions the launch covers and the neighbour-list entries a merge walk
// of each expansion's two lists reads, deg(a) + deg(b).  S = 3 and 4: the
// launch covers rep edges [lo, hi); S = 5: threads [lo, hi), thread k of
// rep edge r being (r, the (k - yoff[r])-th neighbour n2 of v2), whose
// expansions are (n1, n2) for every n1 of v1 (n1 = n2 included: the count
// skips it after loading both bounds).  nsum[v] = sum of the degrees of v's
// neighbours.
template <int S>
static void count_work(Ctx& ctx, const Csr& G, const int* rv1, const int* rv2, i64 R, const i64* yoff, const i64* nsum,
                       i64 lo, i64 hi, double& x, double& entries) {
    const int* OFF = G.off.get();
    const int* ADJ = G.adj.get();
    x = entries = 0.0;
    if (hi <= lo) return;
    if (S == 3) {
        x = (double)(hi - lo);
        entries = (double)reduce<i64>(
            ctx,
            [=] __device__(i64 k) {
                const int v = rv1[lo + k], u = rv2[lo + k];
                return (i64)(OFF[v + 1] - OFF[v]) + (i64)(OFF[u + 1] - OFF[u]);
            },
            hi - lo, (i64)0, OpSum());
    } else if (S == 4) {
        x = (double)reduce<i64>(
            ctx, [=] __device__(i64 k) { const int v = rv1[lo + k]; return (i64)(OFF[v + 1] - OFF[v]); }, hi - lo,
            (i64)0, OpSum());
        entries = (double)reduce<i64>(
            ctx,
            [=] __device__(i64 k) {
                const int v = rv1[lo + k], u = rv2[lo + k];
                return nsum[v] + (i64)(OFF[v + 1] - OFF[v]) * (i64)(OFF[u + 1] - OFF[u]);
            },
            hi - lo, (i64)0, OpSum());
    } else {
        x = (double)reduce<i64>(
            ctx,
            [=] __device__(i64 r) {
                const i64 a = yoff[r] > lo ? yoff[r] : lo, b = yoff[r + 1] < hi ? yoff[r + 1] : hi;
                const int v = rv1[r];
                return a < b ? (i64)(OFF[v + 1] - OFF[v]) * (b - a) : (i64)0;
            },
            R, (i64)0, OpSum());
        entries = (double)reduce<i64>(
            ctx,
            [=] __device__(i64 r) {
                const i64 a = yoff[r] > lo ? yoff[r] : lo, b = yoff[r + 1] < hi ? yoff[r + 1] : hi;
                if (a >= b) return (i64)0;
                const int v = rv1[r], u = rv2[r];
                const i64 d1 = OFF[v + 1] - OFF[v], base = (i64)OFF[u] - yoff[r];
                i64 s = 0;
                for (i64 k = a; k < b; ++k) {
                    const int n2 = ADJ[base + k];
                    s += nsum[v] + d1 * (i64)(OFF[n2 + 1] - OFF[n2]);
                }
                return s;
            },
            R, (i64)0, OpSum());
    }
}

// Runs one stage over all expansions; returns the number of attempted writes.
template <int S>
static i64 run_stage(Ctx& ctx, const Csr& G, const int* rv1, const int* rv2, i64 R, i64 base, i64 budget,
                     Triangles& T) {
    const int* OFF = G.off.get();
    DBuf<i64> xoff;  // 3-cycles: expansion = rep edge, no offsets
    if (S > 3) xoff = ctx.buf<i64>(R + 1);
    i64* xm = ctx.mirror<i64>();
    if (S == 4) {
        scan_exclusive<i64>(ctx, [=] __device__(i64 r) { const int v = rv1[r]; return (i64)(OFF[v + 1] - OFF[v]); },
                            xoff.get(), R, xm);
    } else if (S == 5) {
        scan_exclusive<i64>(
            ctx,
            [=] __device__(i64 r) {
                const int v = rv1[r], u = rv2[r];
                return (i64)(OFF[v + 1] - OFF[v]) * (i64)(OFF[u + 1] - OFF[u]);
            },
            xoff.get(), R, xm);
    }
    i64 X = R;
    if (S > 3) ctx.read_mirror(xoff.get() + R, xm, &X);
    if (ctx.verbose && S > 3) std::printf("%d-cycles: number of expansions: %lld\n", S, (long long)X);
    if (X == 0) return 0;
    const i64 ntiles = (X + kCycTile - 1) / kCycTile;
    DBuf<int> tile_rep;
    if (S > 3) {
        tile_rep = ctx.buf<int>(ntiles);
        k_tile_rep<<<grid_for(R), kBlock, 0, ctx.stream>>>(xoff.get(), R, tile_rep.get());
        RAMA_LAUNCHED();
    }
    CycArgs a{OFF, G.adj.get(), rv1, rv2, S > 3 ? xoff.get() : nullptr, S > 3 ? tile_rep.get() : nullptr, R, X};
    a.tables = G.n < kT5MaxNode && cycle_tables_on() ? 1 : 0;
    a.wide4 = cycle_tables_on() ? 1 : 0;
    DBuf<unsigned char> cnt8 = ctx.buf<unsigned char>(X);
    DBuf<i64> tile = ctx.buf<i64>(ntiles + 1);
    RAMA_REQUIRE(ntiles < ((i64)1 << 31), "too many cycle expansions");
    // profiling only: each count launch's expansions and list entries, and
    // their merge-walk bytes (4 per list entry, 8 of list bounds per expansion)
    const bool work_on = prof::wants(S == 3 ? "cycles3_count" : (S == 4 ? "cycles4_count" : "cycles5_count"));
    DBuf<i64> nsum;
    if (work_on && S > 3) {
        nsum = ctx.buf<i64>(G.n);
        i64* NS = nsum.get();
        const int* ADJ = G.adj.get();
        for_each(ctx, G.n, [=] __device__(i64 v) {
            i64 s = 0;
            for (int k = OFF[v]; k < OFF[v + 1]; ++k) s += OFF[ADJ[k] + 1] - OFF[ADJ[k]];
            NS[v] = s;
        });
    }
    double wx = 0.0, wy = 0.0;
    if (S == 3) {
        if (work_on) count_work<3>(ctx, G, rv1, rv2, R, nullptr, nullptr, 0, R, wx, wy);
        prof::Scope ps(ctx, "cycles3_count", 4.0 * wy + 8.0 * wx, wx, wy);
        k_cycle_count3<<<(unsigned)ntiles, kBlock, 0, ctx.stream>>>(a, cnt8.get(), tile.get());
        RAMA_LAUNCHED();
    }
    // Stages 4 and 5 count in chunks of blocks, each a prefix of the
    // expansion order, and stop once the attempted writes reach the budget
    // left (budget - base): expansions past that point write nothing, so
    // the triangles, the stage's outcome (total >= budget left) and every
    // later decision are those of a full count.  Only the returned total is
    // then a lower bound of the full one; verbose runs count everything, as
    // they print it.
    i64 tiles_used = ntiles;
    if (S > 3) {
        ctx.zero(tile.get(), ntiles);
        DBuf<i64> cnt = ctx.buf<i64>(kCountedLines * 16);
        ctx.zero(cnt.get(), kCountedLines * 16);
        a.counted = cnt.get();
        DBuf<i64> yoff;
        DBuf<int> ytile_rep;
        i64 Y = 0;
        unsigned nblk = 0;
        if (S == 4) {
            nblk = grid_for(R);
        } else {
            // thread space: one thread per (rep edge, neighbour of v2)
            yoff = ctx.buf<i64>(R + 1);
            i64* ym = ctx.mirror<i64>();
            scan_exclusive<i64>(ctx,
                                [=] __device__(i64 r) { const int u = rv2[r]; return (i64)(OFF[u + 1] - OFF[u]); },
                                yoff.get(), R, ym);
            ctx.read_mirror(yoff.get() + R, ym, &Y);
            if (Y > 0) {
                const i64 ytiles = (Y + kCycTile - 1) / kCycTile;
                RAMA_REQUIRE(ytiles < ((i64)1 << 31), "too many cycle expansions");
                ytile_rep = ctx.buf<int>(ytiles);
                k_tile_rep<<<grid_for(R), kBlock, 0, ctx.stream>>>(yoff.get(), R, ytile_rep.get());
                RAMA_LAUNCHED();
                nblk = (unsigned)ytiles;
            }
        }
        static const bool cyc_stats = std::getenv("RAMA_CYCLE_STATS") != nullptr;
        DBuf<unsigned long long> st;
        if (cyc_stats && S == 5) {
            st = ctx.buf<unsigned long long>(2);
            ctx.zero(st.get(), 2);
            a.stats = st.get();
        }
        const bool early = !ctx.verbose && budget > base;
        const i64 need = budget - base;
        // the first chunk: a sixteenth of the blocks (a count that runs to
        // the end pays this launch's tail once; smaller first chunks cost D
        // mode on the 2D grid 5 ms, larger ones overshoot budgets)
        unsigned done = 0, chunk = early ? std::max<unsigned>(nblk / 16, 256u) : nblk;
        i64 hc[2] = {0, X};
        while (done < nblk) {
            const unsigned n = std::min(chunk, nblk - done);
            a.blk0 = done;
            if (work_on) {
                const i64 lo = (i64)done * kBlock, hi = std::min<i64>((i64)(done + n) * kBlock, S == 4 ? R : Y);
                count_work<S>(ctx, G, rv1, rv2, R, S == 5 ? yoff.get() : nullptr, nsum.get(), lo, hi, wx, wy);
            }
            if (S == 4) {
                prof::Scope ps(ctx, "cycles4_count", 4.0 * wy + 8.0 * wx, wx, wy);
                k_cycle_count4<<<n, kBlock, 0, ctx.stream>>>(a, cnt8.get(), tile.get());
                RAMA_LAUNCHED();
            } else {
                prof::Scope ps(ctx, "cycles5_count", 4.0 * wy + 8.0 * wx, wx, wy);
                k_cycle_count5<<<n, kBlock, 0, ctx.stream>>>(a, yoff.get(), ytile_rep.get(), Y, cnt8.get(),
                                                             tile.get());
                RAMA_LAUNCHED();
            }
            done += n;
            if (!early || done >= nblk) break;
            {
                i64 lines[kCountedLines * 16];
                ctx.read(cnt.get(), lines, kCountedLines * 16);
                hc[0] = 0;
                hc[1] = 0;
                for (int k = 0; k < kCountedLines; ++k) {
                    hc[0] += lines[k * 16];
                    hc[1] = std::max(hc[1], lines[k * 16 + 1]);
                }
            }
            if (hc[0] >= need) break;
            // next chunk: cycle_chunk_pct percent (default 125) of the
            // blocks that the rate so far says reach the budget (all the
            // rest when it says they will not)
            const double per_block = (double)hc[0] / (double)done;
            const double grow = 0.01 * (double)tune::get(tune::CycleChunkPct, X);
            const double want = per_block > 0.0 ? grow * (double)(need - hc[0]) / per_block + 64.0 : 1e30;
            chunk = want >= (double)(nblk - done) ? nblk - done : (unsigned)want;
        }
        if (a.stats) {
            unsigned long long h[2];
            ctx.read(a.stats, h, 2);
            std::fprintf(stderr, "5-cycle count: %llu threads, %llu through 2-hop tables (%d blocks)\n", h[0], h[1], (int)done);
            a.stats = nullptr;
        }
        if (done < nblk) {
            // counted expansions [0, x_end): emit the tiles they touch; the
            // rest of the last tile contributes nothing
            const i64 x_end = hc[1];
            tiles_used = (x_end + kCycTile - 1) / kCycTile;
            const i64 tail = std::min<i64>(tiles_used * kCycTile, X) - x_end;
            if (tail > 0) ctx.zero(cnt8.get() + x_end, tail);
        }
    }
    i64* wm = ctx.mirror<i64>();
    scan_exclusive_arr<i64>(ctx, tile.get(), tile.get(), tiles_used, wm);
    i64 W;
    ctx.read_mirror(tile.get() + tiles_used, wm, &W);
    if (W > 0 && base < budget) {
        prof::Scope ps(ctx, S == 3 ? "cycles3_emit" : (S == 4 ? "cycles4_emit" : "cycles5_emit"), 0.0);
        launch_emit<S>(ctx, a, cnt8.get(), tile.get(), tiles_used, W, base, budget, T.t1.get(), T.t2.get(),
                       T.t3.get());
    }
    return W;
}

// One stage over several packed instances.  Instance q's writes go to its
// triangle region (d_out[q] onward) at positions continuing from d_base[q],
// below d_budget[q]; h_total[q] receives its attempted writes (the stage's
// return value in the single-instance form).
template <int S>
void run_stage_packed(Ctx& ctx, const Csr& G, const int* rv1, const int* rv2, i64 R, const int* nslot, int Q,
                      const i64* d_base, const i64* d_budget, const i64* d_out, int* t1, int* t2, int* t3,
                      i64* h_total) {
    for (int q = 0; q < Q; ++q) h_total[q] = 0;
    if (R == 0) return;
    const int* OFF = G.off.get();
    DBuf<i64> xoff;  // 3-cycles: expansion = rep edge, no offsets
    if (S > 3) xoff = ctx.buf<i64>(R + 1);
    i64* xm = ctx.mirror<i64>();
    if (S == 4) {
        scan_exclusive<i64>(ctx, [=] __device__(i64 r) { const int v = rv1[r]; return (i64)(OFF[v + 1] - OFF[v]); },
                            xoff.get(), R, xm);
    } else if (S == 5) {
        scan_exclusive<i64>(
            ctx,
            [=] __device__(i64 r) {
                const int v = rv1[r], u = rv2[r];
                return (i64)(OFF[v + 1] - OFF[v]) * (i64)(OFF[u + 1] - OFF[u]);
            },
            xoff.get(), R, xm);
    }
    i64 X = R;
    if (S > 3) ctx.read_mirror(xoff.get() + R, xm, &X);
    if (X == 0) return;
    const i64 ntiles = (X + kCycTile - 1) / kCycTile;
    RAMA_REQUIRE(ntiles < ((i64)1 << 31), "too many cycle expansions");
    DBuf<int> tile_rep;
    if (S > 3) {
        tile_rep = ctx.buf<int>(ntiles);
        k_tile_rep<<<grid_for(R), kBlock, 0, ctx.stream>>>(xoff.get(), R, tile_rep.get());
        RAMA_LAUNCHED();
    }
    DBuf<i64> slot = ctx.buf<i64>(2 * (i64)Q + 1);  // [0, Q) totals, [Q, 2Q] their exclusive scan
    i64* TOT = slot.get();
    ctx.zero(TOT, Q);
    CycArgs a{OFF, G.adj.get(), rv1, rv2, S > 3 ? xoff.get() : nullptr, S > 3 ? tile_rep.get() : nullptr, R, X};
    a.tables = G.n < kT5MaxNode && cycle_tables_on() ? 1 : 0;
    a.wide4 = cycle_tables_on() ? 1 : 0;
    a.nslot = nslot;
    a.slot_total = TOT;
    a.slot_first = TOT + Q;
    a.slot_base = d_base;
    a.slot_budget = d_budget;
    a.slot_out = d_out;
    DBuf<unsigned char> cnt8 = ctx.buf<unsigned char>(X);
    DBuf<i64> tile = ctx.buf<i64>(ntiles + 1);
    if (S == 3) {
        k_cycle_count3<<<(unsigned)ntiles, kBlock, 0, ctx.stream>>>(a, cnt8.get(), tile.get());
        RAMA_LAUNCHED();
    } else if (S == 4) {
        ctx.zero(tile.get(), ntiles);
        k_cycle_count4<<<grid_for(R), kBlock, 0, ctx.stream>>>(a, cnt8.get(), tile.get());
        RAMA_LAUNCHED();
    } else {
        DBuf<i64> yoff = ctx.buf<i64>(R + 1);
        scan_exclusive<i64>(ctx, [=] __device__(i64 r) { const int u = rv2[r]; return (i64)(OFF[u + 1] - OFF[u]); },
                            yoff.get(), R);
        const i64 Y = ctx.read1(yoff.get() + R);
        ctx.zero(tile.get(), ntiles);
        if (Y > 0) {
            const i64 ytiles = (Y + kCycTile - 1) / kCycTile;
            RAMA_REQUIRE(ytiles < ((i64)1 << 31), "too many cycle expansions");
            DBuf<int> ytile_rep = ctx.buf<int>(ytiles);
            k_tile_rep<<<grid_for(R), kBlock, 0, ctx.stream>>>(yoff.get(), R, ytile_rep.get());
            RAMA_LAUNCHED();
            k_cycle_count5<<<(unsigned)ytiles, kBlock, 0, ctx.stream>>>(a, yoff.get(), ytile_rep.get(), Y,
                                                                        cnt8.get(), tile.get());
            RAMA_LAUNCHED();
        }
    }
    scan_exclusive_arr<i64>(ctx, tile.get(), tile.get(), ntiles);
    scan_exclusive_arr<i64>(ctx, TOT, TOT + Q, Q);
    ctx.d2h(h_total, TOT, Q);
    i64 W = 0;
    for (int q = 0; q < Q; ++q) W += h_total[q];
    if (W > 0) {
        launch_emit<S>(ctx, a, cnt8.get(), tile.get(), ntiles, W, 0, 0, t1, t2, t3);
    }
}
template void run_stage_packed<3>(Ctx&, const Csr&, const int*, const int*, i64, const int*, int, const i64*,
                                  const i64*, const i64*, int*, int*, int*, i64*);
template void run_stage_packed<4>(Ctx&, const Csr&, const int*, const int*, i64, const int*, int, const i64*,
                                  const i64*, const i64*, int*, int*, int*, i64*);
template void run_stage_packed<5>(Ctx&, const Csr&, const int*, const int*, i64, const int*, int, const i64*,
                                  const i64*, const i64*, int*, int*, int*, i64*);

// rearrange_triangles / normalize_triangles (include/rama_utils.h:325-363)
// on the first `valid` slots: drop triangles with a repeated vertex, sort
// lexicographically, remove duplicates.  Returns the new count.
i64 rearrange_triangles(Ctx& ctx, Triangles& T, i64 valid, int n_nodes) {
    if (valid <= 0) return 0;
    int* A = T.t1.get();
    int* B = T.t2.get();
    int* C = T.t3.get();
    const int kb = bits_for(n_nodes > 1 ? n_nodes - 1 : 1);
    DBuf<int> cnt = ctx.buf<int>(1);  // unique triangles
    int* CNT = cnt.get();
    DBuf<u64> keys = ctx.buf<u64>(valid);
    DBuf<u32> vals = ctx.buf<u32>(valid);
    u64* K = keys.get();
    u32* V = vals.get();
    if (3 * kb <= 64) {
        // key = the ordered triple (write_triangle already orders it; kept
        // for foreign input); a triangle with a repeated vertex gets the
        // all-ones key of the sort width, which no real triangle has
        // (first < second < third), so it sorts last and is cut below
        const u64 pad = (1ull << (3 * kb)) - 1ull;
        for_each(ctx, valid, [=] __device__(i64 t) {
            const int a = A[t], b = B[t], c = C[t];
            const bool degen = a == b || a == c || b == c;
            const u32 lo = (u32)min(a, min(b, c)), md = (u32)max(min(a, b), min(max(a, b), c)),
                      hi = (u32)max(a, max(b, c));
            K[t] = degen ? pad : ((u64)lo << (2 * kb)) | ((u64)md << kb) | (u64)hi;
            V[t] = (u32)t;
        });
        {
            prof::Site site_("triangles");
            radix_sort_pairs(ctx, keys, vals, valid, 3 * kb);
        }
        K = keys.get();  // an odd pass count hands back the scratch buffers
        const u64 m1 = (1ull << kb) - 1ull;
        // keep the first of each run of equal keys; decode the triple in place
        select_into(
            ctx, valid, [=] __device__(i64 q) { return K[q] != pad && (q == 0 || K[q] != K[q - 1]); },
            [=] __device__(i64 q, int o) {
                const u64 k = K[q];
                A[o] = (int)(k >> (2 * kb));
                B[o] = (int)((k >> kb) & m1);
                C[o] = (int)(k & m1);
            },
            CNT, ctx.mirror<int>());
        int h;
        ctx.read_mirror(CNT, ctx.mirror<int>(), &h);
        return h;
    }
    // ids too wide for one key: two stable passes, by third id and then by
    // (first, second), on the ordered triples kept in S1..S3 (-1: repeated vertex)
    DBuf<int> s1 = ctx.buf<int>(valid), s2 = ctx.buf<int>(valid), s3 = ctx.buf<int>(valid);
    int* S1 = s1.get();
    int* S2 = s2.get();
    int* S3 = s3.get();
    const u64 pad1 = (1ull << kb) - 1ull, pad2 = (1ull << (2 * kb)) - 1ull;
    for_each(ctx, valid, [=] __device__(i64 t) {
        const int a = A[t], b = B[t], c = C[t];
        const bool degen = a == b || a == c || b == c;
        S1[t] = degen ? -1 : min(a, min(b, c));
        S2[t] = degen ? -1 : max(min(a, b), min(max(a, b), c));
        S3[t] = degen ? -1 : max(a, max(b, c));
        K[t] = degen ? pad1 : (u64)(u32)S3[t];
        V[t] = (u32)t;
    });
    prof::Site site_("triangles");
    radix_sort_pairs(ctx, keys, vals, valid, kb);
    K = keys.get();
    V = vals.get();
    for_each(ctx, valid, [=] __device__(i64 q) {
        const u32 o = V[q];
        K[q] = S1[o] < 0 ? pad2 : ((u64)(u32)S1[o] << kb) | (u64)(u32)S2[o];
    });
    radix_sort_pairs(ctx, keys, vals, valid, 2 * kb);
    K = keys.get();
    V = vals.get();
    select_into(
        ctx, valid,
        [=] __device__(i64 q) {
            const u32 a = V[q];
            if (S1[a] < 0) return false;
            if (q == 0) return true;
            const u32 b = V[q - 1];
            return S1[a] != S1[b] || S2[a] != S2[b] || S3[a] != S3[b];
        },
        [=] __device__(i64 q, int o) {
            const u32 a = V[q];
            A[o] = S1[a];
            B[o] = S2[a];
            C[o] = S3[a];
        },
        CNT, ctx.mirror<int>());
    int h;
    ctx.read_mirror(CNT, ctx.mirror<int>(), &h);
    return h;
}

void normalize_triangles(Ctx& ctx, Triangles& T) { T.n = rearrange_triangles(ctx, T, T.n, 1 << 30); }

// conflicted_cycles_cuda (src/conflicted_cycles_cuda.cu:222-343), followed by
// the normalisation the message-passing constructor applies.
Triangles conflicted_cycles(Ctx& ctx, const DCoo& A, int max_len, float factor, float tol) {
    Triangles T;
    if (max_len > 5) throw Error(RAMA_ERR_INVALID, "max_cycle_length should be <= 5. Received: " + std::to_string(max_len));
    if (max_len < 3 || A.nnz == 0) return T;
    const float* W = A.val.get();
    const int N = A.max_dim();
    // determine_threshold (:170-174): thr = min(c) * tol; attractive c > -thr
    // (is_positive_edge), repulsive c < thr (is_neg_edge).  The minimum and
    // both class sizes come back in one read.
    struct Sizes {
        float mn;
        float pad;
        u64 counts;  // attractive count << 32 | repulsive count
    };
    DBuf<Sizes> sz = ctx.buf<Sizes>(1);
    Sizes* SZ = sz.get();
    Sizes* SM = ctx.mirror<Sizes>();
    reduce_to<float>(ctx, [=] __device__(i64 e) { return W[e]; }, A.nnz, 3.402823466e38f, OpMin(), &SZ->mn,
                     SM ? &SM->mn : nullptr);
    reduce_to<u64>(
        ctx,
        [=] __device__(i64 e) {
            const float t = SZ->mn * tol;
            return ((u64)(W[e] > -1.0f * t) << 32) | (u64)(W[e] < t);
        },
        A.nnz, (u64)0, OpSum(), &SZ->counts, SM ? &SM->counts : nullptr);
    Sizes h;
    ctx.read_mirror(SZ, SM, &h);
    const float thr = h.mn * tol;
    const i64 npos = (i64)(h.counts >> 32);
    const i64 R = (i64)(h.counts & 0xffffffffull);
    if (npos == 0 || R == 0) {
        ctx.last_cycles_empty = true;
        return T;
    }
    Csr G;
    if (A.sym || ctx.last_cycles_empty) {
        // the contraction that made A left its full symmetric CSR, or the
        // last search found nothing, so this dual pass likely leaves A as it
        // is: build that CSR once (the matching reuses it) and take the
        // attractive part out of it
        if (!A.sym) A.sym = std::make_shared<Csr>(symmetric_csr(ctx, A.row.get(), A.col.get(), W, A.nnz, N, false, 0.0f));
        G = csr_filter(ctx, *A.sym, -1.0f * thr, 2 * npos);
    } else {
        G = symmetric_csr(ctx, A.row.get(), A.col.get(), W, A.nnz, N, true, -1.0f * thr, npos);
    }
    DBuf<int> rv1 = ctx.buf<int>(R), rv2 = ctx.buf<int>(R);
    {
        int* X1 = rv1.get();
        int* X2 = rv2.get();
        const int* AR = A.row.get();
        const int* AC = A.col.get();
        select_into_loaded(
            ctx, A.nnz, [=] __device__(i64 e) { return W[e] < thr; },
            [=] __device__(i64 e) { return make_int2(AR[e], AC[e]); },
            [=] __device__(i64, int q, int2 rc) {
                X1[q] = rc.x;
                X2[q] = rc.y;
            },
            reinterpret_cast<int*>(&SZ->pad));
    }
    const i64 budget = (i64)(int)(factor * (float)R);
    T.t1 = ctx.buf<int>(budget > 0 ? budget : 1);
    T.t2 = ctx.buf<int>(budget > 0 ? budget : 1);
    T.t3 = ctx.buf<int>(budget > 0 ? budget : 1);
    i64 count = run_stage<3>(ctx, G, rv1.get(), rv2.get(), R, 0, budget, T);
    if (ctx.verbose) std::printf("3-cycles: found # of triangles: %lld, budget: %lld\n", (long long)count, (long long)budget);
    if (max_len >= 4 && count < budget) {
        count = rearrange_triangles(ctx, T, count, N);
        count += run_stage<4>(ctx, G, rv1.get(), rv2.get(), R, count, budget, T);
        if (ctx.verbose) std::printf("4-cycles: found # of triangles: %lld, budget: %lld\n", (long long)count, (long long)budget);
    }
    if (max_len >= 5 && count < budget) {
        count = rearrange_triangles(ctx, T, count, N);
        count += run_stage<5>(ctx, G, rv1.get(), rv2.get(), R, count, budget, T);
        if (ctx.verbose) std::printf("5-cycles: found # of triangles: %lld, budget: %lld\n", (long long)count, (long long)budget);
    }
    T.raw = count < budget ? count : budget;
    T.n = rearrange_triangles(ctx, T, T.raw, N);
    T.max_id = N - 1;
    ctx.last_cycles_empty = T.raw == 0;
    return T;
}

}  // namespace rama
