// prims.cuh — data-parallel building blocks written for CDNA (wave64):
// wave/block scans and reductions, device-wide exclusive scan, deterministic
// reductions, stream compaction and a stable LSD radix sort of
// (u64 key, u32 value) pairs.  Everything runs on the context's stream.
#pragma once

#include <cstring>
#include <type_traits>

#include "common.hpp"
#include "profiler.hpp"
#include "tuning.hpp"

namespace rama {

constexpr int kBlock = 256;  // 4 waves of 64
constexpr int kWaves = kBlock / 64;

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Logical block index that gives each XCD a contiguous range of blocks.
// Workgroups are dispatched round-robin over the 8 XCDs (block b runs on
// XCD b % 8), and each XCD has its own L2; with this remap neighbouring
// blocks — which touch neighbouring edges and nodes — share an L2.  A
// bijection on [0, gridDim.x).
__device__ __forceinline__ unsigned xcd_block() {
    constexpr unsigned kXcd = 8;
    const unsigned b = blockIdx.x, g = gridDim.x;
    if (g < 2 * kXcd) return b;
    const unsigned q = g / kXcd, rem = g % kXcd, x = b % kXcd, k = b / kXcd;
    return x * q + (x < rem ? x : rem) + k;
}

template <class T>
__device__ __forceinline__ T shfl_up64(T x, int d) {
    return __shfl_up(x, d, 64);
}
template <>
__device__ __forceinline__ long shfl_up64<long>(long x, int d) {
    return (long)__shfl_up((long long)x, d, 64);
}
template <>
__device__ __forceinline__ unsigned long shfl_up64<unsigned long>(unsigned long x, int d) {
    return (unsigned long)__shfl_up((unsigned long long)x, d, 64);
}
template <class T>
__device__ __forceinline__ T shfl_down64(T x, int d) {
    return __shfl_down(x, d, 64);
}
template <>
__device__ __forceinline__ long shfl_down64<long>(long x, int d) {
    return (long)__shfl_down((long long)x, d, 64);
}

// Inclusive wave scan (sum).
template <class T>
__device__ __forceinline__ T wave_incl_scan(T x) {
    const int lane = lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        T y = shfl_up64(x, d);
        if (lane >= d) x += y;
    }
    return x;
}

// Exclusive block scan over kBlock threads; *total receives the block sum.
// Uses the caller-provided shared array `ws` of kWaves elements.
template <class T>
__device__ __forceinline__ T block_excl_scan(T x, T* ws, T* total) {
    const int lane = lane_id(), w = threadIdx.x >> 6;
    T inc = wave_incl_scan(x);
    if (lane == 63) ws[w] = inc;
    __syncthreads();
    T off = T(0), tot = T(0);
#pragma unroll
    for (int k = 0; k < kWaves; ++k) {
        const T s = ws[k];
        if (k < w) off += s;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return off + inc - x;
}

// Block reduction with an associative op (fixed tree order).
template <class T, class Op>
__device__ __forceinline__ T block_reduce(T x, Op op, T* ws) {
    const int lane = lane_id(), w = threadIdx.x >> 6;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) x = op(x, shfl_down64(x, d));
    if (lane == 0) ws[w] = x;
    __syncthreads();
    T r = ws[0];
#pragma unroll
    for (int k = 1; k < kWaves; ++k) r = op(r, ws[k]);
    __syncthreads();
    return r;
}

struct OpSum {
    template <class T>
    __device__ __forceinline__ T operator()(T a, T b) const { return a + b; }
};
struct OpMin {
    template <class T>
    __device__ __forceinline__ T operator()(T a, T b) const { return b < a ? b : a; }
};
struct OpMax {
    template <class T>
    __device__ __forceinline__ T operator()(T a, T b) const { return a < b ? b : a; }
};

// ------------------------------------------------------------------ scan
// Status words of the onesweep sort's decoupled look-back pack
// (epoch:16 | flag:2 | value:46); a word from another epoch reads as "not
// ready".  A look-back that spins too long raises ctx's error flag (checked
// by scan_check) instead of hanging.
constexpr u64 kScanValMask = (1ull << 46) - 1;
constexpr long kScanSpinLimit = 1l << 24;

__device__ __forceinline__ u64 scan_word(u32 epoch, u32 flag, u64 v) {
    return ((u64)epoch << 48) | ((u64)flag << 46) | (v & kScanValMask);
}

// Exclusive scan of non-negative integer sums, reduce-then-scan over a
// bounded grid: block b owns a contiguous chunk of whole tiles; the first
// launch stores each chunk's sum, the second adds the sums of the chunks
// before its own (a block reduction over at most kScanMaxBlocks partials)
// and scans its chunk tile by tile carrying the running prefix.  No atomics
// and no waiting on other blocks.  One chunk: the second launch only.
constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;
constexpr unsigned kScanMaxBlocks = 8192;  // the tuning table picks up to this many

// Items are evaluated and emitted in striped order (item k*kBlock + t of the
// tile by thread t), so the loads inside f and the stores inside out are
// coalesced across each wave; the scan itself runs on the blocked order
// (thread t owns items t*kScanItems .. +kScanItems-1) after a transpose
// through LDS (row stride kScanItems + 1 keeps both orders conflict-free).
constexpr int kScanPad = kScanItems + 1;
__device__ __forceinline__ int scan_slot(int j) { return (j / kScanItems) * kScanPad + j % kScanItems; }

template <class T, class F>
__global__ __launch_bounds__(kBlock) void k_scan_up(F f, i64 n, i64 chunk, T* __restrict__ partials) {
    __shared__ T ws[kWaves];
    const i64 c0 = (i64)blockIdx.x * chunk;
    const i64 c1 = c0 + chunk < n ? c0 + chunk : n;
    T s = T(0);
    i64 i = c0 + threadIdx.x;
    for (; i + (kScanItems - 1) * kBlock < c1; i += kScanItems * kBlock) {
        T x[kScanItems];
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) x[k] = (T)f(i + k * kBlock);
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) s += x[k];
    }
    for (; i < c1; i += kBlock) s += (T)f(i);
    const T r = block_reduce(s, OpSum(), ws);
    if (threadIdx.x == 0) partials[blockIdx.x] = r;
}

// An Out may also load what it will need per item before the scan
// (`load(i, value)`, returning a payload that `emit(i, prefix, value,
// payload)` then receives): every item's loads are then issued together
// instead of one item's loads waiting on the previous item's stores.
template <class O, class = void>
struct OutLoads : std::false_type {
    using type = int;
};
template <class O>
struct OutLoads<O, std::void_t<decltype(std::declval<const O&>().load((i64)0, 0))>> : std::true_type {
    using type = decltype(std::declval<const O&>().load((i64)0, 0));
};

// out(i, exclusive_prefix, value) is called for every i < n; total (if not
// null) receives the sum of all values.
template <class T, class F, class Out>
__global__ __launch_bounds__(kBlock) void k_scan_down(F f, i64 n, i64 chunk, Out out,
                                                      const T* __restrict__ partials, T* total, T* mirror) {
    __shared__ T ws[kWaves];
    __shared__ T sv[kBlock * kScanPad];
    constexpr bool loads = OutLoads<Out>::value;
    using Payload = typename OutLoads<Out>::type;
    T p = T(0);
    for (unsigned b = threadIdx.x; b < blockIdx.x; b += kBlock) p += partials[b];
    T run = block_reduce(p, OpSum(), ws);
    const i64 c0 = (i64)blockIdx.x * chunk;
    const i64 c1 = c0 + chunk < n ? c0 + chunk : n;
    for (i64 tile0 = c0; tile0 < c1; tile0 += kScanTile) {
        T v[kScanItems];
        Payload pl[kScanItems];
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            const int j = k * kBlock + threadIdx.x;
            const i64 i = tile0 + j;
            v[k] = i < c1 ? (T)f(i) : T(0);
            sv[scan_slot(j)] = v[k];
        }
        if constexpr (loads) {
#pragma unroll
            for (int k = 0; k < kScanItems; ++k) {
                const i64 i = tile0 + k * kBlock + threadIdx.x;
                pl[k] = out.load(i < c1 ? i : -1, v[k]);
            }
        }
        __syncthreads();
        T s = T(0);
        T bl[kScanItems];
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            bl[k] = sv[threadIdx.x * kScanPad + k];
            s += bl[k];
        }
        T agg;
        const T excl = block_excl_scan(s, ws, &agg);  // its barriers order the reads above before the writes below
        // blocked exclusive prefixes back into LDS, then emit in striped order
        T r = run + excl;
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            sv[threadIdx.x * kScanPad + k] = r;
            r += bl[k];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            const int j = k * kBlock + threadIdx.x;
            const i64 i = tile0 + j;
            if constexpr (loads) {
                if (i < c1) out.emit(i, sv[scan_slot(j)], v[k], pl[k]);
            } else {
                if (i < c1) out(i, sv[scan_slot(j)], v[k]);
            }
        }
        run += agg;
        __syncthreads();  // the next tile overwrites sv
    }
    if (threadIdx.x == 0 && blockIdx.x == gridDim.x - 1) {
        if (total) *total = run;
        if (mirror) *mirror = run;
    }
}

template <class T>
struct LoadFn {
    const T* p;
    __device__ __forceinline__ T operator()(i64 i) const { return p[i]; }
};

template <class T>
struct StoreExcl {
    T* p;
    __device__ __forceinline__ void operator()(i64 i, T run, T) const { p[i] = run; }
};

// Generic scan: out(i, prefix, value) for all i < n; *total on device, and
// at `mirror` (a Ctx::mirror slot) when that is set.
template <class T, class F, class Out>
void scan_apply(Ctx& ctx, F f, i64 n, Out out, T* total, T* mirror = nullptr) {
    if (n <= 0) {
        if (total) ctx.zero(total, 1);
        if (mirror) *static_cast<T*>(ctx.mirror_host()) = T(0);  // no kernel stores the empty total
        return;
    }
    static_assert(std::is_integral<T>::value, "integer scans only: the chunking must not change the sums");
    const i64 nt = (n + kScanTile - 1) / kScanTile;
    const i64 maxb = std::min<i64>(tune::get(tune::ScanChunks, n), kScanMaxBlocks);
    const i64 per = (nt + maxb - 1) / maxb;  // tiles per block
    const i64 chunk = per * kScanTile;
    const unsigned g = (unsigned)((n + chunk - 1) / chunk);
    T* partials = nullptr;
    if (g > 1) {
        static_assert(sizeof(T) <= 8, "scan type");
        partials = static_cast<T*>(ctx.scan_scratch(kScanMaxBlocks * 8));
        k_scan_up<T, F><<<g, kBlock, 0, ctx.stream>>>(f, n, chunk, partials);
        RAMA_LAUNCHED();
    }
    k_scan_down<T, F, Out><<<g, kBlock, 0, ctx.stream>>>(f, n, chunk, out, partials, total, mirror);
    RAMA_LAUNCHED();
}

// Exclusive scan of f(0..n-1) into out[0..n]; out[n] = total.  `out` may
// alias the storage f reads (item i is written only by the block that owns
// it, after the first launch and after its tile has read all its items).
template <class T, class F>
void scan_exclusive(Ctx& ctx, F f, T* out, i64 n, T* mirror = nullptr) {
    if (n <= 0) {
        ctx.zero(out, 1);
        if (mirror) *static_cast<T*>(ctx.mirror_host()) = T(0);
        return;
    }
    scan_apply<T>(ctx, f, n, StoreExcl<T>{out}, out + n, mirror);
}

template <class T>
void scan_exclusive_arr(Ctx& ctx, const T* in, T* out, i64 n, T* mirror = nullptr) {
    scan_exclusive<T>(ctx, LoadFn<T>{in}, out, n, mirror);
}

// Raises if any look-back gave up (never expected; guards against a hang).
// The flag is cleared before raising, so later solves on the same context
// (batch workers and serving lanes reuse theirs) start clean.
inline void scan_check(Ctx& ctx) {
    if (ctx.scan_counter.empty()) return;
    int h[2];
    ctx.read(ctx.scan_counter.get(), h, 2);
    if (h[1]) {
        ctx.zero(ctx.scan_counter.get(), 2);
        ctx.sync();
        throw Error(RAMA_ERR_INTERNAL, "single-pass scan look-back timed out");
    }
}

// Clears a look-back error flag left by a solve that ended early on another
// error, before a new solve on the context relies on it.
inline void scan_reset(Ctx& ctx) {
    if (!ctx.scan_counter.empty()) ctx.zero(ctx.scan_counter.get() + 1, 1);
}

// ------------------------------------------------------------- reduction
// Every block reduces its grid-stride share and stores the partial; a
// second one-block launch combines the partials in block order, so the
// result is deterministic for a given n.  A single block combines its own
// partial the same way, in the same launch.  No tickets: same-address
// device atomics from every block serialise.
constexpr unsigned kReduceMaxBlocks = 1024;

// a = op(ident, partials[b]) over b ≡ thread (mod kBlock), then the block
// tree: the combination of the partials in block order.
template <class T, class Op>
__device__ __forceinline__ T combine_partials(const T* partials, unsigned nb, T ident, Op op, T* ws) {
    T a = ident;
    for (unsigned b = threadIdx.x; b < nb; b += kBlock) a = op(a, partials[b]);
    return block_reduce(a, op, ws);
}

template <class T, class Op, class F>
__global__ __launch_bounds__(kBlock) void k_reduce1(F f, i64 n, T ident, Op op, T* partials, T* dst, T* mirror) {
    __shared__ T ws[kWaves];
    __shared__ T own;
    T acc = ident;
    for (i64 i = (i64)blockIdx.x * kBlock + threadIdx.x; i < n; i += (i64)gridDim.x * kBlock) acc = op(acc, (T)f(i));
    const T r = block_reduce(acc, op, ws);
    if (gridDim.x > 1) {
        if (threadIdx.x == 0) partials[blockIdx.x] = r;
        return;
    }
    if (threadIdx.x == 0) own = r;
    __syncthreads();
    const T total = combine_partials(&own, 1u, ident, op, ws);
    if (threadIdx.x == 0) {
        *dst = total;
        if (mirror) *mirror = total;
    }
}

template <class T, class Op>
__global__ __launch_bounds__(kBlock) void k_reduce_combine(const T* __restrict__ partials, unsigned nb, T ident,
                                                           Op op, T* dst, T* mirror) {
    __shared__ T ws[kWaves];
    const T total = combine_partials(partials, nb, ident, op, ws);
    if (threadIdx.x == 0) {
        *dst = total;
        if (mirror) *mirror = total;
    }
}

// Deterministic reduction (fixed grid, fixed tree); result stays on device in
// *dst, and is also stored at `mirror` (a Ctx::mirror slot) when that is set.
template <class T, class Op, class F>
void reduce_to(Ctx& ctx, F f, i64 n, T ident, Op op, T* dst, T* mirror = nullptr) {
    const unsigned nb = (unsigned)std::min<i64>(std::max<i64>((n + kBlock - 1) / kBlock, 1), kReduceMaxBlocks);
    static_assert(sizeof(T) <= 8, "reduction type");
    T* scratch = static_cast<T*>(ctx.reduce_scratch(kReduceMaxBlocks * 8));
    k_reduce1<T, Op, F><<<nb, kBlock, 0, ctx.stream>>>(f, n, ident, op, scratch, dst, mirror);
    RAMA_LAUNCHED();
    if (nb > 1) {
        k_reduce_combine<T, Op><<<1, kBlock, 0, ctx.stream>>>(scratch, nb, ident, op, dst, mirror);
        RAMA_LAUNCHED();
    }
}

// Two double sums in one pass, each reduced exactly as reduce_to would
// reduce it alone (same grid, grid-stride order, block tree and partial
// combination), so the results are bit-identical to two separate calls.
// f(i, a, b) gives element i's two terms.
template <class F>
__global__ __launch_bounds__(kBlock) void k_reduce_pair(F f, i64 n, double* partials, double* dst_a, double* dst_b) {
    __shared__ double ws[kWaves];
    __shared__ double own[2];
    double a = 0.0, b = 0.0;
    for (i64 i = (i64)blockIdx.x * kBlock + threadIdx.x; i < n; i += (i64)gridDim.x * kBlock) {
        double x, y;
        f(i, x, y);
        a += x;
        b += y;
    }
    const double ra = block_reduce(a, OpSum(), ws);
    const double rb = block_reduce(b, OpSum(), ws);
    if (gridDim.x > 1) {
        if (threadIdx.x == 0) {
            partials[blockIdx.x] = ra;
            partials[kReduceMaxBlocks + blockIdx.x] = rb;
        }
        return;
    }
    if (threadIdx.x == 0) {
        own[0] = ra;
        own[1] = rb;
    }
    __syncthreads();
    const double ta = combine_partials(own, 1u, 0.0, OpSum(), ws);
    const double tb = combine_partials(own + 1, 1u, 0.0, OpSum(), ws);
    if (threadIdx.x == 0) {
        *dst_a = ta;
        *dst_b = tb;
    }
}

__global__ __launch_bounds__(kBlock) inline void k_reduce_pair_combine(const double* __restrict__ partials, unsigned nb,
                                                                       double* dst_a, double* dst_b) {
    __shared__ double ws[kWaves];
    const double ta = combine_partials(partials, nb, 0.0, OpSum(), ws);
    const double tb = combine_partials(partials + kReduceMaxBlocks, nb, 0.0, OpSum(), ws);
    if (threadIdx.x == 0) {
        *dst_a = ta;
        *dst_b = tb;
    }
}

template <class F>
void reduce_pair_to(Ctx& ctx, F f, i64 n, double* dst_a, double* dst_b) {
    const unsigned nb = (unsigned)std::min<i64>(std::max<i64>((n + kBlock - 1) / kBlock, 1), kReduceMaxBlocks);
    double* scratch = static_cast<double*>(ctx.reduce_scratch(2 * kReduceMaxBlocks * 8));
    k_reduce_pair<F><<<nb, kBlock, 0, ctx.stream>>>(f, n, scratch, dst_a, dst_b);
    RAMA_LAUNCHED();
    if (nb > 1) {
        k_reduce_pair_combine<<<1, kBlock, 0, ctx.stream>>>(scratch, nb, dst_a, dst_b);
        RAMA_LAUNCHED();
    }
}

template <class T, class Op, class F>
T reduce(Ctx& ctx, F f, i64 n, T ident, Op op) {
    DBuf<T> r = ctx.buf<T>(1);
    T* m = ctx.mirror<T>();
    reduce_to(ctx, f, n, ident, op, r.get(), m);
    T h;
    ctx.read_mirror(r.get(), m, &h);
    return h;
}

// ------------------------------------------------ segmented reduction
// Segment q of [beg[q], beg[q+1]) reduced exactly as reduce_to reduces an
// array of that length on its own: the same number of logical blocks
// (seg_blocks(n)), the same grid-stride element order, the same block tree,
// and the same combination of the partials in block order.  So a result is
// bit-identical whether its instance is reduced alone or packed with
// others.  Two launches and no atomics: a bounded grid of workgroups walks
// the logical blocks of all segments (each finds its segment by a binary
// search over the block-count prefix held in LDS) and stores their
// partials; then one workgroup per segment combines that segment's
// partials.  bpre (Q+1) is the prefix of the segments' block counts.
inline unsigned seg_blocks(i64 n) {
    return (unsigned)std::min<i64>(std::max<i64>((n + kBlock - 1) / kBlock, 1), kReduceMaxBlocks);
}

// Owner of block b given the prefix pre[0..Q] of per-segment block counts
// (every segment has at least one block).
__device__ __forceinline__ int block_owner(const int* pre, int Q, int b) {
    int lo = 0, hi = Q - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pre[mid] <= b) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Segments whose block-count prefix fits in LDS for the owner search.
constexpr int kSegLdsMax = 4096;

template <class T, class Op, class F>
__global__ __launch_bounds__(kBlock) void k_seg_partials(F f, const i64* __restrict__ beg,
                                                         const int* __restrict__ bpre, int Q, int total, T ident,
                                                         Op op, T* __restrict__ partials) {
    __shared__ T ws[kWaves];
    __shared__ int spre[kSegLdsMax + 1];
    const bool lds = Q <= kSegLdsMax;
    if (lds)
        for (int k = threadIdx.x; k <= Q; k += kBlock) spre[k] = bpre[k];
    __syncthreads();
    for (int L = blockIdx.x; L < total; L += gridDim.x) {  // uniform per workgroup
        const int q = block_owner(lds ? spre : bpre, Q, L);
        const int p0 = lds ? spre[q] : bpre[q];
        const int nb = (lds ? spre[q + 1] : bpre[q + 1]) - p0;
        const int lb = L - p0;
        const i64 b0 = beg[q], n = beg[q + 1] - b0;
        T acc = ident;
        for (i64 i = (i64)lb * kBlock + threadIdx.x; i < n; i += (i64)nb * kBlock) acc = op(acc, (T)f(b0 + i));
        const T r = block_reduce(acc, op, ws);
        if (threadIdx.x == 0) partials[L] = r;
    }
}

template <class T, class Op>
__global__ __launch_bounds__(kBlock) void k_seg_combine(const int* __restrict__ bpre, T ident, Op op,
                                                        const T* __restrict__ partials, T* dst, int stride) {
    __shared__ T ws[kWaves];
    const int q = blockIdx.x;
    const int p0 = bpre[q], nb = bpre[q + 1] - p0;
    T a = ident;
    for (int k = threadIdx.x; k < nb; k += kBlock) a = op(a, partials[p0 + k]);
    const T total = block_reduce(a, op, ws);
    if (threadIdx.x == 0) dst[(i64)q * stride] = total;
}

// Host side of the segmented reduction: hbeg/dbeg are the segment bounds
// (Q+1) on the host and on the device; dst[q * stride] receives segment q.
template <class T, class Op, class F>
void seg_reduce(Ctx& ctx, F f, const i64* hbeg, const i64* dbeg, int Q, T ident, Op op, T* dst, int stride) {
    if (Q <= 0) return;
    static_assert(sizeof(T) <= 8, "reduction type");
    std::vector<int> bpre((size_t)Q + 1);
    bpre[0] = 0;
    for (int q = 0; q < Q; ++q) bpre[q + 1] = bpre[q] + (int)seg_blocks(hbeg[q + 1] - hbeg[q]);
    DBuf<int> dpre = ctx.buf<int>(Q + 1);
    ctx.upload(dpre.get(), bpre.data(), Q + 1);
    DBuf<T> part = ctx.buf<T>(bpre[Q]);
    constexpr int kMaxGroups = 4096;  // workgroups walking the logical blocks
    const unsigned g = (unsigned)std::min(bpre[Q], kMaxGroups);
    k_seg_partials<T, Op, F><<<g, kBlock, 0, ctx.stream>>>(f, dbeg, dpre.get(), Q, bpre[Q], ident, op, part.get());
    RAMA_LAUNCHED();
    k_seg_combine<T, Op><<<(unsigned)Q, kBlock, 0, ctx.stream>>>(dpre.get(), ident, op, part.get(), dst, stride);
    RAMA_LAUNCHED();
}

// Per-segment integer reductions over contiguous ranges known on the device:
// segment q = [rb[q], rb[q+1]) is covered by kSegSub blocks; each block
// reduces its stride of the range and adds its result to the segment with
// one atomic, so a few atomics per segment replace one per element.
// Integer min/max/sum only (order does not matter).  f(i, acc) folds element
// i into the block-local accumulator; fin(q, acc) publishes it.
constexpr int kSegSub = 8;
template <class Acc, class F, class Red, class Fin>
__global__ __launch_bounds__(kBlock) void k_range_reduce(const i64* __restrict__ rb, int Q, F f, Red red, Fin fin) {
    __shared__ __align__(16) unsigned char ws_raw[kWaves * sizeof(Acc)];  // Acc may have initialisers
    Acc* ws = reinterpret_cast<Acc*>(ws_raw);
    const int q = blockIdx.x / kSegSub, sub = blockIdx.x % kSegSub;
    if (q >= Q) return;
    const i64 lo = rb[q], hi = rb[q + 1];
    Acc acc = Acc();
    for (i64 i = lo + (i64)sub * kBlock + threadIdx.x; i < hi; i += (i64)kSegSub * kBlock) f(i, acc);
    // block combine through LDS (Acc is a small POD with a combine op)
    const int lane = lane_id(), w = threadIdx.x >> 6;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) acc = red(acc, red.shfl_down(acc, d));
    if (lane == 0) ws[w] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        Acc r = ws[0];
        for (int k = 1; k < kWaves; ++k) r = red(r, ws[k]);
        if (hi > lo) fin(q, r);
    }
}
template <class Acc, class F, class Red, class Fin>
void range_reduce(Ctx& ctx, const i64* rb, int Q, F f, Red red, Fin fin) {
    if (Q <= 0) return;
    k_range_reduce<Acc, F, Red, Fin><<<(unsigned)Q * kSegSub, kBlock, 0, ctx.stream>>>(rb, Q, f, red, fin);
    RAMA_LAUNCHED();
}

// ------------------------------------------------------- simple kernels// ------------------------------------------------------- simple kernels
template <class T>
__global__ void k_fill(T* p, i64 n, T v) {
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
template <class T>
void fill(Ctx& ctx, T* p, i64 n, T v) {
    if (n <= 0) return;
    k_fill<T><<<grid_for(n), kBlock, 0, ctx.stream>>>(p, n, v);
    RAMA_LAUNCHED();
}

template <class T>
__global__ void k_iota(T* p, i64 n) {
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = (T)i;
}
template <class T>
void iota(Ctx& ctx, T* p, i64 n) {
    if (n <= 0) return;
    k_iota<T><<<grid_for(n), kBlock, 0, ctx.stream>>>(p, n);
    RAMA_LAUNCHED();
}

// Generic element-wise launch: f(i) for i in [0, n).
template <class F>
__global__ void k_for_each(F f, i64 n) {
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) f(i);
}
template <class F>
void for_each(Ctx& ctx, i64 n, F f) {
    if (n <= 0) return;
    k_for_each<F><<<grid_for(n), kBlock, 0, ctx.stream>>>(f, n);
    RAMA_LAUNCHED();
}

// f(i) for every thread of a grid covering [0, n) — also for i >= n in the
// last block, so wave-wide operations inside f see whole waves; f checks
// the bound itself.
template <class F>
__global__ void k_for_all(F f, i64 n) {
    f((i64)blockIdx.x * blockDim.x + threadIdx.x);
}
template <class F>
void for_all(Ctx& ctx, i64 n, F f) {
    k_for_all<F><<<grid_for(n), kBlock, 0, ctx.stream>>>(f, n);
    RAMA_LAUNCHED();
}

template <class P>
struct FlagFn {
    P pred;
    __device__ __forceinline__ int operator()(i64 i) const { return pred(i) ? 1 : 0; }
};

// Stable compaction helper: pos[i] = #{k < i : pred(k)}, returns the count.
template <class P>
i64 select_positions(Ctx& ctx, P pred, i64 n, DBuf<int>& pos) {
    pos = ctx.buf<int>(n + 1);
    scan_exclusive<int>(ctx, FlagFn<P>{pred}, pos.get(), n);
    return ctx.read1(pos.get() + n);
}

template <class Emit>
struct SelectOut {
    Emit emit;
    __device__ __forceinline__ void operator()(i64 i, int pos, int flag) const {
        if (flag) emit(i, pos);
    }
};

// Stable compaction in one launch: emit(i, position) for every i with
// pred(i); the number selected is left on the device in *d_count.
template <class P, class Emit>
void select_into(Ctx& ctx, i64 n, P pred, Emit emit, int* d_count, int* mirror = nullptr) {
    scan_apply<int>(ctx, FlagFn<P>{pred}, n, SelectOut<Emit>{emit}, d_count, mirror);
}

template <class Load, class Emit>
struct SelectLoadOut {
    Load ld;
    Emit emit_;
    // i < 0: past the end; unselected items load nothing
    __device__ __forceinline__ auto load(i64 i, int flag) const {
        using R = decltype(ld((i64)0));
        return i >= 0 && flag ? ld(i) : R{};
    }
    template <class R>
    __device__ __forceinline__ void emit(i64 i, int pos, int flag, const R& r) const {
        if (flag) emit_(i, pos, r);
    }
};

// select_into whose emit(i, position, load(i)) receives what load(i) read
// for the selected items, all loaded before the block's scan.
template <class P, class Load, class Emit>
void select_into_loaded(Ctx& ctx, i64 n, P pred, Load load, Emit emit, int* d_count, int* mirror = nullptr) {
    scan_apply<int>(ctx, FlagFn<P>{pred}, n, SelectLoadOut<Load, Emit>{load, emit}, d_count, mirror);
}

// ------------------------------------------------------------ radix sort
constexpr int kRsItems = 16;
constexpr int kRsTile = kBlock * kRsItems;  // 4096 keys per tile

// Stable LSD sort of (keys, vals) by the low `bits` bits of the key; K is
// uint32_t or uint64_t.  `keys`/`vals` hold the result; `tk`/`tv` are scratch
// of the same length.  Digits are balanced: ceil(bits/8) passes of at most
// 8 bits each.
template <class K>
void radix_sort_pairs(Ctx& ctx, K* keys, u32* vals, K* tk, u32* tv, i64 n, int bits);

// Convenience: allocates the scratch itself.
template <class K>
void radix_sort_pairs(Ctx& ctx, DBuf<K>& keys, DBuf<u32>& vals, i64 n, int bits);
// As the pointer form, with the sort path (0 three-launch, 1 onesweep,
// 2 top-digit + per-bucket blocks) and threads per tile given instead of
// taken from the tuning table (the tuner times paths with it).
template <class K>
void radix_sort_pairs_shaped(Ctx& ctx, K* keys, u32* vals, K* tk, u32* tv, i64 n, int bits, int path, int threads);

inline int bits_for(i64 max_value) {
    int b = 1;
    while (b < 63 && ((i64)1 << b) <= max_value) ++b;
    return b;
}

}  // namespace rama
