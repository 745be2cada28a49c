// prims.hip — stable LSD radix sort of (u32 or u64 key, u32 value) pairs.
//
// Digits of at most 8 bits, balanced over the key width.  Inside a
// 4096-pair tile every path ranks by wave64 ballot multi-split and reorders
// through LDS so each digit run is written out contiguously; items are
// ranked in (wave, item, lane) order, which is their input order, so each
// pass is stable and the whole sort is stable.  Across tiles, per array size
// (tuning table): three launches per pass (per-tile histograms, their scan,
// the scatter), or one histogram launch for every pass and then one
// "onesweep" launch per pass with decoupled look-back for each digit's
// global offset, or a top-digit onesweep pass followed by one block per
// bucket.  Arrays of at most 8192 pairs are sorted by one block.
#include <cstdlib>

#include "prims.cuh"
#include "tuning.hpp"

namespace rama {

// In-place exclusive scan of 256 LDS counts by wave 0.  Block-uniform call.
__device__ __forceinline__ void scan256(int* a) {
    if (threadIdx.x < 64) {
        const int lane = lane_id();
        int v[4], s = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            v[q] = a[lane * 4 + q];
            s += v[q];
        }
        int inc = s;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int t = __shfl_up(inc, o);
            if (lane >= o) inc += t;
        }
        int run = inc - s;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            a[lane * 4 + q] = run;
            run += v[q];
        }
    }
    __syncthreads();
}


// Three-launch pass for large arrays: per-tile digit histogram
// (digit-major table), a scan of the table, and the scatter.
// Adds one count per active lane to h[d], wave-aggregated: up to two
// groups of lanes sharing a digit (the common case on partly sorted keys,
// where whole waves share their high digits) are added by one leader lane
// each; the remaining lanes add individually.
__device__ __forceinline__ void wave_count(int* h, u32 d, bool ok) {
    const int lane = lane_id();
    unsigned long long active = __ballot(ok);
#pragma unroll
    for (int it = 0; it < 2; ++it) {
        if (!active) return;
        const int leader = __builtin_ctzll(active);
        const u32 dl = (u32)__builtin_amdgcn_readlane((int)d, leader);
        const unsigned long long m = __ballot(d == dl) & active;
        if (lane == leader) atomicAdd(&h[dl], __popcll(m));
        active &= ~m;
    }
    if ((active >> lane) & 1ull) atomicAdd(&h[d], 1);
}

// The radix passes run NT threads per 4096-pair tile: 1024 (4 pairs each,
// few registers per thread, so many waves per SIMD within the same LDS) or
// 512 (8 pairs each), as the tuning table says for the array's tile count.
#define RAMA_SC_SHAPE(NT)                     \
    constexpr int kScThreads = (NT);          \
    constexpr int kScItems = kRsTile / (NT);  \
    constexpr int kScWaves = (NT) / 64;

template <class K, int NT>
__global__ __launch_bounds__(NT) void k_radix_hist(const K* __restrict__ keys, i64 n, int shift, int bits, int* hist,
                                                   int ntiles) {
    RAMA_SC_SHAPE(NT)
    __shared__ int h[256];
    if (threadIdx.x < 256) h[threadIdx.x] = 0;
    __syncthreads();
    const u32 mask = (1u << bits) - 1u;
    const i64 base = (i64)blockIdx.x * kRsTile;
    K key[kScItems];
#pragma unroll
    for (int k = 0; k < kScItems; ++k) {
        const i64 i = base + (i64)k * kScThreads + threadIdx.x;
        key[k] = i < n ? keys[i] : K(0);
    }
#pragma unroll
    for (int k = 0; k < kScItems; ++k) {
        const i64 i = base + (i64)k * kScThreads + threadIdx.x;
        wave_count(h, (u32)(key[k] >> shift) & mask, i < n);
    }
    __syncthreads();
    if (threadIdx.x < (1 << bits)) hist[(i64)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

// Lanes of the wave holding the same digit as this lane (ballot
// multi-split), with a shortcut when the whole wave shares one digit.
__device__ __forceinline__ unsigned long long digit_peers(u32 d, int bits) {
    const u32 d0 = (u32)__builtin_amdgcn_readfirstlane((int)d);
    const unsigned long long all = __ballot(1);
    if (__ballot(d == d0) == all) return all;
    unsigned long long peers = all;
    for (int b = 0; b < bits; ++b) {
        const unsigned long long m = __ballot((d >> b) & 1u);
        peers &= ((d >> b) & 1u) ? m : ~m;
    }
    return peers;
}

template <class K, int NT>
__global__ __launch_bounds__(NT) void k_radix_scatter(const K* __restrict__ kin, const u32* __restrict__ vin,
                                                      K* __restrict__ kout, u32* __restrict__ vout, i64 n, int shift,
                                                      int bits, const int* __restrict__ offs, int ntiles) {
    RAMA_SC_SHAPE(NT)
    // one reorder buffer, used for the keys and then for the values, keeps
    // the tile's LDS at about 50 KB: three tiles per CU instead of two
    __shared__ K sk[kRsTile];
    u32* const sv = reinterpret_cast<u32*>(sk);
    __shared__ int wcnt[kScWaves][256];
    __shared__ int dstart[256];
    __shared__ int gbase[256];

    const int tid = threadIdx.x, lane = lane_id(), w = tid >> 6;
    const u32 mask = (1u << bits) - 1u;
    for (int k = tid; k < kScWaves * 256; k += kScThreads) (&wcnt[0][0])[k] = 0;
    __syncthreads();

    const i64 tile0 = (i64)blockIdx.x * kRsTile;
    const i64 wbase = tile0 + (i64)w * (64 * kScItems);
    K key[kScItems];
    u32 val[kScItems];
    int rnk[kScItems];
#pragma unroll
    for (int k = 0; k < kScItems; ++k) {
        const i64 i = wbase + (i64)k * 64 + lane;
        const bool ok = i < n;
        // streamed once per pass: non-temporal loads (measured faster here)
        key[k] = ok ? __builtin_nontemporal_load(&kin[i]) : ~K(0);
        val[k] = ok ? __builtin_nontemporal_load(&vin[i]) : 0u;
    }
    const unsigned long long lt = (1ull << lane) - 1ull;
#pragma unroll
    for (int k = 0; k < kScItems; ++k) {
        const u32 d = (u32)(key[k] >> shift) & mask;
        const unsigned long long peers = digit_peers(d, bits);
        const int r = __popcll(peers & lt);
        const int c = __popcll(peers);
        const int before = wcnt[w][d];
        __builtin_amdgcn_wave_barrier();
        if (r == 0) wcnt[w][d] = before + c;
        __builtin_amdgcn_wave_barrier();
        rnk[k] = before + r;
    }
    __syncthreads();
    if (tid < 256) {
        const int d = tid;
        int tot = 0;
#pragma unroll
        for (int q = 0; q < kScWaves; ++q) {
            const int c = wcnt[q][d];
            wcnt[q][d] = tot;
            tot += c;
        }
        dstart[d] = tot;
        if (d < (1 << bits)) gbase[d] = offs[(i64)d * ntiles + blockIdx.x];
    }
    __syncthreads();
    scan256(dstart);
    int pos[kScItems];
#pragma unroll
    for (int k = 0; k < kScItems; ++k) {
        const u32 d = (u32)(key[k] >> shift) & mask;
        pos[k] = dstart[d] + wcnt[w][d] + rnk[k];
        sk[pos[k]] = key[k];
    }
    __syncthreads();
    const int nvalid = (int)((n - tile0) < kRsTile ? (n - tile0) : kRsTile);
    // the keys in digit runs; each element's output place is kept for its value
    i64 outp[(kRsTile + kScThreads - 1) / kScThreads];
#pragma unroll
    for (int q = 0; q < (kRsTile + kScThreads - 1) / kScThreads; ++q) {
        const int p = tid + q * kScThreads;
        outp[q] = -1;
        if (p < nvalid) {
            const K k = sk[p];
            const u32 d = (u32)(k >> shift) & mask;
            const i64 o = (i64)gbase[d] + (p - dstart[d]);
            kout[o] = k;
            outp[q] = o;
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kScItems; ++k) sv[pos[k]] = val[k];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < (kRsTile + kScThreads - 1) / kScThreads; ++q) {
        const int p = tid + q * kScThreads;
        if (p < nvalid) vout[outp[q]] = sv[p];
    }
}

struct PassPlan {
    int passes;
    int shift[8];
    int width[8];
};

// Digit histograms of every pass in one read of the keys.  Each block
// stores its partial histograms (row blockIdx.x of `part`); the last block
// to finish (atomic ticket, reset for the next launch) adds the rows in
// block order into `hist`, so nothing needs clearing beforehand.
template <class K>
__global__ __launch_bounds__(kBlock) void k_radix_hist_all(const K* __restrict__ keys, i64 n, PassPlan plan,
                                                           int* __restrict__ part, int* __restrict__ hist,
                                                           unsigned* ticket) {
    __shared__ int h[8 * 256];
    __shared__ bool last;
    const int nb = plan.passes * 256;
    for (int k = threadIdx.x; k < nb; k += kBlock) h[k] = 0;
    __syncthreads();
    for (i64 i = (i64)blockIdx.x * kBlock + threadIdx.x; i < n; i += (i64)gridDim.x * kBlock) {
        const K key = keys[i];
        for (int p = 0; p < plan.passes; ++p)
            atomicAdd(&h[p * 256 + (int)((u32)(key >> plan.shift[p]) & ((1u << plan.width[p]) - 1u))], 1);
    }
    __syncthreads();
    int* mine = part + (i64)blockIdx.x * nb;
    for (int k = threadIdx.x; k < nb; k += kBlock)
        __hip_atomic_store(&mine[k], h[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();  // the row is stored before thread 0 releases it with the ticket
    if (threadIdx.x == 0) {
        const unsigned t = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        last = t == gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    __threadfence();  // acquire: the other blocks' rows are visible to plain loads
    for (int k = threadIdx.x; k < nb; k += kBlock) {
        int sum = 0;
        unsigned b = 0;
        for (; b + 8 <= gridDim.x; b += 8) {
            int v[8];  // eight loads in flight
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = part[(i64)(b + u) * nb + k];
#pragma unroll
            for (int u = 0; u < 8; ++u) sum += v[u];
        }
        for (; b < gridDim.x; ++b) sum += part[(i64)b * nb + k];
        hist[k] = sum;
    }
    if (threadIdx.x == 0) *ticket = 0;
}

// One onesweep pass.  Tiles are drawn from an atomic counter (every smaller
// tile is already running).  A tile publishes its per-digit counts, ranks
// its keys stably (wave64 ballot multi-split), then each of 256 threads looks
// back along its digit for the exclusive prefix over earlier tiles, and the
// tile is reordered in LDS and written out in digit runs.  Status words are
// (epoch | flag | count) as in the single-pass scan.
template <class K, int NT>
__global__ __launch_bounds__(NT) void k_onesweep(const K* __restrict__ kin, const u32* __restrict__ vin,
                                                 K* __restrict__ kout, u32* __restrict__ vout, i64 n, int shift,
                                                 int bits, const int* __restrict__ hist_pass, u64* __restrict__ status,
                                                 int* ctl, u32 epoch) {
    RAMA_SC_SHAPE(NT)
    __shared__ K sk[kRsTile];
    __shared__ u32 sv[kRsTile];
    __shared__ int wcnt[kScWaves][256];
    __shared__ int tcnt[256];
    __shared__ int dstart[256];
    __shared__ int dglob[256];
    __shared__ i64 gbase[256];
    __shared__ int tile_s;

    const int tid = threadIdx.x, lane = lane_id(), w = tid >> 6;
    const u32 mask = (1u << bits) - 1u;
    if (tid == 0) tile_s = atomicAdd(&ctl[0], 1);
    for (int k = tid; k < kScWaves * 256; k += kScThreads) (&wcnt[0][0])[k] = 0;
    if (tid < 256) tcnt[tid] = 0;
    __syncthreads();
    const int tile = tile_s;
    const i64 tile0 = (i64)tile * kRsTile;
    const i64 wbase = tile0 + (i64)w * (64 * kScItems);
    K key[kScItems];
    u32 val[kScItems];
    int rnk[kScItems];
#pragma unroll
    for (int k = 0; k < kScItems; ++k) {
        const i64 i = wbase + (i64)k * 64 + lane;
        const bool ok = i < n;
        key[k] = ok ? kin[i] : ~K(0);  // padding sorts into the last bucket, after all real keys
        val[k] = ok ? vin[i] : 0u;
    }
#pragma unroll
    for (int k = 0; k < kScItems; ++k) wave_count(tcnt, (u32)(key[k] >> shift) & mask, wbase + (i64)k * 64 + lane < n);
    __syncthreads();
    // publish this tile's digit counts early so successors can look back
    u64* my = status + (i64)tile * 256;
    if (tid <= (int)mask) {
        __hip_atomic_store(&my[tid], scan_word(epoch, tile == 0 ? 2u : 1u, (u64)tcnt[tid]), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
    // stable ranks within the tile
    const unsigned long long lt = (1ull << lane) - 1ull;
#pragma unroll
    for (int k = 0; k < kScItems; ++k) {
        const u32 d = (u32)(key[k] >> shift) & mask;
        const unsigned long long peers = digit_peers(d, bits);
        const int r = __popcll(peers & lt);
        const int c = __popcll(peers);
        const int before = wcnt[w][d];
        __builtin_amdgcn_wave_barrier();
        if (r == 0) wcnt[w][d] = before + c;
        __builtin_amdgcn_wave_barrier();
        rnk[k] = before + r;
    }
    __syncthreads();
    if (tid < 256) {
        const int d = tid;
        int tot = 0;
#pragma unroll
        for (int q = 0; q < kScWaves; ++q) {
            const int c = wcnt[q][d];
            wcnt[q][d] = tot;
            tot += c;
        }
        dstart[d] = tot;
        dglob[d] = d <= (int)mask ? hist_pass[d] : 0;
    }
    __syncthreads();
    scan256(dstart);
    scan256(dglob);  // global start of each digit in this pass
    {
        // then look back over the earlier tiles, one thread per digit
        const int d = tid;
        if (d <= (int)mask) {
            i64 prefix = 0;
            if (tile > 0) {
                // windowed look-back: kLookWin predecessor words are loaded
                // together (independent loads in flight), then consumed
                // nearest first until an inclusive prefix is met; a word not
                // yet published restarts the window at that tile
                constexpr int kLookWin = 8;
                int p = tile - 1;
                long spins = 0;
                bool done = false;
                while (!done) {
                    u64 wd[kLookWin];
#pragma unroll
                    for (int k = 0; k < kLookWin; ++k)
                        wd[k] = p - k >= 0 ? __hip_atomic_load(&status[(i64)(p - k) * 256 + d], __ATOMIC_RELAXED,
                                                                __HIP_MEMORY_SCOPE_AGENT)
                                           : 0ull;
                    int used = 0;
#pragma unroll
                    for (int k = 0; k < kLookWin; ++k) {
                        if (done || used < k || p - k < 0) continue;  // stop at the first gap
                        const u32 fl = (u32)(wd[k] >> 46) & 3u;
                        if ((u32)(wd[k] >> 48) != epoch || fl == 0) continue;
                        prefix += (i64)(wd[k] & kScanValMask);
                        used = k + 1;
                        if (fl == 2) done = true;
                    }
                    if (done) break;
                    p -= used;
                    if (used < kLookWin) {
                        // tile 0 always publishes an inclusive word, so p >= 0 here
                        if (++spins > kScanSpinLimit) {
                            ctl[1] = 1;
                            break;
                        }
                        __builtin_amdgcn_s_sleep(1);
                    }
                }
                __hip_atomic_store(&my[d], scan_word(epoch, 2, (u64)(prefix + tcnt[d])), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            }
            gbase[d] = (i64)dglob[d] + prefix;
        }
    }
    __syncthreads();
    if (tid == 0 && tile == (int)gridDim.x - 1) ctl[0] = 0;  // every block has drawn its tile
#pragma unroll
    for (int k = 0; k < kScItems; ++k) {
        const u32 d = (u32)(key[k] >> shift) & mask;
        const int pos = dstart[d] + wcnt[w][d] + rnk[k];
        sk[pos] = key[k];
        sv[pos] = val[k];
    }
    __syncthreads();
    const int nvalid = (int)((n - tile0) < kRsTile ? (n - tile0) : kRsTile);
    for (int p = tid; p < nvalid; p += kScThreads) {
        const K k = sk[p];
        const u32 d = (u32)(k >> shift) & mask;
        const i64 o = gbase[d] + (p - dstart[d]);
        kout[o] = k;
        vout[o] = sv[p];
    }
}

// Arrays of at most kTsTile pairs: every pass in one block, in place.  The
// pairs stay in registers (kTsItems per thread, in (wave, item, lane)
// order); each pass ranks them stably by ballot multi-split, reorders them
// through LDS and reloads them in the same order, so each pass is stable.
// One launch replaces the histogram launch and one launch per pass.
constexpr int kTsThreads = 1024;
constexpr int kTsItems = 8;
constexpr int kTsTile = kTsThreads * kTsItems;
constexpr int kTsWaves = kTsThreads / 64;

template <class K>
struct TsSmem {
    K sk[kTsTile];
    u32 sv[kTsTile];
    int wcnt[kTsWaves][256];
    int dstart[256];
    int base[256];
    int ph[8][256];  // per-pass digit histograms of a bucket too large for LDS
    int misc[2];
};

// Stable ranks of the block's register-held items by the digit (shift,
// bits): afterwards s.wcnt[w][d] counts the items of digit d in earlier
// waves, s.dstart[d] the block's items of digit d, and rnk[k] is item k's
// rank among its wave's items of the same digit.  Block-uniform call.
template <class K>
__device__ __forceinline__ void tile_rank(TsSmem<K>& s, const K (&key)[kTsItems], int shift, int bits,
                                          int (&rnk)[kTsItems]) {
    const int tid = threadIdx.x, lane = lane_id(), w = tid >> 6;
    const u32 mask = (1u << bits) - 1u;
    const unsigned long long lt = (1ull << lane) - 1ull;
    for (int k = tid; k < kTsWaves * 256; k += kTsThreads) (&s.wcnt[0][0])[k] = 0;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kTsItems; ++k) {
        const u32 d = (u32)(key[k] >> shift) & mask;
        const unsigned long long peers = digit_peers(d, bits);
        const int r = __popcll(peers & lt);
        const int c = __popcll(peers);
        const int before = s.wcnt[w][d];
        __builtin_amdgcn_wave_barrier();
        if (r == 0) s.wcnt[w][d] = before + c;
        __builtin_amdgcn_wave_barrier();
        rnk[k] = before + r;
    }
    __syncthreads();
    if (tid < 256) {
        int tot = 0;
#pragma unroll
        for (int q = 0; q < kTsWaves; ++q) {
            const int c = s.wcnt[q][tid];
            s.wcnt[q][tid] = tot;
            tot += c;
        }
        s.dstart[tid] = tot;
    }
    __syncthreads();
}

// Every pass of `plan` over the block's register-held items, through LDS.
template <class K>
__device__ __forceinline__ void tile_sort_regs(TsSmem<K>& s, K (&key)[kTsItems], u32 (&val)[kTsItems],
                                               const PassPlan& plan) {
    const int lane = lane_id(), w = threadIdx.x >> 6;
    const int wbase = w * 64 * kTsItems;
    int rnk[kTsItems];
    for (int p = 0; p < plan.passes; ++p) {
        const int shift = plan.shift[p], bits = plan.width[p];
        const u32 mask = (1u << bits) - 1u;
        tile_rank(s, key, shift, bits, rnk);
        scan256(s.dstart);
#pragma unroll
        for (int k = 0; k < kTsItems; ++k) {
            const u32 d = (u32)(key[k] >> shift) & mask;
            const int pos = s.dstart[d] + s.wcnt[w][d] + rnk[k];
            s.sk[pos] = key[k];
            s.sv[pos] = val[k];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kTsItems; ++k) {
            const int i = wbase + k * 64 + lane;
            key[k] = s.sk[i];
            val[k] = s.sv[i];
        }
        __syncthreads();
    }
}

// Loads up to kTsTile pairs from (ki, vi) [0, n) into registers in
// (wave, item, lane) order; the padding sorts after every real key (it
// starts after them and has the largest digit in every pass).
template <class K>
__device__ __forceinline__ void tile_load(const K* __restrict__ ki, const u32* __restrict__ vi, int n,
                                          K (&key)[kTsItems], u32 (&val)[kTsItems]) {
    const int lane = lane_id(), wbase = (threadIdx.x >> 6) * 64 * kTsItems;
#pragma unroll
    for (int k = 0; k < kTsItems; ++k) {
        const int i = wbase + k * 64 + lane;
        const bool ok = i < n;
        key[k] = ok ? ki[i] : ~K(0);
        val[k] = ok ? vi[i] : 0u;
    }
}

template <class K>
__device__ __forceinline__ void tile_store(K* __restrict__ ko, u32* __restrict__ vo, int n, const K (&key)[kTsItems],
                                           const u32 (&val)[kTsItems]) {
    const int lane = lane_id(), wbase = (threadIdx.x >> 6) * 64 * kTsItems;
#pragma unroll
    for (int k = 0; k < kTsItems; ++k) {
        const int i = wbase + k * 64 + lane;
        if (i < n) {
            ko[i] = key[k];
            vo[i] = val[k];
        }
    }
}

template <class K>
__global__ __launch_bounds__(kTsThreads) void k_tile_sort(K* __restrict__ keys, u32* __restrict__ vals, int n,
                                                          PassPlan plan) {
    __shared__ TsSmem<K> s;
    K key[kTsItems];
    u32 val[kTsItems];
    tile_load(keys, vals, n, key, val);
    tile_sort_regs(s, key, val, plan);
    tile_store(keys, vals, n, key, val);
}

// Second stage of the top-digit sort.  An LSD pass on the top digit has
// grouped the pairs into 256 buckets (kin/vin, stable); block b sorts bucket
// b stably by the remaining low bits (`plan`) and writes it to the same
// range of kout/vout.  Bucket bounds: with stride 0, `bnd` holds the 256
// bucket sizes; otherwise bucket d starts at bnd[d * stride] (a scanned
// digit-major tile table) and the last one ends at n.  A bucket that fits in
// LDS is sorted as a tile.  A larger one is sorted by the block alone: the
// histograms of every pass in one sweep, then per pass its tiles in order,
// each ranked in LDS and scattered after the earlier tiles' items of each
// digit, ping-ponging between the bucket's two ranges.
template <class K>
__global__ __launch_bounds__(kTsThreads) void k_bucket_sort(K* __restrict__ kin, u32* __restrict__ vin,
                                                            K* __restrict__ kout, u32* __restrict__ vout,
                                                            const int* __restrict__ bnd, int stride, int n,
                                                            PassPlan plan) {
    __shared__ TsSmem<K> s;
    const int tid = threadIdx.x, b = blockIdx.x;
    if (stride == 0) {
        if (tid < 64) {
            // start of bucket b: the sizes of the buckets before it
            int acc = 0;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int d = tid * 4 + q;
                acc += d < b ? bnd[d] : 0;
            }
#pragma unroll
            for (int o = 32; o >= 1; o >>= 1) acc += __shfl_xor(acc, o, 64);
            if (tid == 0) {
                s.misc[0] = acc;
                s.misc[1] = bnd[b];
            }
        }
    } else if (tid == 0) {
        const int b0 = bnd[(i64)b * stride];
        s.misc[0] = b0;
        s.misc[1] = (b < 255 ? bnd[(i64)(b + 1) * stride] : n) - b0;
    }
    __syncthreads();
    const int beg = s.misc[0], m = s.misc[1];
    if (m == 0) return;  // uniform
    K* ka = kin + beg;
    u32* va = vin + beg;
    K* kb = kout + beg;
    u32* vb = vout + beg;
    K key[kTsItems];
    u32 val[kTsItems];
    if (m <= kTsTile) {
        tile_load(ka, va, m, key, val);
        tile_sort_regs(s, key, val, plan);
        tile_store(kb, vb, m, key, val);
        return;
    }
    const int lane = lane_id(), w = tid >> 6, wbase = w * 64 * kTsItems;
    int rnk[kTsItems];
    // digit counts do not depend on the order: every pass's histogram in one sweep
    for (int k = tid; k < 8 * 256; k += kTsThreads) (&s.ph[0][0])[k] = 0;
    __syncthreads();
    for (int i = tid; i < m; i += kTsThreads) {
        const K x = ka[i];
        for (int p = 0; p < plan.passes; ++p)
            atomicAdd(&s.ph[p][(u32)(x >> plan.shift[p]) & ((1u << plan.width[p]) - 1u)], 1);
    }
    __syncthreads();
    for (int p = 0; p < plan.passes; ++p) {
        const int shift = plan.shift[p], bits = plan.width[p];
        const u32 mask = (1u << bits) - 1u;
        K* ks = (p & 1) ? kb : ka;
        u32* vs = (p & 1) ? vb : va;
        K* kd = (p & 1) ? ka : kb;
        u32* vd = (p & 1) ? va : vb;
        if (tid < 256) s.base[tid] = s.ph[p][tid];
        __syncthreads();
        scan256(s.base);
        for (int t0 = 0; t0 < m; t0 += kTsTile) {  // uniform
            const int nt = min(kTsTile, m - t0);
            tile_load(ks + t0, vs + t0, nt, key, val);
            tile_rank(s, key, shift, bits, rnk);
#pragma unroll
            for (int k = 0; k < kTsItems; ++k) {
                if (wbase + k * 64 + lane < nt) {
                    const u32 d = (u32)(key[k] >> shift) & mask;
                    const int o = s.base[d] + s.wcnt[w][d] + rnk[k];
                    kd[o] = key[k];
                    vd[o] = val[k];
                }
            }
            __syncthreads();
            // the padding (kTsTile - nt items) all carry the largest digit
            if (tid < 256) s.base[tid] += s.dstart[tid] - (tid == (int)mask ? kTsTile - nt : 0);
            __syncthreads();
        }
    }
    if ((plan.passes & 1) == 0) {
        // an even pass count ends in the input range: copy to the output
        for (int i = tid; i < m; i += kTsThreads) {
            kb[i] = ka[i];
            vb[i] = va[i];
        }
    }
}

// Launch helpers: one instantiation per threads-per-tile shape.
template <class K>
static void launch_hist(Ctx& ctx, int nt, int ntiles, const K* keys, i64 n, int shift, int bits, int* th) {
    if (nt == 512)
        k_radix_hist<K, 512><<<ntiles, 512, 0, ctx.stream>>>(keys, n, shift, bits, th, ntiles);
    else
        k_radix_hist<K, 1024><<<ntiles, 1024, 0, ctx.stream>>>(keys, n, shift, bits, th, ntiles);
    RAMA_LAUNCHED();
}

template <class K>
static void launch_scatter(Ctx& ctx, int nt, int ntiles, const K* ka, const u32* va, K* kb, u32* vb, i64 n,
                           int shift, int bits, const int* th) {
    if (nt == 512)
        k_radix_scatter<K, 512><<<ntiles, 512, 0, ctx.stream>>>(ka, va, kb, vb, n, shift, bits, th, ntiles);
    else
        k_radix_scatter<K, 1024><<<ntiles, 1024, 0, ctx.stream>>>(ka, va, kb, vb, n, shift, bits, th, ntiles);
    RAMA_LAUNCHED();
}

template <class K>
static void launch_onesweep(Ctx& ctx, int nt, int ntiles, const K* ka, const u32* va, K* kb, u32* vb, i64 n,
                            int shift, int bits, const int* hist_pass, u64* st, u32 epoch) {
    if (nt == 512)
        k_onesweep<K, 512><<<ntiles, 512, 0, ctx.stream>>>(ka, va, kb, vb, n, shift, bits, hist_pass, st,
                                                           ctx.scan_counter.get(), epoch);
    else
        k_onesweep<K, 1024><<<ntiles, 1024, 0, ctx.stream>>>(ka, va, kb, vb, n, shift, bits, hist_pass, st,
                                                             ctx.scan_counter.get(), epoch);
    RAMA_LAUNCHED();
}

// Sorts and returns whether the result ended in (tk, tv) instead of
// (keys, vals) (odd number of passes).  path / threads < 0: from the tuning
// table for this array size.
template <class K>
static bool radix_sort_core(Ctx& ctx, K* keys, u32* vals, K* tk, u32* tv, i64 n, int bits, int path = -1,
                            int threads = -1) {
    if (n <= 1 || bits <= 0) return false;
    RAMA_REQUIRE(n < ((i64)1 << 31), "radix sort supports fewer than 2^31 elements");
    RAMA_REQUIRE(bits <= (int)(8 * sizeof(K)), "radix sort key wider than its type");
    const int ntiles = (int)((n + kRsTile - 1) / kRsTile);
    PassPlan plan{};
    plan.passes = (bits + 7) / 8;
    const int base_w = bits / plan.passes, extra = bits % plan.passes;
    for (int p = 0, sh = 0; p < plan.passes; ++p) {
        plan.width[p] = base_w + (p < extra ? 1 : 0);
        plan.shift[p] = sh;
        sh += plan.width[p];
    }
    K* ka = keys;
    u32* va = vals;
    K* kb = tk;
    u32* vb = tv;
    const double pair_bytes = (double)(sizeof(K) + sizeof(u32));
    if (n <= kTsTile) {
        prof::Scope ps(ctx, "radix_tile_sort", 2.0 * pair_bytes * n);
        k_tile_sort<K><<<1, kTsThreads, 0, ctx.stream>>>(keys, vals, (int)n, plan);
        RAMA_LAUNCHED();
        return false;
    }
    if (path < 0) path = (int)tune::get(sizeof(K) == 8 ? tune::SortPathU64 : tune::SortPathU32, n);
    if (threads < 0) threads = (int)tune::get(tune::SortThreads, ntiles);
    if (path == 2 && (bits <= 8 || ntiles > tune::kBucketPathMaxTiles)) path = 1;
    if (path == 2) {
        // one pass on the top 8 bits into 256 buckets (histogram +
        // onesweep), then each bucket's low bits in one block
        PassPlan top{}, low{};
        top.passes = 1;
        top.shift[0] = bits - 8;
        top.width[0] = 8;
        const int lb = bits - 8;
        low.passes = (lb + 7) / 8;
        const int bw = lb / low.passes, ex = lb % low.passes;
        for (int p = 0, sh = 0; p < low.passes; ++p) {
            low.width[p] = bw + (p < ex ? 1 : 0);
            low.shift[p] = sh;
            sh += low.width[p];
        }
        const unsigned hb = (unsigned)std::min<i64>(ntiles, 1024);
        DBuf<int> hist = ctx.buf<int>((i64)(hb + 1) * 256);
        int* hsum = hist.get() + (i64)hb * 256;
        {
            prof::Scope ps(ctx, "radix_hist_all", (double)sizeof(K) * n);
            k_radix_hist_all<K><<<hb, kBlock, 0, ctx.stream>>>(keys, n, top, hist.get(), hsum, ctx.reduce_ticket());
            RAMA_LAUNCHED();
        }
        u32 epoch;
        u64* st = ctx.scan_begin((i64)ntiles * 256, &epoch);
        {
            prof::Scope ps(ctx, "radix_onesweep", 2.0 * pair_bytes * n);
            launch_onesweep<K>(ctx, threads, ntiles, keys, vals, tk, tv, n, top.shift[0], 8, hsum, st, epoch);
        }
        {
            prof::Scope ps(ctx, "radix_bucket_sort", 2.0 * pair_bytes * n);
            k_bucket_sort<K><<<256, kTsThreads, 0, ctx.stream>>>(tk, tv, keys, vals, hsum, 0, (int)n, low);
            RAMA_LAUNCHED();
        }
        return false;
    }
    if (path == 0) {
        // three launches per pass: per-tile histograms, their scan, scatter
        DBuf<int> th = ctx.buf<int>((i64)256 * ntiles + 1);
        for (int p = 0; p < plan.passes; ++p) {
            {
                prof::Scope ps(ctx, "radix_hist", (double)sizeof(K) * n);
                launch_hist<K>(ctx, threads, ntiles, ka, n, plan.shift[p], plan.width[p], th.get());
            }
            scan_exclusive_arr<int>(ctx, th.get(), th.get(), ((i64)1 << plan.width[p]) * ntiles);
            {
                prof::Scope ps(ctx, "radix_scatter", 2.0 * pair_bytes * n);
                launch_scatter<K>(ctx, threads, ntiles, ka, va, kb, vb, n, plan.shift[p], plan.width[p], th.get());
            }
            std::swap(ka, kb);
            std::swap(va, vb);
        }
        return ka != keys;
    }
    // onesweep: every pass's digit totals in one read, one launch per pass
    const unsigned hb = (unsigned)std::min<i64>(ntiles, 1024);
    DBuf<int> hist = ctx.buf<int>((i64)(hb + 1) * plan.passes * 256);
    int* hsum = hist.get() + (i64)hb * plan.passes * 256;
    {
        prof::Scope ps(ctx, "radix_hist_all", (double)sizeof(K) * n);
        k_radix_hist_all<K><<<hb, kBlock, 0, ctx.stream>>>(keys, n, plan, hist.get(), hsum, ctx.reduce_ticket());
        RAMA_LAUNCHED();
    }
    for (int p = 0; p < plan.passes; ++p) {
        u32 epoch;
        u64* st = ctx.scan_begin((i64)ntiles * 256, &epoch);
        {
            // reads and writes one (key, u32 value) pair per element
            prof::Scope ps(ctx, "radix_onesweep", 2.0 * pair_bytes * n);
            launch_onesweep<K>(ctx, threads, ntiles, ka, va, kb, vb, n, plan.shift[p], plan.width[p], hsum + p * 256,
                               st, epoch);
        }
        std::swap(ka, kb);
        std::swap(va, vb);
    }
    return ka != keys;
}

template <class K>
void radix_sort_pairs(Ctx& ctx, K* keys, u32* vals, K* tk, u32* tv, i64 n, int bits) {
    if (radix_sort_core<K>(ctx, keys, vals, tk, tv, n, bits)) {
        ctx.d2d(keys, tk, n);
        ctx.d2d(vals, tv, n);
    }
}

template <class K>
void radix_sort_pairs_shaped(Ctx& ctx, K* keys, u32* vals, K* tk, u32* tv, i64 n, int bits, int path, int threads) {
    if (radix_sort_core<K>(ctx, keys, vals, tk, tv, n, bits, path, threads)) {
        ctx.d2d(keys, tk, n);
        ctx.d2d(vals, tv, n);
    }
}

// Buffer form: an odd pass count hands over the scratch buffers instead of
// copying back.
template <class K>
void radix_sort_pairs(Ctx& ctx, DBuf<K>& keys, DBuf<u32>& vals, i64 n, int bits) {
    if (n <= 1) return;
    DBuf<K> tk = ctx.buf<K>(n);
    DBuf<u32> tv = ctx.buf<u32>(n);
    if (radix_sort_core<K>(ctx, keys.get(), vals.get(), tk.get(), tv.get(), n, bits)) {
        keys = std::move(tk);
        vals = std::move(tv);
    }
}

template void radix_sort_pairs<u32>(Ctx&, u32*, u32*, u32*, u32*, i64, int);
template void radix_sort_pairs<u64>(Ctx&, u64*, u32*, u64*, u32*, i64, int);
template void radix_sort_pairs_shaped<u32>(Ctx&, u32*, u32*, u32*, u32*, i64, int, int, int);
template void radix_sort_pairs_shaped<u64>(Ctx&, u64*, u32*, u64*, u32*, i64, int, int, int);
template void radix_sort_pairs<u32>(Ctx&, DBuf<u32>&, DBuf<u32>&, i64, int);
template void radix_sort_pairs<u64>(Ctx&, DBuf<u64>&, DBuf<u32>&, i64, int);

}  // namespace rama
