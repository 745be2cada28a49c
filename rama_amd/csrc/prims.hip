// prims.hip — stable LSD radix sort of (u64 key, u32 value) pairs.
//
// One digit of at most 8 bits per pass (digits balanced over the key),
// three launches per pass:
//   k_radix_hist     per-tile digit histogram (digit-major table)
//   scan_exclusive   table -> global (digit, tile) offsets
//   k_radix_scatter  wave64 ballot "multi-split" ranking, tile reordered in
//                    LDS so each digit run is written out contiguously.
// Tiles are 4096 pairs (256 threads x 16 items).  Items are ranked in
// (wave, item, lane) order, which is their input order, so each pass is
// stable and the whole sort is stable.
#include "prims.cuh"

namespace rama {

template <class K>
__global__ __launch_bounds__(kBlock) void k_radix_hist(const K* __restrict__ keys, i64 n, int shift, int bits,
                                                       int* hist, int ntiles) {
    __shared__ int h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const u32 mask = (1u << bits) - 1u;
    const i64 base = (i64)blockIdx.x * kRsTile;
#pragma unroll 4
    for (int k = 0; k < kRsItems; ++k) {
        const i64 i = base + (i64)k * kBlock + threadIdx.x;
        if (i < n) atomicAdd(&h[(int)((u32)(keys[i] >> shift) & mask)], 1);
    }
    __syncthreads();
    if (threadIdx.x < (1 << bits)) hist[(i64)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

template <class K>
__global__ __launch_bounds__(kBlock) void k_radix_scatter(const K* __restrict__ kin, const u32* __restrict__ vin,
                                                          K* __restrict__ kout, u32* __restrict__ vout, i64 n,
                                                          int shift, int bits, const int* __restrict__ offs,
                                                          int ntiles) {
    __shared__ K sk[kRsTile];
    __shared__ u32 sv[kRsTile];
    __shared__ int wcnt[kWaves][256];
    __shared__ int dstart[256];
    __shared__ int gbase[256];
    __shared__ int ws[kWaves];

    const int tid = threadIdx.x, lane = lane_id(), w = tid >> 6;
    const u32 mask = (1u << bits) - 1u;
    for (int k = tid; k < kWaves * 256; k += kBlock) (&wcnt[0][0])[k] = 0;
    __syncthreads();

    const i64 tile0 = (i64)blockIdx.x * kRsTile;
    const i64 wbase = tile0 + (i64)w * (64 * kRsItems);
    K key[kRsItems];
    u32 val[kRsItems];
    int rnk[kRsItems];
#pragma unroll
    for (int k = 0; k < kRsItems; ++k) {
        const i64 i = wbase + (i64)k * 64 + lane;
        const bool ok = i < n;
        key[k] = ok ? kin[i] : ~K(0);  // padding sorts into the last bucket, after all real keys
        val[k] = ok ? vin[i] : 0u;
    }
    const unsigned long long lt = (1ull << lane) - 1ull;
#pragma unroll
    for (int k = 0; k < kRsItems; ++k) {
        const u32 d = (u32)(key[k] >> shift) & mask;
        unsigned long long peers = __ballot(1);
        for (int b = 0; b < bits; ++b) {
            const unsigned long long m = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? m : ~m;
        }
        const int r = __popcll(peers & lt);
        const int c = __popcll(peers);
        const int before = wcnt[w][d];
        __builtin_amdgcn_wave_barrier();
        if (r == 0) wcnt[w][d] = before + c;
        __builtin_amdgcn_wave_barrier();
        rnk[k] = before + r;
    }
    __syncthreads();
    // per digit: exclusive over waves, then exclusive over digits
    {
        const int d = tid;
        int tot = 0;
#pragma unroll
        for (int q = 0; q < kWaves; ++q) {
            const int c = wcnt[q][d];
            wcnt[q][d] = tot;
            tot += c;
        }
        int all;
        const int st = block_excl_scan(tot, ws, &all);
        dstart[d] = st;
        if (d < (1 << bits)) gbase[d] = offs[(i64)d * ntiles + blockIdx.x];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kRsItems; ++k) {
        const u32 d = (u32)(key[k] >> shift) & mask;
        const int pos = dstart[d] + wcnt[w][d] + rnk[k];
        sk[pos] = key[k];
        sv[pos] = val[k];
    }
    __syncthreads();
    const int nvalid = (int)((n - tile0) < kRsTile ? (n - tile0) : kRsTile);
    for (int p = tid; p < nvalid; p += kBlock) {
        const K k = sk[p];
        const u32 d = (u32)(k >> shift) & mask;
        const i64 o = (i64)gbase[d] + (p - dstart[d]);
        kout[o] = k;
        vout[o] = sv[p];
    }
}

template <class K>
void radix_sort_pairs(Ctx& ctx, K* keys, u32* vals, K* tk, u32* tv, i64 n, int bits) {
    if (n <= 1 || bits <= 0) return;
    RAMA_REQUIRE(n < ((i64)1 << 31), "radix sort supports fewer than 2^31 elements");
    RAMA_REQUIRE(bits <= (int)(8 * sizeof(K)), "radix sort key wider than its type");
    const int ntiles = (int)((n + kRsTile - 1) / kRsTile);
    const int passes = (bits + 7) / 8;
    const int base_w = bits / passes, extra = bits % passes;
    int maxw = base_w + (extra ? 1 : 0);
    DBuf<int> hist = ctx.buf<int>(((i64)1 << maxw) * ntiles + 1);
    K* ka = keys;
    u32* va = vals;
    K* kb = tk;
    u32* vb = tv;
    int shift = 0;
    const double pair_bytes = (double)(sizeof(K) + sizeof(u32));
    for (int p = 0; p < passes; ++p) {
        const int wbits = base_w + (p < extra ? 1 : 0);
        {
            prof::Scope ps(ctx, "radix_hist", (double)sizeof(K) * n);
            k_radix_hist<K><<<ntiles, kBlock, 0, ctx.stream>>>(ka, n, shift, wbits, hist.get(), ntiles);
            RAMA_LAUNCHED();
        }
        scan_exclusive_arr<int>(ctx, hist.get(), hist.get(), ((i64)1 << wbits) * ntiles);
        {
            // reads and writes one (key, u32 value) pair per element
            prof::Scope ps(ctx, "radix_scatter", 2.0 * pair_bytes * n);
            k_radix_scatter<K><<<ntiles, kBlock, 0, ctx.stream>>>(ka, va, kb, vb, n, shift, wbits, hist.get(), ntiles);
            RAMA_LAUNCHED();
        }
        std::swap(ka, kb);
        std::swap(va, vb);
        shift += wbits;
    }
    if (ka != keys) {
        ctx.d2d(keys, ka, n);
        ctx.d2d(vals, va, n);
    }
}

template <class K>
void radix_sort_pairs(Ctx& ctx, DBuf<K>& keys, DBuf<u32>& vals, i64 n, int bits) {
    if (n <= 1) return;
    DBuf<K> tk = ctx.buf<K>(n);
    DBuf<u32> tv = ctx.buf<u32>(n);
    radix_sort_pairs<K>(ctx, keys.get(), vals.get(), tk.get(), tv.get(), n, bits);
}

template void radix_sort_pairs<u32>(Ctx&, u32*, u32*, u32*, u32*, i64, int);
template void radix_sort_pairs<u64>(Ctx&, u64*, u32*, u64*, u32*, i64, int);
template void radix_sort_pairs<u32>(Ctx&, DBuf<u32>&, DBuf<u32>&, i64, int);
template void radix_sort_pairs<u64>(Ctx&, DBuf<u64>&, DBuf<u32>&, i64, int);

}  // namespace rama
