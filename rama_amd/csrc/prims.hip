// prims.hip — stable LSD radix sort of (u64 key, u32 value) pairs.
//
// One 8-bit digit per pass, three launches per pass:
//   k_radix_hist     per-tile digit histogram (digit-major table)
//   scan_exclusive   table -> global (digit, tile) offsets
//   k_radix_scatter  wave64 ballot "multi-split" ranking, tile reordered in
//                    LDS so each digit run is written out contiguously.
// Tiles are 4096 pairs (256 threads x 16 items).  Items are ranked in
// (wave, item, lane) order, which is their input order, so each pass is
// stable and the whole sort is stable.
#include "prims.cuh"

namespace rama {

__global__ __launch_bounds__(kBlock) void k_radix_hist(const u64* __restrict__ keys, i64 n, int shift, int* hist,
                                                       int ntiles) {
    __shared__ int h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const i64 base = (i64)blockIdx.x * kRsTile;
#pragma unroll 4
    for (int k = 0; k < kRsItems; ++k) {
        const i64 i = base + (i64)k * kBlock + threadIdx.x;
        if (i < n) atomicAdd(&h[(int)((keys[i] >> shift) & 255u)], 1);
    }
    __syncthreads();
    hist[(i64)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

__global__ __launch_bounds__(kBlock) void k_radix_scatter(const u64* __restrict__ kin, const u32* __restrict__ vin,
                                                          u64* __restrict__ kout, u32* __restrict__ vout, i64 n,
                                                          int shift, const int* __restrict__ offs, int ntiles) {
    __shared__ u64 sk[kRsTile];
    __shared__ u32 sv[kRsTile];
    __shared__ int wcnt[kWaves][256];
    __shared__ int dstart[256];
    __shared__ int gbase[256];
    __shared__ int ws[kWaves];

    const int tid = threadIdx.x, lane = lane_id(), w = tid >> 6;
    for (int k = tid; k < kWaves * 256; k += kBlock) (&wcnt[0][0])[k] = 0;
    __syncthreads();

    const i64 tile0 = (i64)blockIdx.x * kRsTile;
    const i64 wbase = tile0 + (i64)w * (64 * kRsItems);
    u64 key[kRsItems];
    u32 val[kRsItems];
    int rnk[kRsItems];
#pragma unroll
    for (int k = 0; k < kRsItems; ++k) {
        const i64 i = wbase + (i64)k * 64 + lane;
        const bool ok = i < n;
        key[k] = ok ? kin[i] : ~0ull;  // padding sorts into the last bucket, after all real keys
        val[k] = ok ? vin[i] : 0u;
    }
    const unsigned long long lt = (1ull << lane) - 1ull;
#pragma unroll
    for (int k = 0; k < kRsItems; ++k) {
        const u32 d = (u32)((key[k] >> shift) & 255u);
        unsigned long long peers = __ballot(1);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
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
        gbase[d] = offs[(i64)d * ntiles + blockIdx.x];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kRsItems; ++k) {
        const u32 d = (u32)((key[k] >> shift) & 255u);
        const int pos = dstart[d] + wcnt[w][d] + rnk[k];
        sk[pos] = key[k];
        sv[pos] = val[k];
    }
    __syncthreads();
    const int nvalid = (int)((n - tile0) < kRsTile ? (n - tile0) : kRsTile);
    for (int p = tid; p < nvalid; p += kBlock) {
        const u64 k = sk[p];
        const u32 d = (u32)((k >> shift) & 255u);
        const i64 o = (i64)gbase[d] + (p - dstart[d]);
        kout[o] = k;
        vout[o] = sv[p];
    }
}

void radix_sort_pairs(Ctx& ctx, u64* keys, u32* vals, u64* tk, u32* tv, i64 n, int bits) {
    if (n <= 1 || bits <= 0) return;
    RAMA_REQUIRE(n < ((i64)1 << 31), "radix sort supports fewer than 2^31 elements");
    const int ntiles = (int)((n + kRsTile - 1) / kRsTile);
    DBuf<int> hist = ctx.buf<int>((i64)256 * ntiles + 1);
    u64* ka = keys;
    u32* va = vals;
    u64* kb = tk;
    u32* vb = tv;
    const int passes = (bits + 7) / 8;
    for (int p = 0; p < passes; ++p) {
        const int shift = 8 * p;
        {
            prof::Scope ps(ctx, "radix_hist", 8.0 * n);
            k_radix_hist<<<ntiles, kBlock, 0, ctx.stream>>>(ka, n, shift, hist.get(), ntiles);
            RAMA_LAUNCHED();
        }
        scan_exclusive_arr<int>(ctx, hist.get(), hist.get(), (i64)256 * ntiles);
        {
            // reads and writes one (u64 key, u32 value) pair per element
            prof::Scope ps(ctx, "radix_scatter", 24.0 * n);
            k_radix_scatter<<<ntiles, kBlock, 0, ctx.stream>>>(ka, va, kb, vb, n, shift, hist.get(), ntiles);
            RAMA_LAUNCHED();
        }
        std::swap(ka, kb);
        std::swap(va, vb);
    }
    if (ka != keys) {
        ctx.d2d(keys, ka, n);
        ctx.d2d(vals, va, n);
    }
}

void radix_sort_pairs(Ctx& ctx, DBuf<u64>& keys, DBuf<u32>& vals, i64 n, int bits) {
    if (n <= 1) return;
    DBuf<u64> tk = ctx.buf<u64>(n);
    DBuf<u32> tv = ctx.buf<u32>(n);
    radix_sort_pairs(ctx, keys.get(), vals.get(), tk.get(), tv.get(), n, bits);
}

}  // namespace rama
