// prims.cuh — data-parallel building blocks written for CDNA (wave64):
// wave/block scans and reductions, device-wide exclusive scan, deterministic
// reductions, stream compaction and a stable LSD radix sort of
// (u64 key, u32 value) pairs.  Everything runs on the context's stream.
#pragma once

#include <cstring>

#include "common.hpp"
#include "profiler.hpp"

namespace rama {

constexpr int kBlock = 256;  // 4 waves of 64
constexpr int kWaves = kBlock / 64;

__device__ __forceinline__ int lane_id() { return __lane_id(); }

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
constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;

template <class T, class F>
__global__ __launch_bounds__(kBlock) void k_scan_partials(F f, i64 n, T* partials) {
    __shared__ T ws[kWaves];
    const i64 base = (i64)blockIdx.x * kScanTile;
    T s = T(0);
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const i64 i = base + (i64)k * kBlock + threadIdx.x;
        if (i < n) s += (T)f(i);
    }
    T r = block_reduce(s, OpSum(), ws);
    if (threadIdx.x == 0) partials[blockIdx.x] = r;
}

// out[i] = sum_{k<i} f(k) for i in [0, n]; blocked per thread.
template <class T, class F>
__global__ __launch_bounds__(kBlock) void k_scan_final(F f, i64 n, const T* offs, T* out) {
    __shared__ T ws[kWaves];
    const i64 base = (i64)blockIdx.x * kScanTile + (i64)threadIdx.x * kScanItems;
    T v[kScanItems];
    T s = T(0);
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const i64 i = base + k;
        v[k] = i < n ? (T)f(i) : T(0);
        s += v[k];
    }
    T tot;
    const T boff = offs ? offs[blockIdx.x] : T(0);
    T run = block_excl_scan(s, ws, &tot) + boff;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const i64 i = base + k;
        if (i < n) out[i] = run;
        run += v[k];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == kBlock - 1) out[n] = boff + tot;
}

template <class T>
struct LoadFn {
    const T* p;
    __device__ __forceinline__ T operator()(i64 i) const { return p[i]; }
};

// Exclusive scan of f(0..n-1) into out[0..n]; out[n] = total.  `out` may
// alias the storage f reads (each tile is read completely before written).
template <class T, class F>
void scan_exclusive(Ctx& ctx, F f, T* out, i64 n) {
    if (n <= 0) {
        ctx.zero(out, 1);
        return;
    }
    const i64 nb = (n + kScanTile - 1) / kScanTile;
    if (nb == 1) {
        k_scan_final<T, F><<<1, kBlock, 0, ctx.stream>>>(f, n, (const T*)nullptr, out);
        RAMA_LAUNCHED();
        return;
    }
    DBuf<T> part = ctx.buf<T>(nb + 1);
    k_scan_partials<T, F><<<(unsigned)nb, kBlock, 0, ctx.stream>>>(f, n, part.get());
    RAMA_LAUNCHED();
    scan_exclusive<T>(ctx, LoadFn<T>{part.get()}, part.get(), nb);
    k_scan_final<T, F><<<(unsigned)nb, kBlock, 0, ctx.stream>>>(f, n, part.get(), out);
    RAMA_LAUNCHED();
}

template <class T>
void scan_exclusive_arr(Ctx& ctx, const T* in, T* out, i64 n) {
    scan_exclusive<T>(ctx, LoadFn<T>{in}, out, n);
}

// ------------------------------------------------------------- reduction
template <class T, class Op, class F>
__global__ __launch_bounds__(kBlock) void k_reduce(F f, i64 n, T ident, Op op, T* out) {
    __shared__ T ws[kWaves];
    T acc = ident;
    for (i64 i = (i64)blockIdx.x * kBlock + threadIdx.x; i < n; i += (i64)gridDim.x * kBlock) acc = op(acc, (T)f(i));
    T r = block_reduce(acc, op, ws);
    if (threadIdx.x == 0) out[blockIdx.x] = r;
}

// Deterministic reduction (fixed grid, fixed tree); result stays on device in *dst.
template <class T, class Op, class F>
void reduce_to(Ctx& ctx, F f, i64 n, T ident, Op op, T* dst) {
    const unsigned nb = (unsigned)std::min<i64>(std::max<i64>((n + kBlock - 1) / kBlock, 1), 1024);
    if (nb == 1) {
        k_reduce<T, Op, F><<<1, kBlock, 0, ctx.stream>>>(f, n, ident, op, dst);
        RAMA_LAUNCHED();
        return;
    }
    DBuf<T> part = ctx.buf<T>(nb);
    k_reduce<T, Op, F><<<nb, kBlock, 0, ctx.stream>>>(f, n, ident, op, part.get());
    RAMA_LAUNCHED();
    k_reduce<T, Op, LoadFn<T>><<<1, kBlock, 0, ctx.stream>>>(LoadFn<T>{part.get()}, (i64)nb, ident, op, dst);
    RAMA_LAUNCHED();
}

template <class T, class Op, class F>
T reduce(Ctx& ctx, F f, i64 n, T ident, Op op) {
    DBuf<T> r = ctx.buf<T>(1);
    reduce_to(ctx, f, n, ident, op, r.get());
    return ctx.read1(r.get());
}

// ------------------------------------------------------- simple kernels
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

// Stable compaction helper: pos[i] = #{k < i : pred(k)}, returns the count.
template <class P>
i64 select_positions(Ctx& ctx, P pred, i64 n, DBuf<int>& pos) {
    pos = ctx.buf<int>(n + 1);
    scan_exclusive<int>(ctx, pred, pos.get(), n);
    return ctx.read1(pos.get() + n);
}

// ------------------------------------------------------------ radix sort
constexpr int kRsItems = 16;
constexpr int kRsTile = kBlock * kRsItems;  // 4096 keys per tile

__global__ __launch_bounds__(kBlock) void k_radix_hist(const u64* __restrict__ keys, i64 n, int shift, int* hist,
                                                       int ntiles);
__global__ __launch_bounds__(kBlock) void k_radix_scatter(const u64* __restrict__ kin, const u32* __restrict__ vin,
                                                          u64* __restrict__ kout, u32* __restrict__ vout, i64 n,
                                                          int shift, const int* __restrict__ offs, int ntiles);

// Stable LSD sort of (keys, vals) by the low `bits` bits of the key.
// `keys`/`vals` hold the result; `tk`/`tv` are scratch of the same length.
void radix_sort_pairs(Ctx& ctx, u64* keys, u32* vals, u64* tk, u32* tv, i64 n, int bits);

// Convenience: allocates the scratch itself.
void radix_sort_pairs(Ctx& ctx, DBuf<u64>& keys, DBuf<u32>& vals, i64 n, int bits);

inline int bits_for(i64 max_value) {
    int b = 1;
    while (b < 63 && ((i64)1 << b) <= max_value) ++b;
    return b;
}

}  // namespace rama
