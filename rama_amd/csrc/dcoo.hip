// dcoo.hip — the dCOO edge store: construction, contraction (relabel, sort,
// segmented reduce-by-key, diagonal check and removal), symmetric CSR,
// offsets, label compression and on-device sanitising.
//
// Reference: include/dCOO.h, src/dCOO.cu, include/rama_utils.h.
//
// Layout in HBM: SoA int32 row / int32 col / fp32 val, sorted by (row, col).
// Sorting packs (row, col) into one u64 key of 2*ceil(log2 N) bits and
// carries the fp32 cost bit pattern as the u32 payload, so a contraction
// moves 12 B per edge per radix pass.
#include "rama_internal.hpp"

namespace rama {

__device__ __forceinline__ u32 f2u(float f) { return __float_as_uint(f); }
__device__ __forceinline__ float u2f(u32 u) { return __uint_as_float(u); }

int max_id(Ctx& ctx, const int* a, const int* b, i64 n) {
    if (n <= 0) return -1;
    return reduce<int>(
        ctx, [=] __device__(i64 e) { return a[e] > b[e] ? a[e] : b[e]; }, n, -1, OpMax());
}

float dcoo_min(Ctx& ctx, const float* val, i64 n) {
    return reduce<float>(ctx, [=] __device__(i64 e) { return val[e]; }, n, 3.402823466e38f, OpMin());
}

// get_lb (include/rama_utils.h:179-190): sum of negative parts in double.
double dcoo_lb(Ctx& ctx, const float* val, i64 n) {
    if (n <= 0) return 0.0;
    return reduce<double>(
        ctx, [=] __device__(i64 e) { const float x = val[e]; return x < 0.0f ? (double)x : 0.0; }, n, 0.0,
        OpSum());
}

// Writes rows = last row + 1 and cols = max col + 1 (src/dCOO.cu:46-51).
static void derive_dims(Ctx& ctx, DCoo& A) {
    if (A.nnz == 0) {
        A.rows = A.cols = 0;
        return;
    }
    const int* col = A.col.get();
    DBuf<int> tmp = ctx.buf<int>(2);
    reduce_to<int>(ctx, [=] __device__(i64 e) { return col[e]; }, A.nnz, -1, OpMax(), tmp.get());
    ctx.d2d(tmp.get() + 1, A.row.get() + A.nnz - 1, 1);
    int h[2];
    ctx.read(tmp.get(), h, 2);
    A.cols = h[0] + 1;
    A.rows = h[1] + 1;
    A.max_node = h[0];  // rows are below their cols
}

// dCOO(col, row, data, directed=true) + init (src/dCOO.cu:28-55): orient
// each edge (min, max) and stable-sort by (row, col).  No deduplication.
DCoo dcoo_from_edges(Ctx& ctx, const int* row_in, const int* col_in, const float* val_in, i64 m) {
    DCoo A;
    A.nnz = m;
    A.row = ctx.buf<int>(m);
    A.col = ctx.buf<int>(m);
    A.val = ctx.buf<float>(m);
    if (m == 0) return A;
    const int mx = max_id(ctx, row_in, col_in, m);
    RAMA_REQUIRE(mx >= 0, "node ids must be non-negative");
    const int kb = bits_for(mx);
    DBuf<u64> keys = ctx.buf<u64>(m);
    DBuf<u32> vals = ctx.buf<u32>(m);
    u64* K = keys.get();
    u32* V = vals.get();
    DBuf<int> bad = ctx.buf<int>(1);
    ctx.zero(bad.get(), 1);
    int* badp = bad.get();
    for_each(ctx, m, [=] __device__(i64 e) {
        const int a = row_in[e], b = col_in[e];
        if (a < 0 || b < 0) badp[0] = 1;
        const u32 lo = (u32)(a < b ? a : b), hi = (u32)(a < b ? b : a);
        K[e] = ((u64)lo << kb) | hi;
        V[e] = f2u(val_in[e]);
    });
    RAMA_REQUIRE(ctx.read1(badp) == 0, "node ids must be non-negative");
    prof::Site site_("init");
    radix_sort_pairs(ctx, keys, vals, m, 2 * kb);
    K = keys.get();  // an odd pass count hands back the scratch buffers
    V = vals.get();
    int* R = A.row.get();
    int* C = A.col.get();
    float* W = A.val.get();
    const u64 mask = ((u64)1 << kb) - 1;
    for_each(ctx, m, [=] __device__(i64 e) {
        const u64 k = K[e];
        R[e] = (int)(k >> kb);
        C[e] = (int)(k & mask);
        W[e] = u2f(V[e]);
    });
    derive_dims(ctx, A);
    return A;
}

// dCOO(j, i, costs, directed, is_sorted=true) (src/dual_solver.cu:32-33).
DCoo dcoo_from_sorted(Ctx& ctx, DBuf<int>&& row, DBuf<int>&& col, DBuf<float>&& val, i64 m) {
    DCoo A;
    A.row = std::move(row);
    A.col = std::move(col);
    A.val = std::move(val);
    A.nnz = m;
    derive_dims(ctx, A);
    return A;
}

// Per-segment reduction of the sorted contraction keys (packed instances,
// which need the segments for their per-instance statistics): one thread
// per (row, col) segment sums its costs in double in sorted order and
// rounds once; self-loops are flagged (negative -> bad contraction).
// The segment count S is read from the device; the grid covers the bound n.
__global__ __launch_bounds__(kBlock) void k_contract_reduce(const u64* __restrict__ K, const u32* __restrict__ V,
                                                           const int* __restrict__ seg_start, const int* dS, i64 n,
                                                           int kb, u64* __restrict__ okey, float* __restrict__ oval,
                                                           int* __restrict__ keep, int* bad, double* diag) {
    __shared__ double ws[kWaves];
    __shared__ int wbad[kWaves];
    const i64 S = *dS;
    const i64 s = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    double dsum = 0.0;
    int isbad = 0;
    if (s < S) {
        const i64 b = seg_start[s];
        const i64 e = s + 1 < S ? seg_start[s + 1] : n;
        double acc = 0.0;
        for (i64 k = b; k < e; ++k) acc += (double)__uint_as_float(V[k]);
        const u64 key = K[b];
        const u64 mask = ((u64)1 << kb) - 1;
        const int r = (int)(key >> kb), c = (int)(key & mask);
        const float f = (float)acc;
        okey[s] = key;
        oval[s] = f;
        keep[s] = r != c;
        if (r == c) {
            isbad = f < 0.0f;
            dsum = (double)f;
        }
    }
    // one atomic per block for the self-loop total and the bad flag
    const double bsum = block_reduce(dsum, OpSum(), ws);
    const int bbad = block_reduce(isbad, OpMax(), wbad);
    if (threadIdx.x == 0) {
        if (bsum != 0.0) atomicAdd(diag, bsum);
        if (bbad) *bad = 1;
    }
}

// Sum, in order and in double, of the costs of the run of keys equal to k
// that starts at its head e (the caller rounds it once).  The head's cost
// and the next key are loaded together, so a run of one edge, the common
// case, costs one round of loads; the additions keep their order.
__device__ __forceinline__ double head_run_sum(const u64* __restrict__ K, const u32* __restrict__ V, u64 k, i64 e,
                                               i64 n) {
    const u32 v0 = V[e];
    const u64 k1 = e + 1 < n ? K[e + 1] : ~k;
    double acc = 0.0;
    acc += (double)__uint_as_float(v0);
    if (k1 == k) {
        acc += (double)__uint_as_float(V[e + 1]);
        for (i64 x = e + 2; x < n && K[x] == k; ++x) acc += (double)__uint_as_float(V[x]);
    }
    return acc;
}

// Statistics of the sorted contraction keys (grid-stride over the n sorted
// keys; a key whose row exceeds its col is a mirrored copy, left out):
// largest col, largest col of an off-diagonal edge, largest row; each
// self-loop run (the diagonal) is summed in sorted order and rounded to
// float, a negative one raises the bad-contraction flag, and their total
// goes to *diag.  Each block stores its results in `part` (4 ints) and
// `dpart`; a one-block launch combines them in block order (no atomics:
// same-address atomics from every block serialise).
__global__ __launch_bounds__(kBlock) void k_contract_stats(const u64* __restrict__ K, const u32* __restrict__ V,
                                                          i64 n, int kb, int* __restrict__ part,
                                                          double* __restrict__ dpart) {
    __shared__ int ws[kWaves];
    __shared__ double wd[kWaves];
    const u64 mask = ((u64)1 << kb) - 1;
    int all = -1, kept = -1, bad = 0, last = -1;
    double dsum = 0.0;
    const i64 stride = (i64)gridDim.x * kBlock;
    for (i64 e0 = (i64)blockIdx.x * kBlock + threadIdx.x; e0 < n; e0 += 4 * stride) {
        // four keys (and their predecessors) in flight per thread
        u64 kk[4], pk[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const i64 e = e0 + u * stride;
            kk[u] = e < n ? K[e] : 0;
            pk[u] = e < n && e > 0 ? K[e - 1] : ~(u64)0;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const i64 e = e0 + u * stride;
            if (e >= n) break;
            const u64 k = kk[u];
            const u64 hi = k >> kb;
            const int c = (int)(k & mask);
            if (hi > (u64)c) continue;  // mirrored copy or its padding
            const int r = (int)hi;
            all = c > all ? c : all;
            last = r > last ? r : last;
            if (r != c) {
                kept = c > kept ? c : kept;
            } else if (pk[u] != k) {
                double acc = 0.0;
                for (i64 x = e; x < n && K[x] == k; ++x) acc += (double)__uint_as_float(V[x]);
                const float f = (float)acc;  // the merged self-loop cost, as the edge would carry it
                bad |= f < 0.0f ? 1 : 0;
                dsum += (double)f;
            }
        }
    }
    all = block_reduce(all, OpMax(), ws);
    kept = block_reduce(kept, OpMax(), ws);
    bad = block_reduce(bad, OpMax(), ws);
    last = block_reduce(last, OpMax(), ws);
    dsum = block_reduce(dsum, OpSum(), wd);
    if (threadIdx.x == 0) {
        part[4 * blockIdx.x] = all;
        part[4 * blockIdx.x + 1] = kept;
        part[4 * blockIdx.x + 2] = bad;
        part[4 * blockIdx.x + 3] = last;
        dpart[blockIdx.x] = dsum;
    }
}

// `merged`, when set, is the merge scan's packed total (forward edges << 32
// | symmetric entries): its forward count becomes sc[1].
__global__ __launch_bounds__(kBlock) void k_contract_stats_combine(const int* __restrict__ part,
                                                                  const double* __restrict__ dpart, unsigned nb,
                                                                  const i64* merged, int* sc, double* diag,
                                                                  int* mirror) {
    __shared__ int ws[kWaves];
    __shared__ double wd[kWaves];
    int all = -1, kept = -1, bad = 0, last = -1;
    double d = 0.0;
    for (unsigned b = threadIdx.x; b < nb; b += kBlock) {
        const int x = part[4 * b], y = part[4 * b + 1], z = part[4 * b + 2], w = part[4 * b + 3];
        all = x > all ? x : all;
        kept = y > kept ? y : kept;
        bad = z > bad ? z : bad;
        last = w > last ? w : last;
        d += dpart[b];
    }
    all = block_reduce(all, OpMax(), ws);
    kept = block_reduce(kept, OpMax(), ws);
    bad = block_reduce(bad, OpMax(), ws);
    last = block_reduce(last, OpMax(), ws);
    d = block_reduce(d, OpSum(), wd);
    if (threadIdx.x == 0) {
        if (merged) sc[1] = (int)(*merged >> 32);
        sc[2] = kept;
        sc[3] = all;
        sc[4] = last;
        sc[5] = bad;
        *diag = d;
        if (mirror) {  // sc[1], the kept edge count, is final: the merge's scan ran before
            mirror[1] = sc[1];
            mirror[2] = kept;
            mirror[3] = all;
            mirror[4] = last;
            mirror[5] = bad;
        }
    }
}

// dCOO::contract_cuda + has_bad_contractions + remove_diagonal
// (src/dCOO.cu:57-95,97-142; src/rama_cuda.cu:18-30,136-140).
// Counts stay on the device until one read-back at the end.
//
// Lists of at most kSymContractMax edges also yield the symmetric CSR of the
// contracted graph (B.sym) from the same sort: each edge enters twice, as
// (lo, hi) and mirrored as (hi, lo) (a self-loop's mirror as a padding key
// above every real one), so the sorted runs of keys are the rows of the
// symmetric matrix in order.  Both copies of an edge sit in the same order
// within their runs, so a mirrored run sums to the same float as its
// forward run.  The next cycle search filters this CSR instead of sorting
// the transpose again.  (Long lists would double the largest sort of the
// round to save a smaller one.)
constexpr i64 kSymContractMax = (i64)1 << 18;

static void sym_row_offsets(Ctx& ctx, const int* rows, i64 m, int n_nodes, int* off);

DCoo contract(Ctx& ctx, const DCoo& A, const int* node_map, int max_label, bool* bad, double* diag_sum) {
    const i64 n = A.nnz;
    DCoo B;
    if (n == 0) {
        *bad = false;
        if (diag_sum) *diag_sum = 0.0;
        return B;
    }
    const int kb = bits_for(max_label < 1 ? 1 : max_label);
    const bool dbl = n <= kSymContractMax;
    const i64 nk = dbl ? 2 * n : n;
    const u64 pad = (u64)1 << (2 * kb);
    DBuf<u64> keys = ctx.buf<u64>(nk);
    DBuf<u32> vals = ctx.buf<u32>(nk);
    u64* K = keys.get();
    u32* V = vals.get();
    const int* R = A.row.get();
    const int* C = A.col.get();
    const float* W = A.val.get();
    // scalars on device: [1] kept edges, [2] max kept col, [3] max col, [4] last row, [5] bad; diag as double
    DBuf<int> sc = ctx.buf<int>(8);
    DBuf<double> dg = ctx.buf<double>(1);
    int* SC = sc.get();
    int* SM = ctx.mirror<int>();
    double* DG = dg.get();
    for_each(ctx, n, [=] __device__(i64 e) {
        const int a = node_map[R[e]], b = node_map[C[e]];
        const u32 lo = (u32)(a < b ? a : b), hi = (u32)(a < b ? b : a);
        K[e] = ((u64)lo << kb) | hi;
        V[e] = f2u(W[e]);
        if (dbl) {
            K[n + e] = lo != hi ? (((u64)hi << kb) | lo) : pad;
            V[n + e] = f2u(W[e]);
        }
        if (e == 0) {
            for (int k = 0; k < 8; ++k) SC[k] = 0;
            *DG = 0.0;
        }
    });
    prof::Site site_("contract");
    radix_sort_pairs(ctx, keys, vals, nk, dbl ? 2 * kb + 1 : 2 * kb);
    K = keys.get();  // an odd pass count hands back the scratch buffers
    V = vals.get();

    const u64 mask = ((u64)1 << kb) - 1;
    B.row = ctx.buf<int>(n);
    B.col = ctx.buf<int>(n);
    B.val = ctx.buf<float>(n);
    int* BR = B.row.get();
    int* BC = B.col.get();
    float* BW = B.val.get();
    DBuf<int> srow, sadj;
    DBuf<float> sw;
    DBuf<i64> merged;
    if (dbl) {
        // the head of each run of equal off-diagonal keys sums its run and
        // writes one symmetric entry; a forward head (row < col) also writes
        // the merged edge.  One scan numbers both: forward heads count in
        // the high half, all heads in the low half.
        prof::Site site2_("contract_merge");
        srow = ctx.buf<int>(nk);
        sadj = ctx.buf<int>(nk);
        sw = ctx.buf<float>(nk);
        merged = ctx.buf<i64>(1);
        int* SR = srow.get();
        int* SA = sadj.get();
        float* SW = sw.get();
        scan_apply<i64>(
            ctx,
            [=] __device__(i64 e) -> i64 {
                const u64 k = K[e];
                if (e > 0 && K[e - 1] == k) return 0;
                const u64 r = k >> kb, c = k & mask;
                if (r == c || r > mask) return 0;  // diagonal, padding
                return r < c ? ((i64)1 << 32) | 1 : 1;
            },
            nk,
            [=] __device__(i64 e, i64 pre, i64 flag) {
                if (!flag) return;
                const u64 k = K[e];
                const int r = (int)(k >> kb), c = (int)(k & mask);
                const float f = (float)head_run_sum(K, V, k, e, nk);
                const int p = (int)(u32)pre;
                SR[p] = r;
                SA[p] = c;
                SW[p] = f;
                if (flag >> 32) {
                    const int q = (int)(pre >> 32);
                    BR[q] = r;
                    BC[q] = c;
                    BW[q] = f;
                }
            },
            merged.get());
    } else {
        // reduce_by_key + remove_diagonal in one pass: the head of each run
        // of equal keys whose ends differ sums the run's costs in sorted
        // order (double, rounded once) and writes the merged edge
        prof::Site site2_("contract_merge");
        select_into(
            ctx, n,
            [=] __device__(i64 e) {
                const u64 k = K[e];
                return (e == 0 || K[e - 1] != k) && (k >> kb) != (k & mask);
            },
            [=] __device__(i64 e, int q) {
                const u64 k = K[e];
                BR[q] = (int)(k >> kb);
                BC[q] = (int)(k & mask);
                BW[q] = (float)head_run_sum(K, V, k, e, n);
            },
            SC + 1);
    }
    // dims as the reference computes them before removing the diagonal
    // (src/dCOO.cu:91-92), the largest endpoint among the kept edges, and
    // the self-loop check
    {
        const unsigned nb = (unsigned)std::min<i64>(std::max<i64>((nk + kBlock - 1) / kBlock, 1), kReduceMaxBlocks);
        char* scratch = static_cast<char*>(ctx.reduce_scratch((size_t)kReduceMaxBlocks * 24));
        int* part = reinterpret_cast<int*>(scratch);
        double* dpart = reinterpret_cast<double*>(scratch + (size_t)kReduceMaxBlocks * 16);
        prof::Scope ps(ctx, "contract_stats", 12.0 * nk);
        k_contract_stats<<<nb, kBlock, 0, ctx.stream>>>(K, V, nk, kb, part, dpart);
        RAMA_LAUNCHED();
        k_contract_stats_combine<<<1, kBlock, 0, ctx.stream>>>(part, dpart, nb, dbl ? merged.get() : nullptr, SC,
                                                               DG, SM);
        RAMA_LAUNCHED();
    }
    int h[8];
    ctx.read_mirror(SC, SM, h, 8);
    B.nnz = h[1];
    B.cols = h[3] + 1;
    B.rows = h[4] + 1;
    B.max_node = h[2];
    *bad = h[5] != 0;
    if (diag_sum) *diag_sum = ctx.read1(dg.get());
    if (dbl && B.nnz > 0) {
        // rows 0 .. max_dim() - 1, as the cycle search sizes its CSR
        auto G = std::make_shared<Csr>();
        G->n = B.max_dim();
        G->nnz = 2 * B.nnz;
        G->off = ctx.buf<int>((i64)G->n + 1);
        sym_row_offsets(ctx, srow.get(), G->nnz, G->n, G->off.get());
        G->adj = std::move(sadj);
        G->w = std::move(sw);
        B.sym = std::move(G);
    }
    return B;
}

// compute_offsets (include/rama_utils.h:403-413): out[v] = #{ids < v} for
// v in [0, max_value+1], one thread per row: out[r] = lower_bound(ids, r).
// Each block bounds its 256 answers by two searches and stages that window
// of ids in LDS when it is short.  (One thread per boundary of the ids would
// leave a long run of empty rows, as a sparse forest over many nodes has,
// to a single thread.)
constexpr int kOffWin = 2048;
// Sorted ids read through an accessor: an int array, the cols of the
// transpose's sort (u32) or the rows packed in the low half of 8-byte words.
struct IntIds {
    const int* p;
    __device__ __forceinline__ int operator[](i64 i) const { return p[i]; }
};
struct U32Ids {
    const u32* p;
    __device__ __forceinline__ int operator[](i64 i) const { return (int)p[i]; }
};
struct LowIds {
    const u64* p;
    __device__ __forceinline__ int operator[](i64 i) const { return (int)(u32)p[i]; }
};
template <class Ids>
__device__ __forceinline__ i64 lower_bound_ids(Ids a, i64 lo, i64 hi, int x) {
    while (lo < hi) {
        const i64 mid = (lo + hi) >> 1;
        if (a[mid] < x) lo = mid + 1; else hi = mid;
    }
    return lo;
}
template <class Ids>
__global__ __launch_bounds__(kBlock) void k_offsets_rows(Ids ids, i64 n, int end_value, int* __restrict__ out) {
    __shared__ i64 range[2];
    __shared__ int sa[kOffWin];
    const i64 r0 = (i64)blockIdx.x * kBlock;
    const i64 r1 = (r0 + kBlock - 1 < end_value ? r0 + kBlock - 1 : end_value);
    if (threadIdx.x == 0) range[0] = lower_bound_ids(ids, 0, n, (int)r0);
    if (threadIdx.x == 64) range[1] = lower_bound_ids(ids, 0, n, (int)r1);
    __syncthreads();
    const i64 lo = range[0], hi = range[1];
    const i64 r = r0 + threadIdx.x;
    if (hi - lo <= kOffWin) {  // uniform
        const int w = (int)(hi - lo);
        for (int k = threadIdx.x; k < w; k += kBlock) sa[k] = ids[lo + k];
        __syncthreads();
        if (r <= r1) {
            int a = 0, b = w;
            while (a < b) {
                const int mid = (a + b) >> 1;
                if (sa[mid] < (int)r) a = mid + 1; else b = mid;
            }
            out[r] = (int)(lo + a);
        }
    } else if (r <= r1) {
        out[r] = (int)lower_bound_ids(ids, lo, hi, (int)r);
    }
}

template <class Ids>
static void offsets_of(Ctx& ctx, Ids ids, i64 n, int max_value, int* out) {
    const int end_value = max_value + 1;
    k_offsets_rows<Ids><<<grid_for((i64)end_value + 1), kBlock, 0, ctx.stream>>>(ids, n, end_value, out);
    RAMA_LAUNCHED();
}

void compute_offsets(Ctx& ctx, const int* sorted_ids, i64 n, int max_value, int* out) {
    offsets_of(ctx, IntIds{sorted_ids}, n, max_value, out);
}

// Row offsets off[0..n_nodes] of m entries with ascending rows: each entry
// fills the rows from its predecessor's to its own, or, with far more rows
// than entries (long runs of empty rows), one thread per row.
static void sym_row_offsets(Ctx& ctx, const int* rows, i64 m, int n_nodes, int* off) {
    if ((i64)n_nodes > 4 * m) {
        offsets_of(ctx, IntIds{rows}, m, n_nodes - 1, off);
        return;
    }
    for_each(ctx, m + 1, [=] __device__(i64 i) {
        const int pr = i > 0 ? rows[i - 1] : -1, cr = i < m ? rows[i] : n_nodes;
        for (int r = pr + 1; r <= cr && r <= n_nodes; ++r) off[r] = (int)i;
    });
}

// to_undirected + dCOO(symmetric) + compute_row_offsets
// (include/rama_utils.h:107-131, src/dCOO.cu:144-147,164-172,
//  src/conflicted_cycles_cuda.cu:208-217).  With `filter`, only edges with
// cost > keep_above enter (is_positive_edge).
//
// Input: a (row, col)-sorted directed edge list with row < col.  Row r of the
// symmetric matrix is [r' of edges (r', r), ascending] ++ [c of edges (r, c),
// ascending], because every r' < r < c.  The first part is the transpose: a
// stable sort of the edges by col alone (equal cols keep the input's
// ascending rows), on 8-byte (u32 col, u32 edge) pairs.  The second part is
// the edge list itself.  Duplicated edges keep their input order, as the
// reference's stable sort of the doubled list does.
template <class Keep>
static Csr symmetric_csr_impl(Ctx& ctx, const int* row_in, const int* col_in, const float* val_in, i64 m,
                              int n_nodes, bool filter, Keep keep, i64 known_count, bool sparse_rows = false) {
    Csr G;
    G.n = n_nodes;
    i64 P = m;
    if (filter) {
        if (known_count < 0) {
            P = reduce<i64>(ctx, [=] __device__(i64 e) { return (i64)keep(e); }, m, (i64)0, OpSum());
        } else {
            P = known_count;
        }
    }
    const i64 N2 = 2 * P;
    G.nnz = N2;
    G.off = ctx.buf<int>((i64)n_nodes + 1);
    G.adj = ctx.buf<int>(N2);
    G.w = ctx.buf<float>(N2);
    if (N2 == 0) {
        ctx.zero(G.off.get(), (i64)n_nodes + 1);
        return G;
    }
    // the kept edges as (row, cost) pairs in one 8-byte word each (one
    // gather per transposed entry instead of two) and their cols, and the
    // transpose's sort input: (col, kept edge index)
    DBuf<u64> rvb = ctx.buf<u64>(P);
    DBuf<int> fc;
    DBuf<u32> tk = ctx.buf<u32>(P), tv = ctx.buf<u32>(P);
    u64* RV = rvb.get();
    u32* TK = tk.get();
    u32* TV = tv.get();
    const int* col = col_in;
    if (filter) {
        fc = ctx.buf<int>(P);
        int* FC = fc.get();
        DBuf<int> cnt = ctx.buf<int>(1);
        select_into_loaded(
            ctx, m, [=] __device__(i64 e) { return keep(e); },
            [=] __device__(i64 e) { return make_int3(row_in[e], col_in[e], __float_as_int(val_in[e])); },
            [=] __device__(i64, int q, int3 rcw) {
                RV[q] = ((u64)(u32)rcw.z << 32) | (u64)(u32)rcw.x;
                FC[q] = rcw.y;
                TK[q] = (u32)rcw.y;
                TV[q] = (u32)q;
            },
            cnt.get());
        col = FC;
    } else {
        for_each(ctx, P, [=] __device__(i64 e) {
            RV[e] = ((u64)__float_as_uint(val_in[e]) << 32) | (u64)(u32)row_in[e];
            TK[e] = (u32)col_in[e];
            TV[e] = (u32)e;
        });
    }
    {
        prof::Site site_("csr");
        radix_sort_pairs<u32>(ctx, tk, tv, P, bits_for(n_nodes > 1 ? n_nodes - 1 : 1));
    }
    TK = tk.get();  // an odd pass count hands back the scratch buffers
    TV = tv.get();
    DBuf<int> offF = ctx.buf<int>((i64)n_nodes + 1), offT = ctx.buf<int>((i64)n_nodes + 1);
    int* OFw = offF.get();
    int* OTw = offT.get();
    // row offsets of the forward part (edge rows) and of the backward part
    // (sorted cols).  Most rows hold entries: one launch over the entries,
    // each filling the rows from its predecessor's to its own.  The rows of
    // a spanning forest (its smaller ends gather at low ids), or far more
    // rows than entries, leave long runs of empty rows that would fall to
    // single threads: one thread per row there.
    if (sparse_rows || (i64)n_nodes > 4 * P) {
        offsets_of(ctx, LowIds{RV}, P, n_nodes - 1, OFw);
        offsets_of(ctx, U32Ids{TK}, P, n_nodes - 1, OTw);
    } else {
        for_each(ctx, P + 1, [=] __device__(i64 i) {
            const int pr = i > 0 ? (int)(u32)RV[i - 1] : -1, cr = i < P ? (int)(u32)RV[i] : n_nodes;
            for (int r = pr + 1; r <= cr && r <= n_nodes; ++r) OFw[r] = (int)i;
            const int pc = i > 0 ? (int)TK[i - 1] : -1, cc = i < P ? (int)TK[i] : n_nodes;
            for (int r = pc + 1; r <= cc && r <= n_nodes; ++r) OTw[r] = (int)i;
        });
    }
    int* OFF = G.off.get();
    const int* OF = OFw;
    const int* OT = OTw;
    int* ADJ = G.adj.get();
    float* WW = G.w.get();
    // one launch: row offsets of the symmetric matrix; backward part of row
    // r (transposed entries, in order); forward part (the row's segment)
    for_each(ctx, std::max<i64>(P, (i64)n_nodes + 1), [=] __device__(i64 i) {
        if (i <= n_nodes) OFF[i] = OF[i] + OT[i];
        if (i < P) {
            {
                const int r = (int)TK[i];
                const u64 rv = RV[TV[i]];
                const i64 o = (i64)OF[r] + OT[r] + (i - OT[r]);
                ADJ[o] = (int)(u32)rv;
                WW[o] = __uint_as_float((u32)(rv >> 32));
            }
            {
                const u64 rv = RV[i];
                const int r = (int)(u32)rv;
                const i64 o = (i64)OF[r] + OT[r] + (OT[r + 1] - OT[r]) + (i - OF[r]);
                ADJ[o] = col[i];
                WW[o] = __uint_as_float((u32)(rv >> 32));
            }
        }
    });
    return G;
}

Csr symmetric_csr(Ctx& ctx, const int* row, const int* col, const float* val, i64 m, int n_nodes, bool filter,
                  float keep_above, i64 known_count) {
    return symmetric_csr_impl(ctx, row, col, val, m, n_nodes, filter,
                              [=] __device__(i64 e) { return val[e] > keep_above; }, known_count);
}

Csr symmetric_csr_forest(Ctx& ctx, const int* row, const int* col, const float* val, i64 m, int n_nodes) {
    return symmetric_csr_impl(ctx, row, col, val, m, n_nodes, false, [=] __device__(i64) { return true; }, -1, true);
}

// Stable filter of a CSR's entries: position of every kept entry by one
// scan, then row offsets and kept entries in one launch.  Each row keeps its
// order, so filtering the full symmetric CSR by cost gives exactly the
// symmetric CSR of the edges that pass.
Csr csr_filter(Ctx& ctx, const Csr& F, float keep_above, i64 kept) {
    Csr G;
    G.n = F.n;
    G.nnz = kept;
    G.off = ctx.buf<int>((i64)F.n + 1);
    G.adj = ctx.buf<int>(kept > 0 ? kept : 1);
    G.w = ctx.buf<float>(kept > 0 ? kept : 1);
    if (F.nnz == 0) {
        ctx.zero(G.off.get(), (i64)F.n + 1);
        return G;
    }
    const i64 nnz = F.nnz;
    const int n = F.n;
    DBuf<int> pos = ctx.buf<int>(nnz + 1);
    int* P = pos.get();
    const int* FO = F.off.get();
    const int* FA = F.adj.get();
    const float* FW = F.w.get();
    int* GO = G.off.get();
    int* GA = G.adj.get();
    float* GW = G.w.get();
    scan_exclusive<int>(ctx, [=] __device__(i64 i) { return (int)(FW[i] > keep_above); }, P, nnz);
    for_each(ctx, std::max<i64>(nnz, (i64)n + 1), [=] __device__(i64 i) {
        if (i <= n) GO[i] = P[FO[i]];
        if (i < nnz && FW[i] > keep_above) {
            GA[P[i]] = FA[i];
            GW[P[i]] = FW[i];
        }
    });
    return G;
}

// Packed instances: edge e is kept when its cost exceeds the threshold of
// its instance, keep_above[nslot[row[e]]].
Csr symmetric_csr_slots(Ctx& ctx, const int* row, const int* col, const float* val, i64 m, int n_nodes,
                        const int* nslot, const float* keep_above, i64 known_count) {
    return symmetric_csr_impl(ctx, row, col, val, m, n_nodes, true,
                              [=] __device__(i64 e) { return val[e] > keep_above[nslot[row[e]]]; }, known_count);
}

// compress_label_sequence (include/rama_utils.h:160-177); returns the number
// of distinct labels.
int compress_labels(Ctx& ctx, const int* data, i64 n, int max_label, int* out, bool want_count) {
    DBuf<int> mask = ctx.buf<int>((i64)max_label + 2);
    ctx.zero(mask.get(), (i64)max_label + 2);
    int* M = mask.get();
    for_each(ctx, n, [=] __device__(i64 k) { M[data[k]] = 1; });
    scan_exclusive_arr<int>(ctx, M, M, (i64)max_label + 1);
    for_each(ctx, n, [=] __device__(i64 k) { out[k] = M[data[k]]; });
    return want_count ? ctx.read1(M + max_label + 1) : -1;
}

// compress_label_sequence of a matching map (map[v] = v, or the smaller end
// of v's pair): the labels in use are exactly the v with map[v] == v, so
// their ranks come from one scan without marking a mask first.  Equal to
// compress_labels(map, n, n - 1, out).
void compress_matching(Ctx& ctx, const int* map, i64 n, int* out) {
    if (n <= 0) return;
    DBuf<int> rank = ctx.buf<int>(n + 1);
    int* RK = rank.get();
    scan_exclusive<int>(ctx, [=] __device__(i64 v) { return (int)(map[v] == (int)v); }, RK, n);
    for_each(ctx, n, [=] __device__(i64 v) { out[v] = RK[map[v]]; });
}

// ------------------------------------------------------------------
// Packed instances: per-instance statistics of a (row, col)-sorted edge
// list whose node ids are global (instance q owns ids from nbase[q]).
// Local ids are global - nbase[q].  Fields as the single-instance dCOO keeps
// them: cols = max col + 1, rows = last row + 1, max_node = largest kept col.
void slot_stats_clear(Ctx& ctx, SlotStats* st, int Q) {
    for_each(ctx, Q, [=] __device__(i64 q) { st[q] = SlotStats{-1, -1, -1, 0, 0}; });
}

// rb[q] = first position of slot q in a sorted key array (q = 0..Q), the
// slot of key k being the last q with base[q] <= k.
template <class K, class Key>
__global__ void k_slot_bounds(Key key, i64 n, const int* __restrict__ base, int Q, i64* __restrict__ rb) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q > Q) return;
    const K x = (K)base[q];
    i64 lo = 0, hi = n;
    while (lo < hi) {
        const i64 mid = (lo + hi) >> 1;
        if ((K)key(mid) < x) lo = mid + 1; else hi = mid;
    }
    rb[q] = lo;
}

struct MaxAcc {
    int v = -1;
};
struct MaxRed {
    __device__ MaxAcc operator()(MaxAcc a, MaxAcc b) const { return MaxAcc{a.v > b.v ? a.v : b.v}; }
    __device__ MaxAcc shfl_down(MaxAcc a, int d) const { return MaxAcc{__shfl_down(a.v, d, 64)}; }
};

// Stats of a sorted edge list (every edge kept): ranges by binary search,
// the largest col by a range reduction, the last row and count from the
// ranges.
void slot_stats_edges(Ctx& ctx, const int* row, const int* col, i64 n, const int* nbase, int Q, SlotStats* st) {
    DBuf<i64> rbuf = ctx.buf<i64>(Q + 1);
    i64* RB = rbuf.get();
    auto rowkey = [=] __device__(i64 i) { return row[i]; };
    k_slot_bounds<int><<<grid_for(Q + 1), kBlock, 0, ctx.stream>>>(rowkey, n, nbase, Q, RB);
    RAMA_LAUNCHED();
    slot_stats_clear(ctx, st, Q);
    range_reduce<MaxAcc>(
        ctx, RB, Q, [=] __device__(i64 i, MaxAcc& a) { a.v = col[i] > a.v ? col[i] : a.v; }, MaxRed(),
        [=] __device__(int q, MaxAcc r) { atomicMax(&st[q].max_col, r.v - nbase[q]); });
    for_each(ctx, Q, [=] __device__(i64 q) {
        const i64 lo = RB[q], hi = RB[q + 1];
        st[q].max_kept_col = st[q].max_col;
        st[q].max_row = hi > lo ? row[hi - 1] - nbase[q] : -1;
        st[q].kept = (int)(hi - lo);
    });
}

struct ContractAcc {
    int mc = -1;    // largest col, diagonal included
    int mk = -1;    // largest col of a kept segment
    int kept = 0;   // kept segments
    int bad = 0;    // negative self-loop
};
struct ContractRed {
    __device__ ContractAcc operator()(ContractAcc a, ContractAcc b) const {
        return ContractAcc{a.mc > b.mc ? a.mc : b.mc, a.mk > b.mk ? a.mk : b.mk, a.kept + b.kept, a.bad | b.bad};
    }
    __device__ ContractAcc shfl_down(ContractAcc a, int d) const {
        return ContractAcc{__shfl_down(a.mc, d, 64), __shfl_down(a.mk, d, 64), __shfl_down(a.kept, d, 64),
                           __shfl_down(a.bad, d, 64)};
    }
};

// contract for packed instances: node_map gives every node of a continuing
// instance its new global id (new_nslot / new_nbase describe the new id
// space).  B receives the contracted, diagonal-free union; st[q] the new
// dims, edge count and bad-contraction flag of instance q.
void contract_packed(Ctx& ctx, const DCoo& A, const int* node_map, int max_label, const int* new_nslot,
                     const int* new_nbase, int Q, DCoo& B, SlotStats* st) {
    const i64 n = A.nnz;
    B = DCoo();
    slot_stats_clear(ctx, st, Q);
    if (n == 0) return;
    const int kb = bits_for(max_label < 1 ? 1 : max_label);
    DBuf<u64> keys = ctx.buf<u64>(n);
    DBuf<u32> vals = ctx.buf<u32>(n);
    u64* K = keys.get();
    u32* V = vals.get();
    const int* R = A.row.get();
    const int* C = A.col.get();
    const float* W = A.val.get();
    DBuf<int> sc = ctx.buf<int>(8);
    DBuf<double> dg = ctx.buf<double>(1);
    int* SC = sc.get();
    double* DG = dg.get();
    for_each(ctx, n, [=] __device__(i64 e) {
        const int a = node_map[R[e]], b = node_map[C[e]];
        const u32 lo = (u32)(a < b ? a : b), hi = (u32)(a < b ? b : a);
        K[e] = ((u64)lo << kb) | hi;
        V[e] = f2u(W[e]);
        if (e == 0) {
            for (int k = 0; k < 8; ++k) SC[k] = 0;
            *DG = 0.0;
        }
    });
    {
        prof::Site site_("contract");
        radix_sort_pairs(ctx, keys, vals, n, 2 * kb);
    }
    K = keys.get();
    V = vals.get();
    DBuf<int> seg_start = ctx.buf<int>(n);
    int* SS = seg_start.get();
    select_into(
        ctx, n, [=] __device__(i64 e) { return e == 0 || K[e] != K[e - 1]; },
        [=] __device__(i64 e, int pos) { SS[pos] = (int)e; }, SC + 0);
    DBuf<u64> okey = ctx.buf<u64>(n);
    DBuf<float> oval = ctx.buf<float>(n);
    DBuf<int> keep = ctx.buf<int>(n);
    k_contract_reduce<<<grid_for(n), kBlock, 0, ctx.stream>>>(K, V, SS, SC + 0, n, kb, okey.get(), oval.get(),
                                                                keep.get(), SC + 5, DG);
    RAMA_LAUNCHED();
    const u64 mask = ((u64)1 << kb) - 1;
    const u64* OK = okey.get();
    const float* OV = oval.get();
    const int* KP = keep.get();
    // per-instance dims over all segments (diagonal included), kept edges
    // and bad contractions (a negative self-loop): segment ranges of each
    // instance by binary search, then range reductions
    {
        DBuf<i64> rbuf = ctx.buf<i64>(Q + 1);
        i64* RB = rbuf.get();
        const int* NB = new_nbase;
        auto rowkey = [=] __device__(i64 s) -> i64 { return s < SC[0] ? (i64)(OK[s] >> kb) : ((i64)1 << 40); };
        k_slot_bounds<i64><<<grid_for(Q + 1), kBlock, 0, ctx.stream>>>(rowkey, n, NB, Q, RB);
        RAMA_LAUNCHED();
        range_reduce<ContractAcc>(
            ctx, RB, Q,
            [=] __device__(i64 s, ContractAcc& a) {
                const int gc = (int)(OK[s] & mask);
                const bool kp = KP[s] != 0;
                a.mc = gc > a.mc ? gc : a.mc;
                if (kp) {
                    a.mk = gc > a.mk ? gc : a.mk;
                    a.kept += 1;
                } else if (OV[s] < 0.0f) {
                    a.bad = 1;
                }
            },
            ContractRed(),
            [=] __device__(int q, ContractAcc r) {
                atomicMax(&st[q].max_col, r.mc - NB[q]);
                if (r.mk >= 0) atomicMax(&st[q].max_kept_col, r.mk - NB[q]);
                atomicAdd(&st[q].kept, r.kept);
                if (r.bad) st[q].bad = 1;
            });
        for_each(ctx, Q, [=] __device__(i64 q) {
            const i64 lo = RB[q], hi = RB[q + 1];
            st[q].max_row = hi > lo ? (int)(OK[hi - 1] >> kb) - NB[q] : -1;
        });
    }
    B.row = ctx.buf<int>(n);
    B.col = ctx.buf<int>(n);
    B.val = ctx.buf<float>(n);
    int* BR = B.row.get();
    int* BC = B.col.get();
    float* BW = B.val.get();
    select_into(
        ctx, n, [=] __device__(i64 s) { return s < SC[0] && KP[s] != 0; },
        [=] __device__(i64 s, int q) {
            BR[q] = (int)(OK[s] >> kb);
            BC[q] = (int)(OK[s] & mask);
            BW[q] = OV[s];
        },
        SC + 1);
    B.nnz = -1;  // the caller sums the instances' kept counts
}

// map_node_labels (src/rama_cuda.cu:32-52)
void map_node_labels(Ctx& ctx, const int* cur, i64 cur_size, int* orig, i64 orig_size) {
    for_each(ctx, orig_size, [=] __device__(i64 n) {
        const int m = orig[n];
        if (m < cur_size) orig[n] = cur[m];
    });
}

// compute_sanitized_graph (include/rama_utils.h:429-466): orient, sort,
// drop repeated edges (first kept), renumber used node ids consecutively.
Sanitized sanitize_graph(Ctx& ctx, DBuf<int>& i, DBuf<int>& j, DBuf<float>& c, i64& m) {
    Sanitized s;
    if (m == 0) return s;
    const int mx = max_id(ctx, i.get(), j.get(), m);
    const int kb = bits_for(mx);
    DBuf<u64> keys = ctx.buf<u64>(m);
    DBuf<u32> vals = ctx.buf<u32>(m);
    u64* K = keys.get();
    u32* V = vals.get();
    const int* I = i.get();
    const int* J = j.get();
    const float* Cc = c.get();
    for_each(ctx, m, [=] __device__(i64 e) {
        const u32 a = (u32)I[e], b = (u32)J[e];
        K[e] = ((u64)(a < b ? a : b) << kb) | (a < b ? b : a);
        V[e] = f2u(Cc[e]);
    });
    prof::Site site_("sanitize");
    radix_sort_pairs(ctx, keys, vals, m, 2 * kb);
    K = keys.get();  // an odd pass count hands back the scratch buffers
    V = vals.get();
    DBuf<int> pos;
    const i64 m2 = select_positions(
        ctx, [=] __device__(i64 e) { return (int)(e == 0 || K[e] != K[e - 1]); }, m, pos);
    DBuf<int> ni = ctx.buf<int>(m2), nj = ctx.buf<int>(m2);
    DBuf<float> nc = ctx.buf<float>(m2);
    s.size = (i64)mx + 1;
    s.ids = ctx.buf<int>((i64)mx + 2);
    ctx.zero(s.ids.get(), (i64)mx + 2);
    int* NI = ni.get();
    int* NJ = nj.get();
    float* NC = nc.get();
    int* IDS = s.ids.get();
    const int* PP = pos.get();
    const u64 mask = ((u64)1 << kb) - 1;
    for_each(ctx, m, [=] __device__(i64 e) {
        if (e == 0 || K[e] != K[e - 1]) {
            const int q = PP[e];
            const int a = (int)(K[e] >> kb), b = (int)(K[e] & mask);
            NI[q] = a;
            NJ[q] = b;
            NC[q] = u2f(V[e]);
            IDS[a] = 1;
            IDS[b] = 1;
        }
    });
    scan_exclusive_arr<int>(ctx, IDS, IDS, (i64)mx + 1);
    for_each(ctx, m2, [=] __device__(i64 e) {
        NI[e] = IDS[NI[e]];
        NJ[e] = IDS[NJ[e]];
    });
    i = std::move(ni);
    j = std::move(nj);
    c = std::move(nc);
    m = m2;
    return s;
}

// desanitize_node_labels (include/rama_utils.h:468-493); out has s.size entries.
void desanitize_labels(Ctx& ctx, const Sanitized& s, const int* labels, int* out) {
    const int* IDS = s.ids.get();
    const i64 sz = s.size;
    for_each(ctx, sz, [=] __device__(i64 n) {
        const int cur = IDS[n];
        out[n] = (n == sz - 1 || cur != IDS[n + 1]) ? labels[cur] : -1;
    });
}

}  // namespace rama
