// matching.hip — greedy mutual-best-neighbour matching on the symmetric
// CSR of the current graph (src/maximum_matching_vertex_based.cu:10-137).
//
// Per round: one kernel picks each unmatched vertex's strongest eligible
// neighbour (strictly above the running maximum initialised to the
// threshold, first maximum wins), one kernel matches mutual pairs, and one
// reduction counts matched vertices.  The result is deterministic: mutual
// pairs are disjoint, so the racing writes of a pair write the same values.
#include "rama_internal.hpp"

namespace rama {

// kGroup lanes per vertex: the group scans the row in kGroup-wide chunks and
// chunk winners are combined keeping the first maximum (strict >, ties to
// the lowest position), which equals the sequential scan of
// get_best_neighbour (src/maximum_matching_vertex_based.cu:10-28).
constexpr int kGroup = 8;

__global__ __launch_bounds__(kBlock) void k_pick_best(int n, float thr, const int* __restrict__ off,
                                                      const int* __restrict__ adj, const float* __restrict__ w,
                                                      const int* __restrict__ matched, int* __restrict__ best) {
    const i64 t = (i64)blockIdx.x * kBlock + threadIdx.x;
    const int v = (int)(t / kGroup);
    const int sub = (int)(t % kGroup);
    const bool active = v < n && !matched[v];
    const int b = active ? off[v] : 0, e = active ? off[v + 1] : 0;
    float mx = thr;
    int nid = -1;
    // all lanes of a group run the same number of chunks (shuffles stay uniform)
    for (int base = b; base < e; base += kGroup) {
        const int l = base + sub;
        bool ok = false;
        float cv = 0.0f;
        int cn = -1;
        if (l < e) {
            cv = w[l];
            cn = adj[l];
            ok = !(cv <= thr) && !matched[cn] && cn != v;
        }
        float bv = ok ? cv : -3.402823466e38f;
        int bp = ok ? l : 0x7fffffff;
        int bn = ok ? cn : -1;
#pragma unroll
        for (int d = kGroup / 2; d >= 1; d >>= 1) {
            const float ov = __shfl_xor(bv, d, 64);
            const int op = __shfl_xor(bp, d, 64);
            const int on = __shfl_xor(bn, d, 64);
            if (ov > bv || (ov == bv && op < bp)) {
                bv = ov;
                bp = op;
                bn = on;
            }
        }
        if (bp != 0x7fffffff && bv > mx) {
            mx = bv;
            nid = bn;
        }
    }
    if (active && sub == 0) best[v] = nid;
}

__global__ void k_match(int n, const int* __restrict__ best, int* matched, int* map, int* running) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    if (matched[v]) return;
    const int vb = best[v];
    if (vb == -1 || matched[vb]) return;
    if (best[vb] == v) {
        map[vb > v ? vb : v] = vb < v ? vb : v;
        matched[v] = 1;
        matched[vb] = 1;
        *running = 1;
    }
}

// determine_matching_threshold (src/maximum_matching_vertex_based.cu:76-86):
// mean of the non-negative costs (summed in double) times the multiplier.
static bool positive_mean_threshold(Ctx& ctx, const float* w, i64 n, float mult, float* thr) {
    DBuf<double> s = ctx.buf<double>(2);
    reduce_to<double>(
        ctx, [=] __device__(i64 e) { const float x = w[e]; return x >= 0.0f ? (double)x : 0.0; }, n, 0.0, OpSum(),
        s.get());
    reduce_to<double>(
        ctx, [=] __device__(i64 e) { return w[e] >= 0.0f ? 1.0 : 0.0; }, n, 0.0, OpSum(), s.get() + 1);
    double h[2];
    ctx.read(s.get(), h, 2);
    if (h[1] == 0.0) return false;
    const float sf = (float)h[0];
    *thr = mult * sf / (float)(i64)h[1];
    return true;
}

int matching_vertex_based(Ctx& ctx, const Csr& U, float mean_multiplier, DBuf<int>& map) {
    const int n = U.n;
    map = ctx.buf<int>(n);
    iota(ctx, map.get(), n);
    float thr = 0.0f;
    if (n == 0 || !positive_mean_threshold(ctx, U.w.get(), U.nnz, mean_multiplier, &thr) || thr < 0.0f) return 0;
    DBuf<int> best = ctx.buf<int>(n), matched = ctx.buf<int>(n), scal = ctx.buf<int>(2);
    fill(ctx, best.get(), n, -1);
    ctx.zero(matched.get(), n);
    int prev = 0;
    const int* M = matched.get();
    for (int t = 0; t < 10; ++t) {
        ctx.zero(scal.get() + 1, 1);
        const i64 threads = (i64)n * kGroup;
        {
            // offsets + (neighbour, cost, neighbour's matched flag) per CSR entry + best per vertex
            prof::Scope ps(ctx, "match_pick_best", 12.0 * U.nnz + 12.0 * n);
            k_pick_best<<<grid_for(threads), kBlock, 0, ctx.stream>>>(n, thr, U.off.get(), U.adj.get(), U.w.get(),
                                                                      matched.get(), best.get());
            RAMA_LAUNCHED();
        }
        k_match<<<grid_for(n), kBlock, 0, ctx.stream>>>(n, best.get(), matched.get(), map.get(), scal.get() + 1);
        RAMA_LAUNCHED();
        reduce_to<int>(ctx, [=] __device__(i64 v) { return M[v]; }, n, 0, OpSum(), scal.get());
        int h[2];
        ctx.read(scal.get(), h, 2);
        const int cur = h[0];
        const float rel = (cur - prev) / (prev + 1.0f);
        prev = cur;
        if (ctx.verbose) std::printf("matched sum: %d, rel_increase: %g\n", cur, rel);
        if (!h[1] || rel < 0.1f) break;
    }
    return prev;
}

}  // namespace rama
