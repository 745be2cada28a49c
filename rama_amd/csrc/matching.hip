// matching.hip — greedy mutual-best-neighbour matching on the symmetric
// CSR of the current graph (src/maximum_matching_vertex_based.cu:10-137).
//
// Per round: one kernel picks each unmatched vertex's strongest eligible
// neighbour (strictly above the running maximum initialised to the
// threshold, first maximum wins) and one kernel matches mutual pairs and
// counts them.  The loop's stop rule runs on the device, so all ten rounds
// are queued at once and the host reads the outcome once.  The result is
// deterministic: mutual pairs are disjoint, so racing writes agree.
#include "rama_internal.hpp"
#include "tuning.hpp"

namespace rama {

// kGroup lanes per vertex: the group scans the row in kGroup-wide chunks and
// chunk winners are combined keeping the first maximum (strict >, ties to
// the lowest position), which equals the sequential scan of
// get_best_neighbour (src/maximum_matching_vertex_based.cu:10-28).
constexpr int kGroup = 8;
constexpr int kMaxRounds = 10;

// Device state of the matching loop: threshold inputs and, per round, the
// number of newly matched vertices and whether any pair matched.
struct MatchState {
    double pos_sum;
    double pos_cnt;
    int newly[kMaxRounds];
    int running[kMaxRounds];
};

// Threshold (determine_matching_threshold, :76-86) and whether round t runs:
// the stop rule `!still_running || rel_increase < 0.1` (:121-127) replayed
// from the per-round counts, identically in every block.
__device__ __forceinline__ bool match_round_active(const MatchState* st, float mult, int t, float* thr) {
    if (st->pos_cnt == 0.0) return false;
    *thr = mult * (float)st->pos_sum / (float)(i64)st->pos_cnt;
    if (*thr < 0.0f) return false;
    int prev = 0;
    for (int r = 0; r < t; ++r) {
        const int cur = prev + st->newly[r];
        const float rel = (cur - prev) / (prev + 1.0f);
        if (!st->running[r] || rel < 0.1f) return false;
        prev = cur;
    }
    return true;
}

// State of the single-instance loop.  The per-round counts are spread over
// kParts counters on separate 128-byte lines: device-scope atomics on one
// address from thousands of blocks serialise, on distinct lines they run in
// parallel.  `head.newly` receives the per-round totals for the host.
constexpr int kParts = 16;
constexpr int kLine = 32;  // ints per 128-byte line
struct MatchHead {
    double pos_sum;
    double pos_cnt;
    int newly[kMaxRounds];
};
struct MatchLoop {
    MatchHead head;
    alignas(128) int part[kMaxRounds][kParts * kLine];
};

// Round t's count, summed over the parts; wave-uniform, all lanes active.
__device__ __forceinline__ int loop_newly(const MatchLoop* st, int t) {
    const int lane = lane_id();
    int x = lane < kParts ? st->part[t][lane * kLine] : 0;
#pragma unroll
    for (int d = kParts / 2; d >= 1; d >>= 1) x += __shfl_xor(x, d, 64);
    return __shfl(x, 0, 64);
}

// match_round_active for the single-instance loop (same rule; a round with
// no new pair has `still_running` false).  Called by whole waves.
__device__ __forceinline__ bool loop_round_active(const MatchLoop* st, float mult, int t, float* thr) {
    if (st->head.pos_cnt == 0.0) return false;
    *thr = mult * (float)st->head.pos_sum / (float)(i64)st->head.pos_cnt;
    if (*thr < 0.0f) return false;
    int prev = 0;
    for (int r = 0; r < t; ++r) {
        const int nw = loop_newly(st, r);
        const int cur = prev + nw;
        const float rel = (cur - prev) / (prev + 1.0f);
        if (nw == 0 || rel < 0.1f) return false;
        prev = cur;
    }
    return true;
}

// Bounded grids (tuning table: pick_blocks, match_blocks): the rounds after
// the loop has stopped cost a launch each.

// Grid-stride over (vertex, lane) pairs.  A vertex's flag and offsets are
// loaded together (not the offsets after the flag), and each lane takes two
// entries per step (positions l and l + kGroup, the earlier one winning
// ties), so a row of up to 2·kGroup entries costs one chain of dependent
// loads: offsets, then entries, then the neighbours' flags.
__global__ __launch_bounds__(kBlock) void k_pick_best(int n, float mult, int round, const MatchLoop* __restrict__ st,
                                                      const int* __restrict__ off, const int* __restrict__ adj,
                                                      const float* __restrict__ w, const int* __restrict__ matched,
                                                      int* __restrict__ best) {
    float thr;
    if (!loop_round_active(st, mult, round, &thr)) return;  // uniform over the grid
    const i64 total = (i64)n * kGroup;
    const i64 stride = (i64)gridDim.x * kBlock;
    for (i64 t = (i64)xcd_block() * kBlock + threadIdx.x; t < total; t += stride) {
        const int v = (int)(t / kGroup);
        const int sub = (int)(t % kGroup);
        const int mv = matched[v];
        const int b0 = off[v], e0 = off[v + 1];
        const bool active = !mv;
        const int b = active ? b0 : 0, e = active ? e0 : 0;
        float mx = thr;
        int nid = -1;
        // all lanes of a group run the same number of steps (shuffles stay uniform)
        for (int base = b; base < e; base += 2 * kGroup) {
            const int l0 = base + sub, l1 = l0 + kGroup;
            float c0 = 0.0f, c1 = 0.0f;
            int n0 = -1, n1 = -1;
            if (l0 < e) {
                c0 = w[l0];
                n0 = adj[l0];
            }
            if (l1 < e) {
                c1 = w[l1];
                n1 = adj[l1];
            }
            const int m0 = l0 < e ? matched[n0] : 1;
            const int m1 = l1 < e ? matched[n1] : 1;
            const bool ok0 = l0 < e && !(c0 <= thr) && !m0 && n0 != v;
            const bool ok1 = l1 < e && !(c1 <= thr) && !m1 && n1 != v;
            float bv = ok0 ? c0 : -3.402823466e38f;
            int bp = ok0 ? l0 : 0x7fffffff;
            int bn = ok0 ? n0 : -1;
            if (ok1 && (!ok0 || c1 > c0)) {
                bv = c1;
                bp = l1;
                bn = n1;
            }
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
}

// match_neighbours (:49-74).  A pair matches when both ends are unmatched
// and each is the other's best; such pairs are disjoint.  Only the smaller
// end of a pair acts (writes the map, both flags and counts 2 vertices).
// The larger end may read its own flag while the smaller end sets it, and
// skips either way (its best is smaller than itself): the matched set and
// the count are the same in any execution order.
__global__ __launch_bounds__(kBlock) void k_match(int n, float mult, int round, MatchLoop* st,
                                                  const int* __restrict__ best, int* matched, int* map) {
    __shared__ int ws[kWaves];
    float thr;
    if (!loop_round_active(st, mult, round, &thr)) return;
    int add = 0;
    const int stride = (int)gridDim.x * kBlock;
    for (int v = (int)xcd_block() * kBlock + threadIdx.x; v < n; v += stride) {
        const int mv = matched[v];
        const int vb = best[v];
        // vb's flag need not be read: a vertex matched in an earlier round
        // keeps its partner in best[] (pick_best skips matched vertices), so
        // best[vb] == v with v unmatched means vb is unmatched too
        if (!mv && vb > v && best[vb] == v) {
            map[vb] = v;
            matched[v] = 1;
            matched[vb] = 1;
            add += 2;
        }
    }
    const int tot = block_reduce(add, OpSum(), ws);
    if (threadIdx.x == 0 && tot) atomicAdd(&st->part[round][(blockIdx.x % kParts) * kLine], tot);
}

// head.newly[r] = round r's count, for the host (one wave); the whole head
// is also stored at `mirror` when that is set.
__global__ void k_match_totals(MatchLoop* st, MatchHead* mirror) {
    for (int r = 0; r < kMaxRounds; ++r) {
        const int x = loop_newly(st, r);
        if (threadIdx.x == 0) {
            st->head.newly[r] = x;
            if (mirror) mirror->newly[r] = x;
        }
    }
    if (mirror && threadIdx.x == 0) {
        mirror->pos_sum = st->head.pos_sum;
        mirror->pos_cnt = st->head.pos_cnt;
    }
}

int matching_vertex_based(Ctx& ctx, const Csr& U, float mean_multiplier, DBuf<int>& map) {
    const int n = U.n;
    map = ctx.buf<int>(n);
    if (n == 0) return 0;
    DBuf<MatchLoop> stb = ctx.buf<MatchLoop>(1);
    MatchLoop* st = stb.get();
    DBuf<int> best = ctx.buf<int>(n), matched = ctx.buf<int>(n);
    {
        // one launch: identity mapping, no best neighbour, nothing matched,
        // cleared loop state (the first int of each counter line is the counter)
        int* M = map.get();
        int* B = best.get();
        int* D = matched.get();
        constexpr int kCounters = kMaxRounds * kParts;
        for_each(ctx, std::max<i64>(n, kCounters), [=] __device__(i64 v) {
            if (v < n) {
                M[v] = (int)v;
                B[v] = -1;
                D[v] = 0;
            }
            if (v < kCounters) st->part[v / kParts][(v % kParts) * kLine] = 0;
            if (v == 0) {
                st->head.pos_sum = 0.0;
                st->head.pos_cnt = 0.0;
            }
        });
    }
    const float* W = U.w.get();
    // sum and count of the non-negative costs, in one pass
    reduce_pair_to(
        ctx,
        [=] __device__(i64 e, double& a, double& b) {
            const float x = W[e];
            a = x >= 0.0f ? (double)x : 0.0;
            b = x >= 0.0f ? 1.0 : 0.0;
        },
        U.nnz, &st->head.pos_sum, &st->head.pos_cnt);
    const i64 threads = (i64)n * kGroup;
    const unsigned pick_grid = std::min<unsigned>(grid_for(threads), (unsigned)tune::get(tune::PickBlocks, threads));
    const unsigned match_grid = std::min<unsigned>(grid_for(n), (unsigned)tune::get(tune::MatchBlocks, n));
    // the stop rule runs on the device; the host only skips launches once it
    // knows the loop has ended: the first kFirstRounds rounds are queued,
    // the state is read, and the remaining rounds are queued only if the
    // loop is still running
    constexpr int kFirstRounds = 3;
    auto stopped = [&](const MatchHead& h, int rounds) {
        if (!(h.pos_cnt > 0.0) || mean_multiplier * (float)h.pos_sum / (float)(i64)h.pos_cnt < 0.0f) return true;
        int prev = 0;
        for (int t = 0; t < rounds; ++t) {
            const int cur = prev + h.newly[t];
            const float rel = (cur - prev) / (prev + 1.0f);
            prev = cur;
            if (h.newly[t] == 0 || rel < 0.1f) return true;
        }
        return false;
    };
    auto read_head = [&](MatchHead& h) {
        MatchHead* m = ctx.mirror<MatchHead>();
        k_match_totals<<<1, 64, 0, ctx.stream>>>(st, m);
        RAMA_LAUNCHED();
        ctx.read_mirror(&st->head, m, &h);
    };
    MatchHead h;
    bool have = false;  // h holds the final state
    for (int t = 0; t < kMaxRounds; ++t) {
        if (t == kFirstRounds) {
            read_head(h);
            if (stopped(h, kFirstRounds)) {
                have = true;
                break;
            }
        }
        {
            // offsets + (neighbour, cost, neighbour's matched flag) per CSR entry + best per vertex
            prof::Scope ps(ctx, "match_pick_best", 12.0 * U.nnz + 12.0 * n);
            k_pick_best<<<pick_grid, kBlock, 0, ctx.stream>>>(n, mean_multiplier, t, st, U.off.get(), U.adj.get(),
                                                              U.w.get(), matched.get(), best.get());
            RAMA_LAUNCHED();
        }
        k_match<<<match_grid, kBlock, 0, ctx.stream>>>(n, mean_multiplier, t, st, best.get(), matched.get(),
                                                       map.get());
        RAMA_LAUNCHED();
    }
    if (!have) read_head(h);
    // replay the stop rule on the host for the returned count
    int prev = 0;
    if (h.pos_cnt > 0.0 && mean_multiplier * (float)h.pos_sum / (float)(i64)h.pos_cnt >= 0.0f) {
        for (int t = 0; t < kMaxRounds; ++t) {
            const int cur = prev + h.newly[t];
            const float rel = (cur - prev) / (prev + 1.0f);
            prev = cur;
            if (ctx.verbose) std::printf("matched sum: %d, rel_increase: %g\n", cur, rel);
            if (h.newly[t] == 0 || rel < 0.1f) break;
        }
    }
    return prev;
}

// ------------------------------------------------------------------
// Packed instances (disjoint node ranges, nslot[v] = instance of node v):
// each instance has its own MatchState, threshold and stop rule; a vertex
// takes part in a round only while its instance's loop runs.  The matched
// pairs of an instance are the pairs it would match alone.
__global__ __launch_bounds__(kBlock) void k_pick_best_packed(int n, float mult, int round,
                                                             const int* __restrict__ nslot,
                                                             const MatchState* __restrict__ sts,
                                                             const int* __restrict__ off,
                                                             const int* __restrict__ adj,
                                                             const float* __restrict__ w,
                                                             const int* __restrict__ matched,
                                                             int* __restrict__ best) {
    const i64 t = (i64)xcd_block() * kBlock + threadIdx.x;
    const int v = (int)(t / kGroup);
    const int sub = (int)(t % kGroup);
    float thr = 0.0f;
    // all lanes of a group share v, so they agree on `active`
    const bool active = v < n && match_round_active(sts + nslot[v], mult, round, &thr) && !matched[v];
    const int b = active ? off[v] : 0, e = active ? off[v + 1] : 0;
    float mx = thr;
    int nid = -1;
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

__global__ __launch_bounds__(kBlock) void k_match_packed(int n, float mult, int round, const int* __restrict__ nslot,
                                                         MatchState* sts, const int* __restrict__ best, int* matched,
                                                         int* map) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    const int q = v < n ? nslot[v] : 0;
    float thr;
    int add = 0;
    if (v < n && match_round_active(sts + q, mult, round, &thr) && !matched[v]) {
        const int vb = best[v];
        // as in k_match: best[vb] == v already implies vb is unmatched
        // (vb is in v's instance, whose rounds run or stop together)
        if (vb > v && best[vb] == v) {
            map[vb] = v;
            matched[v] = 1;
            matched[vb] = 1;
            add = 2;
        }
    }
    // per-instance count: one atomic per distinct instance in the wave
    const int lane = lane_id();
    unsigned long long pending = __ballot(add != 0);
    while (pending) {
        const int leader = __builtin_ctzll(pending);
        const int ql = __builtin_amdgcn_readlane(q, leader);
        const unsigned long long m = __ballot(q == ql) & pending;
        if (lane == leader) {
            atomicAdd(&sts[ql].newly[round], 2 * __popcll(m));
            sts[ql].running[round] = 1;
        }
        pending &= ~m;
    }
}

// Matching of packed instances on their joint symmetric CSR U (n rows).
// Instance q's CSR entries are [hbeg[q], hbeg[q+1]) (dbeg on the device).
// Fills map (global ids) and matched_out[q] = matched vertices of q.
void matching_packed(Ctx& ctx, const Csr& U, const int* nslot, int Q, const i64* hbeg, const i64* dbeg,
                     float mean_multiplier, DBuf<int>& map, int* matched_out) {
    const int n = U.n;
    map = ctx.buf<int>(n > 0 ? n : 1);
    DBuf<MatchState> stb = ctx.buf<MatchState>(Q);
    MatchState* st = stb.get();
    DBuf<int> best = ctx.buf<int>(n > 0 ? n : 1), matched = ctx.buf<int>(n > 0 ? n : 1);
    {
        int* M = map.get();
        int* B = best.get();
        int* D = matched.get();
        for_each(ctx, std::max<i64>(n, Q), [=] __device__(i64 v) {
            if (v < n) {
                M[v] = (int)v;
                B[v] = -1;
                D[v] = 0;
            }
            if (v < Q) {
                st[v].pos_sum = 0.0;
                st[v].pos_cnt = 0.0;
                for (int r = 0; r < kMaxRounds; ++r) {
                    st[v].newly[r] = 0;
                    st[v].running[r] = 0;
                }
            }
        });
    }
    const float* W = U.w.get();
    constexpr int kStride = (int)(sizeof(MatchState) / sizeof(double));
    static_assert(sizeof(MatchState) % sizeof(double) == 0, "MatchState layout");
    seg_reduce<double>(ctx, [=] __device__(i64 e) { const float x = W[e]; return x >= 0.0f ? (double)x : 0.0; },
                       hbeg, dbeg, Q, 0.0, OpSum(), &st->pos_sum, kStride);
    seg_reduce<double>(ctx, [=] __device__(i64 e) { return W[e] >= 0.0f ? 1.0 : 0.0; }, hbeg, dbeg, Q, 0.0,
                       OpSum(), &st->pos_cnt, kStride);
    std::vector<MatchState> h((size_t)Q);
    auto replay = [&](const MatchState& x, int rounds, int* count) {
        // the stop rule of filter_edges_by_matching_vertex_based (:121-127)
        int prev = 0;
        bool stopped = true;
        if (x.pos_cnt > 0.0 && mean_multiplier * (float)x.pos_sum / (float)(i64)x.pos_cnt >= 0.0f) {
            stopped = false;
            for (int t = 0; t < rounds; ++t) {
                const int cur = prev + x.newly[t];
                const float rel = (cur - prev) / (prev + 1.0f);
                prev = cur;
                if (!x.running[t] || rel < 0.1f) {
                    stopped = true;
                    break;
                }
            }
        }
        if (count) *count = prev;
        return stopped;
    };
    bool have = false;
    if (n > 0) {
        const i64 threads = (i64)n * kGroup;
        constexpr int kFirstRounds = 3;
        for (int t = 0; t < kMaxRounds; ++t) {
            if (t == kFirstRounds) {
                ctx.d2h(h.data(), st, Q);
                bool all = true;
                for (int q = 0; q < Q && all; ++q) all = replay(h[q], kFirstRounds, nullptr);
                if (all) {
                    have = true;
                    break;
                }
            }
            k_pick_best_packed<<<grid_for(threads), kBlock, 0, ctx.stream>>>(
                n, mean_multiplier, t, nslot, st, U.off.get(), U.adj.get(), U.w.get(), matched.get(), best.get());
            RAMA_LAUNCHED();
            k_match_packed<<<grid_for(n), kBlock, 0, ctx.stream>>>(n, mean_multiplier, t, nslot, st, best.get(),
                                                                   matched.get(), map.get());
            RAMA_LAUNCHED();
        }
    }
    if (!have) ctx.d2h(h.data(), st, Q);
    for (int q = 0; q < Q; ++q) replay(h[q], kMaxRounds, &matched_out[q]);
}

}  // namespace rama
