// lockstep.hip — a batch of instances solved together in one packed graph.
//
// Instance q owns the node ids [base[q], base[q] + span[q]) of one joint
// graph.  Edges never cross instances, so the sorts, contractions, CSR
// builds, cycle searches, message passing and matching of a round run once
// for all of them: a round costs about as many launches as one instance's
// round, whatever the batch size.  Everything the single-instance solver
// decides from scalars is kept per instance — cycle thresholds and triangle
// budgets, message-passing bounds and stall tests, matching thresholds and
// stop rules, the matching→forest crossover, termination, the dCOO rows/cols
// bookkeeping — and every floating-point reduction of an instance runs in
// the order of its own solve, so each instance's labels and lower bound are
// bit-identical to solving it alone.  Instances leave the pack when they
// finish; one that switches to the spanning-forest contraction continues on
// a worker thread with the single-instance rounds from that round on.
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <deque>
#include <map>
#include <mutex>
#include <thread>

#include "rama_internal.hpp"

namespace rama {

namespace {

struct Slot {
    int id;          // instance index in the batch
    int span;        // node ids [base, base + span)
    int rows, cols, max_node;
    i64 nnz;
    int n0;          // number of labels
    double lb;       // lower bound of the initial dual
    int iter;        // the instance's next round (0 = first)
};

struct Pack {
    std::vector<Slot> s;
    std::vector<int> base;  // Q+1, host
    DBuf<int> dbase;        // Q+1, device
    DBuf<int> nslot;        // node -> slot
    int span_total = 0;
    DCoo A;                 // joint graph, global node ids, sorted by (row, col)
    DBuf<int> labels;       // labels of every slot, concatenated in slot order
    std::vector<i64> lab;   // Q+1 prefix of n0
    int Q() const { return (int)s.size(); }
};

__global__ void k_node_slot(const int* __restrict__ base, int Q, int n, int* __restrict__ nslot) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    int lo = 0, hi = Q - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (base[mid] <= v) lo = mid; else hi = mid - 1;
    }
    nslot[v] = lo;
}

// Largest q in [0, Q) with pre[q] <= x (pre non-decreasing, Q+1 entries).
__device__ __forceinline__ int seg_of(const i64* pre, int Q, i64 x) {
    int lo = 0, hi = Q - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pre[mid] <= x) lo = mid; else hi = mid - 1;
    }
    return lo;
}

template <class T>
DBuf<T> upload_vec(Ctx& ctx, const std::vector<T>& h, const char* fn = __builtin_FUNCTION(),
                   int line = __builtin_LINE()) {
    DBuf<T> d = ctx.buf<T>(std::max<i64>((i64)h.size(), 1));
    ctx.upload(d.get(), h.data(), (i64)h.size(), fn, line);
    return d;
}

std::vector<i64> prefix(const std::vector<i64>& n) {
    std::vector<i64> p(n.size() + 1, 0);
    for (size_t i = 0; i < n.size(); ++i) p[i + 1] = p[i] + n[i];
    return p;
}

// Bases, the device copy and the node -> slot map from the slots' spans.
void pack_index(Ctx& ctx, Pack& P) {
    const int Q = P.Q();
    P.base.assign((size_t)Q + 1, 0);
    for (int q = 0; q < Q; ++q) P.base[q + 1] = P.base[q] + P.s[q].span;
    P.span_total = P.base[Q];
    P.dbase = upload_vec(ctx, P.base);
    P.nslot = ctx.buf<int>(std::max(P.span_total, 1));
    if (P.span_total > 0) {
        k_node_slot<<<grid_for(P.span_total), kBlock, 0, ctx.stream>>>(P.dbase.get(), Q, P.span_total,
                                                                        P.nslot.get());
        RAMA_LAUNCHED();
    }
    P.lab.assign((size_t)Q + 1, 0);
    for (int q = 0; q < Q; ++q) P.lab[q + 1] = P.lab[q] + P.s[q].n0;
    // the joint graph's own dims only bound its ids (sort key widths)
    P.A.rows = P.A.cols = P.span_total;
    P.A.max_node = P.span_total - 1;
}

std::vector<i64> edge_prefix(const Pack& P) {
    std::vector<i64> n((size_t)P.Q());
    for (int q = 0; q < P.Q(); ++q) n[q] = P.s[q].nnz;
    return prefix(n);
}

// Read per-slot stats and apply them as the dCOO dims of the slots in `which`.
void apply_stats(Ctx& ctx, Pack& P, const SlotStats* d_st, const std::vector<char>& which) {
    const int Q = P.Q();
    std::vector<SlotStats> h((size_t)Q);
    ctx.d2h(h.data(), d_st, Q);
    for (int q = 0; q < Q; ++q) {
        if (!which[q]) continue;
        P.s[q].cols = h[q].max_col + 1;
        P.s[q].rows = h[q].max_row + 1;
        P.s[q].max_node = h[q].max_kept_col;
        P.s[q].nnz = h[q].kept;
    }
}

// Per-slot lower bound of the current graph (get_lb), reduced as dcoo_lb
// reduces the instance's own edge array.
std::vector<double> packed_dcoo_lb(Ctx& ctx, Pack& P) {
    const int Q = P.Q();
    std::vector<double> out((size_t)Q, 0.0);
    const std::vector<i64> eb = edge_prefix(P);
    DBuf<i64> deb = upload_vec(ctx, eb);
    DBuf<double> d = ctx.buf<double>(Q);
    const float* W = P.A.val.get();
    seg_reduce<double>(
        ctx, [=] __device__(i64 e) { const float x = W[e]; return x < 0.0f ? (double)x : 0.0; }, eb.data(),
        deb.get(), Q, 0.0, OpSum(), d.get(), 1);
    ctx.d2h(out.data(), d.get(), Q);
    return out;
}

// ------------------------------------------------------- cycle search
__device__ __forceinline__ int float_order(float x) {
    const int i = __float_as_int(x);
    return i >= 0 ? i : i ^ 0x7fffffff;
}
__device__ __forceinline__ float order_float(int i) { return __int_as_float(i >= 0 ? i : i ^ 0x7fffffff); }

struct CycScalars {
    int min_ord;  // smallest cost of the slot (order-preserving int)
    int npos;     // attractive edges
    int nrep;     // repulsive edges
    int pad;
};

// Per-slot update of the cycle-search scalars: min of `mn`, sums of `pos`
// and `rep`, one atomic each per distinct slot in the wave.  Wave-uniform.
__device__ __forceinline__ void cyc_scalars_add(CycScalars* cs, int q, bool on, int mn, int pos, int rep) {
    const int lane = lane_id();
    unsigned long long pending = __ballot(on);
    while (pending) {
        const int leader = __builtin_ctzll(pending);
        const int ql = __builtin_amdgcn_readlane(q, leader);
        const bool mine = ((pending >> lane) & 1ull) && q == ql;
        const unsigned long long m = __ballot(mine);
        int a = mine ? mn : 0x7fffffff, b = mine ? pos : 0, c = mine ? rep : 0;
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) {
            const int oa = __shfl_down(a, d, 64);
            a = oa < a ? oa : a;
            b += __shfl_down(b, d, 64);
            c += __shfl_down(c, d, 64);
        }
        if (lane == 0) {
            if (a != 0x7fffffff) atomicMin(&cs[ql].min_ord, a);
            if (b) atomicAdd(&cs[ql].npos, b);
            if (c) atomicAdd(&cs[ql].nrep, c);
        }
        pending &= ~m;
    }
}

struct MinAcc {
    int v = 0x7fffffff;
};
struct MinRed {
    __device__ MinAcc operator()(MinAcc a, MinAcc b) const { return MinAcc{a.v < b.v ? a.v : b.v}; }
    __device__ MinAcc shfl_down(MinAcc a, int d) const { return MinAcc{__shfl_down(a.v, d, 64)}; }
};
struct CountAcc {
    int pos = 0;
    int rep = 0;
};
struct CountRed {
    __device__ CountAcc operator()(CountAcc a, CountAcc b) const { return CountAcc{a.pos + b.pos, a.rep + b.rep}; }
    __device__ CountAcc shfl_down(CountAcc a, int d) const {
        return CountAcc{__shfl_down(a.pos, d, 64), __shfl_down(a.rep, d, 64)};
    }
};
struct CheckAcc {
    int mx = -1;  // largest id
    int neg = 0;  // a negative id
};
struct CheckRed {
    __device__ CheckAcc operator()(CheckAcc a, CheckAcc b) const {
        return CheckAcc{a.mx > b.mx ? a.mx : b.mx, a.neg | b.neg};
    }
    __device__ CheckAcc shfl_down(CheckAcc a, int d) const {
        return CheckAcc{__shfl_down(a.mx, d, 64), __shfl_down(a.neg, d, 64)};
    }
};

struct PackedTriangles {
    Triangles U;              // joint triangle list, sorted, slot after slot
    std::vector<i64> raw;     // per slot: raw count (the search's triangles_v1.size())
    std::vector<i64> n;       // per slot: triangles after normalisation
};

// out[q] = first position in the sorted node-id array `key` (length n) at
// or above base[q], for q = 0..Q.
__global__ void k_node_bounds(const int* __restrict__ key, i64 n, const int* __restrict__ base, int Q,
                              i64* __restrict__ out) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q > Q) return;
    const int x = base[q];
    i64 lo = 0, hi = n;
    while (lo < hi) {
        const i64 mid = (lo + hi) >> 1;
        if (key[mid] < x) lo = mid + 1; else hi = mid;
    }
    out[q] = lo;
}

// rearrange_triangles on the regions of the slots in `take` (valid[q] slots
// each, from out[q]).  With `compact`, the sorted unique union is returned in
// `U`; otherwise each slot's triangles are written back to the start of its
// region.  Returns the per-slot counts.
std::vector<i64> rearrange_packed(Ctx& ctx, Pack& P, int* t1, int* t2, int* t3, const std::vector<i64>& out,
                                  const std::vector<i64>& valid, const std::vector<char>& take, bool compact,
                                  Triangles* U) {
    const int Q = P.Q();
    std::vector<i64> cnt((size_t)Q, 0);
    std::vector<i64> v((size_t)Q, 0);
    i64 total = 0;
    for (int q = 0; q < Q; ++q) {
        v[q] = take[q] ? valid[q] : 0;
        total += v[q];
    }
    if (total == 0) {
        if (compact && U) *U = Triangles();
        return cnt;
    }
    DBuf<i64> dout = upload_vec(ctx, out);
    DBuf<i64> dv = upload_vec(ctx, v);
    const i64* DO = dout.get();
    const i64* DV = dv.get();
    Triangles T;
    T.t1 = ctx.buf<int>(total);
    T.t2 = ctx.buf<int>(total);
    T.t3 = ctx.buf<int>(total);
    int* A1 = T.t1.get();
    int* A2 = T.t2.get();
    int* A3 = T.t3.get();
    DBuf<int> c = ctx.buf<int>(1);
    select_into(
        ctx, out[Q],
        [=] __device__(i64 p) {
            const int q = seg_of(DO, Q, p);
            return p - DO[q] < DV[q];
        },
        [=] __device__(i64 p, int k) {
            A1[k] = t1[p];
            A2[k] = t2[p];
            A3[k] = t3[p];
        },
        c.get());
    const i64 u = rearrange_triangles(ctx, T, total, std::max(P.span_total, 2));
    // the unique list is sorted by first vertex, so slots are contiguous in it
    DBuf<i64> pb = ctx.buf<i64>(Q + 1);
    k_node_bounds<<<grid_for(Q + 1), kBlock, 0, ctx.stream>>>(A1, u, P.dbase.get(), Q, pb.get());
    RAMA_LAUNCHED();
    std::vector<i64> h((size_t)Q + 1);
    ctx.d2h(h.data(), pb.get(), Q + 1);
    for (int q = 0; q < Q; ++q) cnt[q] = h[q + 1] - h[q];
    const int* NS = P.nslot.get();
    if (compact) {
        T.n = u;
        T.max_id = P.span_total - 1;
        *U = std::move(T);
        return cnt;
    }
    const std::vector<i64> first = prefix(cnt);
    DBuf<i64> df = upload_vec(ctx, first);
    const i64* DF = df.get();
    for_each(ctx, u, [=] __device__(i64 k) {
        const int q = NS[A1[k]];
        const i64 dst = DO[q] + (k - DF[q]);
        t1[dst] = A1[k];
        t2[dst] = A2[k];
        t3[dst] = A3[k];
    });
    return cnt;
}

// conflicted_cycles for every slot with part[q] (others get none).
PackedTriangles packed_cycles(Ctx& ctx, Pack& P, int max_len, float factor, float tol,
                              const std::vector<char>& part) {
    const int Q = P.Q();
    PackedTriangles R;
    R.raw.assign((size_t)Q, 0);
    R.n.assign((size_t)Q, 0);
    if (max_len > 5) throw Error(RAMA_ERR_INVALID, "max_cycle_length should be <= 5. Received: " + std::to_string(max_len));
    if (max_len < 3 || P.A.nnz == 0) return R;
    const i64 m = P.A.nnz;
    const int* AR = P.A.row.get();
    const int* AC = P.A.col.get();
    const float* W = P.A.val.get();
    const int* NS = P.nslot.get();
    // determine_threshold per slot: min cost, then both class sizes, by
    // range reductions over each slot's edges
    DBuf<CycScalars> dcs = ctx.buf<CycScalars>(Q);
    CycScalars* CS = dcs.get();
    for_each(ctx, Q, [=] __device__(i64 q) { CS[q] = CycScalars{0x7fffffff, 0, 0, 0}; });
    const std::vector<i64> eb = edge_prefix(P);
    DBuf<i64> deb = upload_vec(ctx, eb);
    range_reduce<MinAcc>(
        ctx, deb.get(), Q,
        [=] __device__(i64 e, MinAcc& a) {
            const int o = float_order(W[e]);
            a.v = o < a.v ? o : a.v;
        },
        MinRed(), [=] __device__(int q, MinAcc r) { atomicMin(&CS[q].min_ord, r.v); });
    range_reduce<CountAcc>(
        ctx, deb.get(), Q,
        [=] __device__(i64 e, CountAcc& a) {
            const float t = order_float(CS[NS[AR[e]]].min_ord) * tol;
            a.pos += W[e] > -1.0f * t;
            a.rep += W[e] < t;
        },
        CountRed(),
        [=] __device__(int q, CountAcc r) {
            atomicAdd(&CS[q].npos, r.pos);
            atomicAdd(&CS[q].nrep, r.rep);
        });
    std::vector<CycScalars> hs((size_t)Q);
    ctx.d2h(hs.data(), CS, Q);
    std::vector<float> keep_above((size_t)Q), rep_below((size_t)Q);
    std::vector<i64> budget((size_t)Q, 0);
    i64 npos = 0, nrep = 0;
    const float inf = 3.402823466e38f;
    for (int q = 0; q < Q; ++q) {
        keep_above[q] = inf;
        rep_below[q] = -inf;
        if (!part[q] || P.s[q].nnz == 0) continue;
        // bit pattern back to float exactly as the device decodes it
        int i = hs[q].min_ord;
        i = i >= 0 ? i : i ^ 0x7fffffff;
        float mn;
        std::memcpy(&mn, &i, 4);
        const float thr = mn * tol;
        if (hs[q].npos == 0 || hs[q].nrep == 0) continue;  // no triangles for this slot
        keep_above[q] = -1.0f * thr;
        rep_below[q] = thr;
        budget[q] = (i64)(int)(factor * (float)hs[q].nrep);
        npos += hs[q].npos;
        nrep += hs[q].nrep;
    }
    if (nrep == 0) return R;
    DBuf<float> dka = upload_vec(ctx, keep_above), drb = upload_vec(ctx, rep_below);
    const float* RB = drb.get();
    Csr G = symmetric_csr_slots(ctx, AR, AC, W, m, P.span_total, NS, dka.get(), npos);
    DBuf<int> rv1 = ctx.buf<int>(nrep), rv2 = ctx.buf<int>(nrep);
    {
        int* X1 = rv1.get();
        int* X2 = rv2.get();
        DBuf<int> cnt = ctx.buf<int>(1);
        select_into(
            ctx, m, [=] __device__(i64 e) { return W[e] < RB[NS[AR[e]]]; },
            [=] __device__(i64 e, int k) {
                X1[k] = AR[e];
                X2[k] = AC[e];
            },
            cnt.get());
    }
    // one region of budget[q] triangle slots per slot
    const std::vector<i64> out = prefix(budget);
    const i64 cap = std::max<i64>(out[Q], 1);
    int* t1;
    int* t2;
    int* t3;
    DBuf<int> b1 = ctx.buf<int>(cap), b2 = ctx.buf<int>(cap), b3 = ctx.buf<int>(cap);
    t1 = b1.get();
    t2 = b2.get();
    t3 = b3.get();
    DBuf<i64> dout = upload_vec(ctx, out), dbud = upload_vec(ctx, budget);
    std::vector<i64> count((size_t)Q, 0), total((size_t)Q, 0), base((size_t)Q, 0);
    {
        DBuf<i64> dbase = upload_vec(ctx, base);
        run_stage_packed<3>(ctx, G, rv1.get(), rv2.get(), nrep, NS, Q, dbase.get(), dbud.get(), dout.get(), t1, t2,
                            t3, total.data());
        for (int q = 0; q < Q; ++q) count[q] = budget[q] > 0 ? total[q] : 0;
    }
    for (int S = 4; S <= max_len; ++S) {
        std::vector<char> elig((size_t)Q, 0);
        bool any = false;
        for (int q = 0; q < Q; ++q) {
            elig[q] = budget[q] > 0 && count[q] < budget[q];
            any = any || elig[q];
        }
        if (!any) break;
        const std::vector<i64> uq = rearrange_packed(ctx, P, t1, t2, t3, out, count, elig, false, nullptr);
        for (int q = 0; q < Q; ++q) {
            if (elig[q]) count[q] = uq[q];
            base[q] = elig[q] ? count[q] : budget[q];  // a full slot emits nothing
        }
        DBuf<i64> dbase = upload_vec(ctx, base);
        if (S == 4)
            run_stage_packed<4>(ctx, G, rv1.get(), rv2.get(), nrep, NS, Q, dbase.get(), dbud.get(), dout.get(), t1,
                                t2, t3, total.data());
        else
            run_stage_packed<5>(ctx, G, rv1.get(), rv2.get(), nrep, NS, Q, dbase.get(), dbud.get(), dout.get(), t1,
                                t2, t3, total.data());
        for (int q = 0; q < Q; ++q)
            if (elig[q]) count[q] += total[q];
    }
    std::vector<char> all((size_t)Q, 0);
    for (int q = 0; q < Q; ++q) {
        R.raw[q] = std::min(count[q], budget[q]);
        all[q] = R.raw[q] > 0;
    }
    R.n = rearrange_packed(ctx, P, t1, t2, t3, out, R.raw, all, true, &R.U);
    return R;
}

// -------------------------------------------------------------- dual
// dual_update_cycle_length for the slots with on[q]: fin[q] and num_tri[q]
// receive the bound and raw triangle count.
void packed_dual_update(Ctx& ctx, Pack& P, int c, int steps, float factor, float tol, const std::vector<char>& on,
                        std::vector<double>& fin, std::vector<i64>& num_tri) {
    const int Q = P.Q();
    PackedTriangles T = packed_cycles(ctx, P, c, factor, tol, on);
    bool need_lb = false, any_mp = false;
    std::vector<char> mp_on((size_t)Q, 0);
    for (int q = 0; q < Q; ++q) {
        if (!on[q]) continue;
        num_tri[q] = T.raw[q];
        if (T.raw[q] == 0) need_lb = true;
        else mp_on[q] = 1, any_mp = true;
    }
    if (need_lb) {
        const std::vector<double> lb = packed_dcoo_lb(ctx, P);
        for (int q = 0; q < Q; ++q)
            if (on[q] && T.raw[q] == 0) fin[q] = lb[q];
    }
    if (!any_mp) return;
    MPState S;
    mp_build(ctx, P.A, T.U, S);
    T.U = Triangles();
    // triangles and message-passing edges of each slot
    std::vector<i64> tb = prefix(T.n);
    DBuf<i64> dtb = upload_vec(ctx, tb);
    DBuf<i64> deb = ctx.buf<i64>(Q + 1);
    k_node_bounds<<<grid_for(Q + 1), kBlock, 0, ctx.stream>>>(S.ei.get(), S.E, P.dbase.get(), Q, deb.get());
    RAMA_LAUNCHED();
    std::vector<i64> eb((size_t)Q + 1);
    ctx.d2h(eb.data(), deb.get(), Q + 1);
    eb[Q] = S.E;
    ctx.upload(deb.get() + Q, &eb[Q], 1);
    // each slot's edges in a triangle (the edge launch covers only these)
    DBuf<i64> dab = ctx.buf<i64>(Q + 1);
    mp_active_ranges(ctx, S, deb.get(), Q, dab.get());
    std::vector<i64> ab((size_t)Q + 1);
    ctx.d2h(ab.data(), dab.get(), Q + 1);
    std::vector<int> tblk((size_t)Q + 1, 0), eblk((size_t)Q + 1, 0), onv((size_t)Q, 0);
    std::vector<i64> pbeg((size_t)Q, 0);
    for (int q = 0; q < Q; ++q) {
        const int gt = (int)grid_for(tb[q + 1] - tb[q]), ge = (int)mp_edge_blocks(ab[q + 1] - ab[q]);
        tblk[q + 1] = tblk[q] + gt;
        eblk[q + 1] = eblk[q] + ge;
        pbeg[q] = (q ? pbeg[q - 1] + (tblk[q] - tblk[q - 1]) + (eblk[q] - eblk[q - 1]) : 0);
        onv[q] = mp_on[q];
    }
    DBuf<int> dtblk = upload_vec(ctx, tblk), deblk = upload_vec(ctx, eblk), don = upload_vec(ctx, onv);
    DBuf<i64> dpbeg = upload_vec(ctx, pbeg);
    MPPack pk{Q, dtb.get(), deb.get(), dab.get(), dtblk.get(), deblk.get(), dpbeg.get()};
    std::vector<double> lb((size_t)Q, 0.0);
    mp_run_packed(ctx, S, pk, tblk[Q], eblk[Q], steps, tb.data(), eb.data(), don.get(), lb.data());
    for (int q = 0; q < Q; ++q)
        if (mp_on[q]) fin[q] = lb[q];
    // the reparametrised graph; dims of the slots that changed
    P.A.row = std::move(S.ei);
    P.A.col = std::move(S.ej);
    P.A.val = std::move(S.ec);
    P.A.nnz = S.E;
    for (int q = 0; q < Q; ++q) P.s[q].nnz = eb[q + 1] - eb[q];
    DBuf<SlotStats> st = ctx.buf<SlotStats>(Q);
    slot_stats_clear(ctx, st.get(), Q);
    slot_stats_edges(ctx, P.A.row.get(), P.A.col.get(), P.A.nnz, P.dbase.get(), Q, st.get());
    apply_stats(ctx, P, st.get(), mp_on);
}

// dual_solver for every slot with part[q]; returns the bounds.
std::vector<double> packed_dual(Ctx& ctx, Pack& P, int max_len, int num_iter, float factor, int num_outer, float tol,
                                const std::vector<char>& part) {
    const int Q = P.Q();
    std::vector<double> fin((size_t)Q, 0.0), prev((size_t)Q, 0.0);
    if (max_len < 3 || num_iter == 0 || num_outer == 0) {
        const std::vector<double> lb = packed_dcoo_lb(ctx, P);
        for (int q = 0; q < Q; ++q)
            if (part[q]) fin[q] = lb[q];
        return fin;
    }
    std::vector<char> act = part;
    std::vector<i64> num_tri((size_t)Q, 0);
    for (int o = 0; o < num_outer; ++o) {
        bool any = false;
        for (int q = 0; q < Q; ++q) any = any || act[q];
        if (!any) break;
        for (int c = 3; c <= max_len; ++c) packed_dual_update(ctx, P, c, num_iter, factor, tol, act, fin, num_tri);
        for (int q = 0; q < Q; ++q) {
            if (!act[q]) continue;
            if (num_tri[q] == 0 || (o > 0 && (fin[q] - prev[q]) < 1e-3)) act[q] = 0;
            prev[q] = fin[q];
        }
    }
    return fin;
}

// Concatenates the label ranges [src_off[c], src_off[c] + n[c]) of `src`
// into a new buffer, in list order, with one launch.
DBuf<int> gather_ranges(Ctx& ctx, const int* src, const std::vector<i64>& src_off, const std::vector<i64>& n) {
    const std::vector<i64> dst = prefix(n);
    const int C = (int)n.size();
    DBuf<int> g = ctx.buf<int>(std::max<i64>(dst[C], 1));
    if (dst[C] == 0) return g;
    DBuf<i64> dd = upload_vec(ctx, dst), ds = upload_vec(ctx, src_off);
    const i64* DD = dd.get();
    const i64* DS = ds.get();
    int* G = g.get();
    for_each(ctx, dst[C], [=] __device__(i64 x) {
        const int c = seg_of(DD, C, x);
        G[x] = src[DS[c] + (x - DD[c])];
    });
    return g;
}

// Labels of the listed slots, copied out in one read (slot after slot).
std::vector<int32_t> gather_labels(Ctx& ctx, const Pack& P, const std::vector<int>& slots) {
    std::vector<i64> off, n;
    i64 tot = 0;
    for (int q : slots) {
        off.push_back(P.lab[q]);
        n.push_back(P.s[q].n0);
        tot += P.s[q].n0;
    }
    std::vector<int32_t> h((size_t)tot);
    if (tot == 0) return h;
    DBuf<int> g = gather_ranges(ctx, P.labels.get(), off, n);
    ctx.d2h(h.data(), g.get(), tot);
    return h;
}

// ------------------------------------------------------------ handoff
struct Handoff {
    int id = 0;
    int next_iter = 0;
    std::vector<int32_t> row, col;  // local ids
    std::vector<float> val;
    int rows = 0, cols = 0, max_node = -1;
    std::vector<int32_t> labels;
    double lb = 0.0;
};

}  // namespace

// ------------------------------------------------------------- driver

// The lockstep solver: instances are admitted in groups (each group runs
// its initial dual on arrival), and every call of round() advances all
// admitted instances by one contraction round.  Instances leave when they
// finish; those that switch to the forest contraction are continued by
// `workers` threads with the single-instance rounds.
class Lockstep {
   public:
    Lockstep(Ctx& ctx, const Options& o, int workers, LockstepSink& sink) : ctx_(ctx), o_(o), sink_(sink) {
        times_ = std::getenv("RAMA_STAGE_TIMES") != nullptr;
        t0_ = tlast_ = std::chrono::steady_clock::now();
        for (int w = 0; w < std::max(1, workers); ++w) pool_.emplace_back([this] { worker(); });
    }
    ~Lockstep() { finish(); }

    static constexpr int kMaxSlots = kLockstepMaxSlots;
    int slots() const { return P_.Q(); }
    bool empty() const { return P_.Q() == 0; }

    // Admits k instances (edges [edge_off[q], edge_off[q+1]) of i, j, c, own
    // 0-based ids) under the external ids ids[q].
    void admit(int k, const int64_t* edge_off, const int32_t* i, const int32_t* j, const float* c, const int* ids);
    void round();
    // Waits for the handed-off instances.
    void finish() {
        if (pool_.empty()) return;
        {
            std::lock_guard<std::mutex> l(qmu_);
            closing_ = true;
        }
        qcv_.notify_all();
        for (auto& t : pool_) t.join();
        pool_.clear();
        mark("workers done", -1, 0);
        if (times_)
            for (auto& kv : stage_ms_)
                std::fprintf(stderr, "packed total %-14s %9.3f ms\n", kv.first.c_str(), kv.second);
    }

   private:
    void worker();
    void hand_off(Handoff&& h) {
        {
            std::lock_guard<std::mutex> l(qmu_);
            tasks_.push_back(std::move(h));
        }
        qcv_.notify_one();
    }
    void mark(const char* what, int rounds, int slots) {
        if (!times_) return;
        ctx_.sync();
        const auto now = std::chrono::steady_clock::now();
        stage_ms_[what] += std::chrono::duration<double, std::milli>(now - tlast_).count();
        tlast_ = now;
        if (rounds >= 0 && std::string(what) != "round") return;
        const double ms = std::chrono::duration<double, std::milli>(now - t0_).count();
        std::fprintf(stderr, "packed %-14s %9.3f ms  round %d  slots %d\n", what, ms, rounds, slots);
    }
    // Publishes the listed slots as finished (labels from the pack).
    void publish(const std::vector<int>& slots) {
        scan_check(ctx_);  // no result leaves a pack whose sorts may be mis-ordered
        const std::vector<int32_t> hl = gather_labels(ctx_, P_, slots);
        i64 o = 0;
        for (int q : slots) {
            sink_.done(P_.s[q].id, 0, "", hl.data() + o, P_.s[q].n0, P_.s[q].lb);
            o += P_.s[q].n0;
        }
    }
    // Keeps only the slots with keep[q] (edges, labels, ids shifted down).
    void compact(const std::vector<char>& keep);
    // Hands the slots in `which` to the workers (next round of each: its iter).
    void hand_off_slots(const std::vector<int>& which);

    Ctx& ctx_;
    Options o_;
    LockstepSink& sink_;
    Pack P_;
    bool times_ = false;
    std::chrono::steady_clock::time_point t0_, tlast_;
    std::map<std::string, double> stage_ms_;
    std::mutex qmu_;
    std::condition_variable qcv_;
    std::deque<Handoff> tasks_;
    bool closing_ = false;
    std::vector<std::thread> pool_;
};

void Lockstep::worker() {
    std::unique_ptr<Ctx> wctx;
    for (;;) {
        Handoff h;
        {
            std::unique_lock<std::mutex> l(qmu_);
            qcv_.wait(l, [&] { return closing_ || !tasks_.empty(); });
            if (tasks_.empty()) return;
            h = std::move(tasks_.front());
            tasks_.pop_front();
        }
        try {
            if (!wctx) wctx.reset(new Ctx(ctx_.device, nullptr, false));
            Ctx& w = *wctx;
            DCoo A;
            A.nnz = (i64)h.val.size();
            A.row = w.buf<int>(A.nnz);
            A.col = w.buf<int>(A.nnz);
            A.val = w.buf<float>(A.nnz);
            w.h2d(A.row.get(), h.row.data(), A.nnz);
            w.h2d(A.col.get(), h.col.data(), A.nnz);
            w.h2d(A.val.get(), h.val.data(), A.nnz);
            A.rows = h.rows;
            A.cols = h.cols;
            A.max_node = h.max_node;
            SolveResult out;
            out.num_labels = (i64)h.labels.size();
            out.labels = w.buf<int>(out.num_labels);
            w.h2d(out.labels.get(), h.labels.data(), out.num_labels);
            out.lb = h.lb;
            solve_rounds(w, A, out, o_, h.next_iter, false);
            std::vector<int32_t> lab((size_t)out.num_labels);
            if (out.num_labels > 0) w.d2h(lab.data(), out.labels.get(), out.num_labels);
            sink_.done(h.id, 0, "", lab.data(), out.num_labels, h.lb);
        } catch (const Error& e) {
            sink_.done(h.id, e.status, e.what(), nullptr, 0, h.lb);
        } catch (const std::exception& e) {
            sink_.done(h.id, RAMA_ERR_INTERNAL, e.what(), nullptr, 0, h.lb);
        }
    }
}

void Lockstep::compact(const std::vector<char>& keep) {
    const int Q = P_.Q();
    std::vector<int> stay;
    for (int q = 0; q < Q; ++q)
        if (keep[q]) stay.push_back(q);
    if ((int)stay.size() == Q) return;
    if (stay.empty()) {
        P_ = Pack();
        return;
    }
    std::vector<int> kp((size_t)Q, 0), shift((size_t)Q, 0);
    int removed = 0;
    for (int q = 0; q < Q; ++q) {
        shift[q] = removed;
        kp[q] = keep[q] ? 1 : 0;
        if (!keep[q]) removed += P_.s[q].span;
    }
    DBuf<int> dkeep = upload_vec(ctx_, kp), dshift = upload_vec(ctx_, shift);
    const int* KP = dkeep.get();
    const int* SH = dshift.get();
    const int* NS = P_.nslot.get();
    DCoo A2;
    i64 m2 = 0;
    for (int q : stay) m2 += P_.s[q].nnz;
    A2.nnz = m2;
    A2.row = ctx_.buf<int>(std::max<i64>(m2, 1));
    A2.col = ctx_.buf<int>(std::max<i64>(m2, 1));
    A2.val = ctx_.buf<float>(std::max<i64>(m2, 1));
    {
        const int* AR = P_.A.row.get();
        const int* AC = P_.A.col.get();
        const float* AW = P_.A.val.get();
        int* R2 = A2.row.get();
        int* C2 = A2.col.get();
        float* W2 = A2.val.get();
        DBuf<int> cnt = ctx_.buf<int>(1);
        select_into(
            ctx_, P_.A.nnz, [=] __device__(i64 e) { return KP[NS[AR[e]]] != 0; },
            [=] __device__(i64 e, int p) {
                const int q = NS[AR[e]];
                R2[p] = AR[e] - SH[q];
                C2[p] = AC[e] - SH[q];
                W2[p] = AW[e];
            },
            cnt.get());
    }
    std::vector<i64> off, n;
    for (int q : stay) {
        off.push_back(P_.lab[q]);
        n.push_back(P_.s[q].n0);
    }
    DBuf<int> labels = gather_ranges(ctx_, P_.labels.get(), off, n);
    std::vector<Slot> ns;
    for (int q : stay) ns.push_back(P_.s[q]);
    P_.s = ns;
    P_.A = std::move(A2);
    P_.labels = std::move(labels);
    pack_index(ctx_, P_);
}

void Lockstep::hand_off_slots(const std::vector<int>& which) {
    if (which.empty()) return;
    scan_check(ctx_);
    const int Q = P_.Q();
    std::vector<int> isin((size_t)Q, 0);
    std::vector<i64> hoff;
    i64 mh = 0;
    for (int q : which) {
        isin[q] = 1;
        hoff.push_back(mh);
        mh += P_.s[q].nnz;
    }
    std::vector<int32_t> er((size_t)mh), ec((size_t)mh);
    std::vector<float> ev((size_t)mh);
    if (mh > 0) {
        // the slots' edges, gathered on the device in slot order
        DBuf<int> din = upload_vec(ctx_, isin);
        const int* IN = din.get();
        const int* NS = P_.nslot.get();
        const int* AR = P_.A.row.get();
        const int* AC = P_.A.col.get();
        const float* AW = P_.A.val.get();
        DBuf<int> gr = ctx_.buf<int>(mh), gc = ctx_.buf<int>(mh);
        DBuf<float> gw = ctx_.buf<float>(mh);
        int* GR = gr.get();
        int* GC = gc.get();
        float* GW = gw.get();
        DBuf<int> cnt = ctx_.buf<int>(1);
        select_into(
            ctx_, P_.A.nnz, [=] __device__(i64 e) { return IN[NS[AR[e]]] != 0; },
            [=] __device__(i64 e, int p) {
                GR[p] = AR[e];
                GC[p] = AC[e];
                GW[p] = AW[e];
            },
            cnt.get());
        ctx_.d2h(er.data(), gr.get(), mh);
        ctx_.d2h(ec.data(), gc.get(), mh);
        ctx_.d2h(ev.data(), gw.get(), mh);
    }
    const std::vector<int32_t> hl = gather_labels(ctx_, P_, which);
    i64 lo = 0;
    for (size_t x = 0; x < which.size(); ++x) {
        const int q = which[x];
        const Slot& s = P_.s[q];
        Handoff h;
        h.id = s.id;
        h.next_iter = s.iter;
        h.rows = s.rows;
        h.cols = s.cols;
        h.max_node = s.max_node;
        h.lb = s.lb;
        h.labels.assign(hl.begin() + lo, hl.begin() + lo + s.n0);
        lo += s.n0;
        h.row.resize((size_t)s.nnz);
        h.col.resize((size_t)s.nnz);
        h.val.resize((size_t)s.nnz);
        for (i64 e = 0; e < s.nnz; ++e) {
            h.row[e] = er[hoff[x] + e] - P_.base[q];
            h.col[e] = ec[hoff[x] + e] - P_.base[q];
            h.val[e] = ev[hoff[x] + e];
        }
        hand_off(std::move(h));
    }
}

void Lockstep::admit(int k, const int64_t* edge_off, const int32_t* i, const int32_t* j, const float* c,
                     const int* ids) {
    if (k <= 0) return;
    const i64 mall = edge_off[k] - edge_off[0];
    const int64_t e0 = edge_off[0];
    std::vector<int> hchk((size_t)2 * k, 0);
    DBuf<int> ri, rj;
    DBuf<float> rc;
    DBuf<i64> deoff;
    if (mall > 0) {
        // raw edges up, then per-instance checks (max id, negative ids)
        ri = ctx_.buf<int>(mall);
        rj = ctx_.buf<int>(mall);
        rc = ctx_.buf<float>(mall);
        ctx_.h2d(ri.get(), i + e0, mall);
        ctx_.h2d(rj.get(), j + e0, mall);
        ctx_.h2d(rc.get(), c + e0, mall);
        std::vector<i64> eoff((size_t)k + 1);
        for (int q = 0; q <= k; ++q) eoff[q] = edge_off[q] - e0;
        deoff = upload_vec(ctx_, eoff);
        DBuf<int> chk = ctx_.buf<int>(2 * (i64)k);
        int* CK = chk.get();
        for_each(ctx_, 2 * (i64)k, [=] __device__(i64 x) { CK[x] = x < k ? -1 : 0; });
        const int* I = ri.get();
        const int* J = rj.get();
        range_reduce<CheckAcc>(
            ctx_, deoff.get(), k,
            [=] __device__(i64 e, CheckAcc& a) {
                const int x = I[e] > J[e] ? I[e] : J[e];
                a.mx = x > a.mx ? x : a.mx;
                a.neg |= (I[e] < 0 || J[e] < 0) ? 1 : 0;
            },
            CheckRed(),
            [=] __device__(int q, CheckAcc r) {
                atomicMax(&CK[q], r.mx);
                if (r.neg) CK[k + q] = 1;
            });
        ctx_.d2h(hchk.data(), CK, 2 * (i64)k);
    }
    const int Q0 = P_.Q();
    std::vector<int> ibase((size_t)k, -1);
    i64 m = 0, off = P_.span_total;
    for (int q = 0; q < k; ++q) {
        const i64 mq = edge_off[q + 1] - edge_off[q];
        if (mq <= 0) {
            sink_.done(ids[q], RAMA_ERR_INVALID, "the graph has no edges", nullptr, 0, 0.0);
            continue;
        }
        if (hchk[k + q]) {
            sink_.done(ids[q], RAMA_ERR_INVALID, "node ids must be non-negative", nullptr, 0, 0.0);
            continue;
        }
        if (P_.Q() >= kMaxSlots) {
            sink_.done(ids[q], RAMA_ERR_INVALID, "too many instances in one lockstep pack", nullptr, 0, 0.0);
            continue;
        }
        if (off + hchk[q] + 1 >= ((i64)1 << 31)) {
            sink_.done(ids[q], RAMA_ERR_INVALID, "batch too large for 32-bit node ids", nullptr, 0, 0.0);
            continue;
        }
        Slot s{};
        s.id = ids[q];
        s.span = hchk[q] + 1;
        ibase[q] = (int)off;
        off += s.span;
        m += mq;
        P_.s.push_back(s);
    }
    const int Q = P_.Q();
    if (Q == Q0) return;
    {
        // the new instances' edges with global ids above every admitted one,
        // sorted among themselves: appending keeps the joint list sorted
        DBuf<int> dib = upload_vec(ctx_, ibase);
        const int* IB = dib.get();
        const int* I = ri.get();
        const int* J = rj.get();
        const float* Cc = rc.get();
        const i64* EO = deoff.get();
        DBuf<int> di = ctx_.buf<int>(m), dj = ctx_.buf<int>(m);
        DBuf<float> dc = ctx_.buf<float>(m);
        int* GI = di.get();
        int* GJ = dj.get();
        float* GC = dc.get();
        DBuf<int> cnt = ctx_.buf<int>(1);
        select_into(
            ctx_, mall, [=] __device__(i64 e) { return IB[seg_of(EO, k, e)] >= 0; },
            [=] __device__(i64 e, int p) {
                const int b = IB[seg_of(EO, k, e)];
                GI[p] = I[e] + b;
                GJ[p] = J[e] + b;
                GC[p] = Cc[e];
            },
            cnt.get());
        ri.reset();
        rj.reset();
        rc.reset();
        // dCOO(col=i, row=j) + init; the orientation makes both the same
        DCoo An = dcoo_from_edges(ctx_, dj.get(), di.get(), dc.get(), m);
        if (Q0 == 0) {
            P_.A = std::move(An);
        } else {
            DCoo A2;
            A2.nnz = P_.A.nnz + An.nnz;
            A2.row = ctx_.buf<int>(A2.nnz);
            A2.col = ctx_.buf<int>(A2.nnz);
            A2.val = ctx_.buf<float>(A2.nnz);
            ctx_.d2d(A2.row.get(), P_.A.row.get(), P_.A.nnz);
            ctx_.d2d(A2.col.get(), P_.A.col.get(), P_.A.nnz);
            ctx_.d2d(A2.val.get(), P_.A.val.get(), P_.A.nnz);
            ctx_.d2d(A2.row.get() + P_.A.nnz, An.row.get(), An.nnz);
            ctx_.d2d(A2.col.get() + P_.A.nnz, An.col.get(), An.nnz);
            ctx_.d2d(A2.val.get() + P_.A.nnz, An.val.get(), An.nnz);
            P_.A = std::move(A2);
        }
    }
    std::vector<char> fresh((size_t)Q, 0);
    for (int q = Q0; q < Q; ++q) fresh[q] = 1;
    pack_index(ctx_, P_);
    {
        DBuf<SlotStats> st = ctx_.buf<SlotStats>(Q);
        slot_stats_edges(ctx_, P_.A.row.get(), P_.A.col.get(), P_.A.nnz, P_.dbase.get(), Q, st.get());
        apply_stats(ctx_, P_, st.get(), fresh);
    }
    for (int q = Q0; q < Q; ++q) P_.s[q].n0 = std::max(P_.s[q].rows, P_.s[q].cols);
    mark("admitted", 0, Q);
    // initial dual of the new instances (the bound)
    {
        const std::vector<double> lb = packed_dual(ctx_, P_, o_.max_cycle_length_lb, o_.num_dual_itr_lb,
                                                   o_.tri_memory_factor, o_.num_outer_itr_dual, 1e-4f, fresh);
        for (int q = Q0; q < Q; ++q) P_.s[q].lb = lb[q];
    }
    mark("initial dual", 0, Q);
    // labels: the old ones, then identity labels for the new instances
    {
        pack_index(ctx_, P_);  // label offsets with the new instances' n0
        DBuf<int> L2 = ctx_.buf<int>(std::max<i64>(P_.lab[Q], 1));
        if (Q0 > 0 && P_.lab[Q0] > 0) ctx_.d2d(L2.get(), P_.labels.get(), P_.lab[Q0]);
        int* L = L2.get();
        DBuf<i64> dlab = upload_vec(ctx_, P_.lab);
        const i64* DL = dlab.get();
        const i64 from = P_.lab[Q0];
        for_each(ctx_, P_.lab[Q] - from, [=] __device__(i64 y) {
            const i64 x = from + y;
            L[x] = (int)(x - DL[seg_of(DL, Q, x)]);
        });
        P_.labels = std::move(L2);
    }
    if (o_.only_compute_lb) {
        std::vector<char> keep((size_t)Q, 1);
        for (int q = Q0; q < Q; ++q) {
            sink_.done(P_.s[q].id, 0, "", nullptr, 0, P_.s[q].lb);
            keep[q] = 0;
        }
        compact(keep);
        return;
    }
    if (o_.matching_thresh_crossover_ratio > 1.0f) {
        // the forest contraction from round 0 on: straight to the workers
        std::vector<int> which;
        std::vector<char> keep((size_t)Q, 1);
        for (int q = Q0; q < Q; ++q) {
            which.push_back(q);
            keep[q] = 0;
        }
        hand_off_slots(which);
        compact(keep);
    }
}

void Lockstep::round() {
    int Q = P_.Q();
    if (Q == 0) return;
    mark("round", P_.s[0].iter, Q);
    // the primal dual of every slot past its first round
    {
        std::vector<char> on((size_t)Q, 0);
        bool any = false;
        for (int q = 0; q < Q; ++q) {
            on[q] = P_.s[q].iter > 0;
            any = any || on[q];
        }
        if (any)
            packed_dual(ctx_, P_, o_.max_cycle_length_primal, o_.num_dual_itr_primal, 1.0f, 1, 1e-4f, on);
    }
    mark("s:dual", 0, Q);
    // matching on the joint symmetric CSR
    Csr U = symmetric_csr(ctx_, P_.A.row.get(), P_.A.col.get(), P_.A.val.get(), P_.A.nnz, P_.span_total, false, 0.0f);
    DBuf<i64> dhb = ctx_.buf<i64>(Q + 1);
    {
        const int* OFF = U.off.get();
        const int* BS = P_.dbase.get();
        i64* HB = dhb.get();
        for_each(ctx_, Q + 1, [=] __device__(i64 q) { HB[q] = OFF[BS[q]]; });
    }
    std::vector<i64> hb((size_t)Q + 1);
    ctx_.d2h(hb.data(), dhb.get(), Q + 1);
    DBuf<int> map;
    std::vector<int> matched((size_t)Q, 0);
    matching_packed(ctx_, U, P_.nslot.get(), Q, hb.data(), dhb.get(), o_.mean_multiplier_mm, map,
                    matched.data());
    U = Csr();
    mark("s:matching", 0, Q);
    // per slot: finished (nothing to contract) or contracted, maybe switching
    std::vector<int> un((size_t)Q, 0), cont((size_t)Q, 0);
    std::vector<char> sw((size_t)Q, 0);
    std::vector<int> finished;
    for (int q = 0; q < Q; ++q) {
        if (matched[q] == 0) {
            finished.push_back(q);
            continue;
        }
        cont[q] = 1;
        un[q] = P_.s[q].max_node + 1;
        sw[q] = (float)matched[q] < (float)P_.s[q].rows * o_.matching_thresh_crossover_ratio;
    }
    if (!finished.empty()) publish(finished);
    if (finished.size() == (size_t)Q) {
        P_ = Pack();
        return;
    }
    DBuf<int> dun = upload_vec(ctx_, un), dcont = upload_vec(ctx_, cont);
    const int* UN = dun.get();
    const int* CT = dcont.get();
    const int* NS = P_.nslot.get();
    const int* BS = P_.dbase.get();
    const int* MP = map.get();
    // compress_label_sequence of the matching map, per slot
    const i64 n = P_.span_total;
    DBuf<int> rank = ctx_.buf<int>(n + 1);
    int* RK = rank.get();
    scan_exclusive<int>(
        ctx_,
        [=] __device__(i64 v) {
            const int q = NS[v];
            return (int)(CT[q] && (int)v - BS[q] < UN[q] && MP[v] == (int)v);
        },
        RK, n);
    // new bases (rank at the slot's base) and spans, read back
    DBuf<int> dnb = ctx_.buf<int>(2 * (i64)Q);
    {
        int* NB = dnb.get();
        for_each(ctx_, Q, [=] __device__(i64 q) {
            NB[q] = RK[BS[q]];
            NB[Q + q] = RK[BS[q] + UN[q]] - RK[BS[q]];
        });
    }
    std::vector<int> hnb((size_t)2 * Q);
    ctx_.d2h(hnb.data(), dnb.get(), 2 * (i64)Q);
    // labels (map_node_labels per slot: entries below Un move to the
    // compressed label of their representative)
    {
        int* L = P_.labels.get();
        DBuf<i64> dlab = upload_vec(ctx_, P_.lab);
        const i64* DL = dlab.get();
        for_each(ctx_, P_.lab[Q], [=] __device__(i64 x) {
            const int q = seg_of(DL, Q, x);
            if (!CT[q]) return;
            const int v = L[x];
            if (v < UN[q]) L[x] = RK[MP[BS[q] + v]] - RK[BS[q]];
        });
    }
    // edges of finished slots leave before the contraction
    if (!finished.empty()) {
        std::vector<char> keep((size_t)Q, 0);
        for (int q = 0; q < Q; ++q) keep[q] = cont[q] != 0;
        DCoo A2;
        const int* AR = P_.A.row.get();
        const int* AC = P_.A.col.get();
        const float* AW = P_.A.val.get();
        i64 m2 = 0;
        for (int q = 0; q < Q; ++q)
            if (cont[q]) m2 += P_.s[q].nnz;
        A2.nnz = m2;
        A2.row = ctx_.buf<int>(std::max<i64>(m2, 1));
        A2.col = ctx_.buf<int>(std::max<i64>(m2, 1));
        A2.val = ctx_.buf<float>(std::max<i64>(m2, 1));
        int* R2 = A2.row.get();
        int* C2 = A2.col.get();
        float* W2 = A2.val.get();
        DBuf<int> cnt = ctx_.buf<int>(1);
        select_into(
            ctx_, P_.A.nnz, [=] __device__(i64 e) { return CT[NS[AR[e]]] != 0; },
            [=] __device__(i64 e, int p) {
                R2[p] = AR[e];
                C2[p] = AC[e];
                W2[p] = AW[e];
            },
            cnt.get());
        P_.A = std::move(A2);
    }
    // contraction of the continuing slots (new slot index = rank among them)
    std::vector<int> cidx, nbase_new;
    for (int q = 0; q < Q; ++q)
        if (cont[q]) {
            cidx.push_back(q);
            nbase_new.push_back(hnb[q]);
        }
    const int Qc = (int)cidx.size();
    int total_new = 0;
    for (int q : cidx) total_new += hnb[Q + q];
    nbase_new.push_back(total_new);
    DBuf<int> dnbn = upload_vec(ctx_, nbase_new);
    DBuf<int> nns = ctx_.buf<int>(std::max(total_new, 1));
    if (total_new > 0) {
        k_node_slot<<<grid_for(total_new), kBlock, 0, ctx_.stream>>>(dnbn.get(), Qc, total_new, nns.get());
        RAMA_LAUNCHED();
    }
    DBuf<int> nmap = ctx_.buf<int>(std::max<i64>(n, 1));
    {
        int* NM = nmap.get();
        for_each(ctx_, n, [=] __device__(i64 v) { NM[v] = RK[MP[v]]; });
    }
    DCoo Bc;
    DBuf<SlotStats> st = ctx_.buf<SlotStats>(Qc);
    contract_packed(ctx_, P_.A, nmap.get(), std::max(total_new - 1, 0), nns.get(), dnbn.get(), Qc, Bc, st.get());
    std::vector<SlotStats> hst((size_t)Qc);
    ctx_.d2h(hst.data(), st.get(), Qc);
    // the pack continues over the contracted slots in the new id space
    std::vector<i64> off, cnt_lab;
    std::vector<Slot> ns;
    i64 nnz = 0;
    for (int c = 0; c < Qc; ++c) {
        Slot s = P_.s[cidx[c]];
        off.push_back(P_.lab[cidx[c]]);
        cnt_lab.push_back(s.n0);
        s.cols = hst[c].max_col + 1;
        s.rows = hst[c].max_row + 1;
        s.max_node = hst[c].max_kept_col;
        s.nnz = hst[c].kept;
        s.span = hnb[Q + cidx[c]];
        s.iter += 1;
        nnz += s.nnz;
        ns.push_back(s);
    }
    Bc.nnz = nnz;
    DBuf<int> labels = gather_ranges(ctx_, P_.labels.get(), off, cnt_lab);
    std::vector<char> swc;
    for (int c = 0; c < Qc; ++c) swc.push_back(sw[cidx[c]]);
    P_.s = ns;
    P_.A = std::move(Bc);
    P_.labels = std::move(labels);
    pack_index(ctx_, P_);
    mark("s:contract", 0, Qc);
    // slots leaving: failed (bad contraction), done (no edges left), handed
    // off (switching to the forest contraction)
    std::vector<int> done, hand;
    std::vector<char> keep((size_t)Qc, 1);
    for (int c = 0; c < Qc; ++c) {
        if (hst[c].bad) {
            sink_.done(P_.s[c].id, RAMA_ERR_BAD_CONTRACTION, "Found bad contractions", nullptr, 0, P_.s[c].lb);
            keep[c] = 0;
        } else if (P_.s[c].nnz == 0) {
            done.push_back(c);  // A.nnz == 0 ends the rounds
            keep[c] = 0;
        } else if (swc[c]) {
            hand.push_back(c);
            keep[c] = 0;
        }
    }
    if (!done.empty()) publish(done);
    hand_off_slots(hand);
    compact(keep);
}

namespace {
struct BatchSink : LockstepSink {
    PackedBatch& B;
    explicit BatchSink(PackedBatch& b) : B(b) {}
    void done(int id, int status, const std::string& error, const int32_t* labels, i64 n, double lb) override {
        B.lbs[id] = lb;
        if (status == 0) {
            const i64 cap = B.label_off[id + 1] - B.label_off[id];
            if (n > cap) {
                status = RAMA_ERR_INVALID;
            } else {
                B.n_labels[id] = n;
                if (n > 0) std::copy(labels, labels + n, B.labels_out + B.label_off[id]);
            }
        }
        if (B.status) B.status[id] = status;
        if (status) {
            std::lock_guard<std::mutex> l(*B.emu);
            if (B.first_error->empty())
                *B.first_error = status == RAMA_ERR_INVALID && error.empty()
                                     ? "label buffer too small for instance " + std::to_string(id)
                                     : error;
        }
    }
};
}  // namespace

// Solves B.k instances in lockstep on ctx; instances that switch to the
// forest contraction are continued by `workers` threads.  Per-instance
// failures go to status; a failure of the whole pack throws.
void solve_packed(Ctx& ctx, PackedBatch& B, int workers) {
    scan_reset(ctx);
    BatchSink sink(B);
    Lockstep L(ctx, B.o, workers, sink);
    RAMA_REQUIRE(B.k <= Lockstep::kMaxSlots, "at most 4096 instances per packed batch");
    std::vector<int> ids((size_t)B.k);
    for (int q = 0; q < B.k; ++q) ids[q] = q;
    L.admit(B.k, B.edge_off, B.i, B.j, B.c, ids.data());
    while (!L.empty()) L.round();
    scan_check(ctx);
    L.finish();
}

// Opaque handle for callers outside this file (the serving loop).
struct LockstepHandle {
    Lockstep engine;
    LockstepHandle(Ctx& ctx, const Options& o, int workers, LockstepSink& sink) : engine(ctx, o, workers, sink) {}
};
LockstepHandle* lockstep_create(Ctx& ctx, const Options& o, int workers, LockstepSink& sink) {
    return new LockstepHandle(ctx, o, workers, sink);
}
void lockstep_admit(LockstepHandle* h, int k, const int64_t* edge_off, const int32_t* i, const int32_t* j,
                    const float* c, const int* ids) {
    h->engine.admit(k, edge_off, i, j, c, ids);
}
void lockstep_round(LockstepHandle* h) { h->engine.round(); }
int lockstep_slots(const LockstepHandle* h) { return h->engine.slots(); }
void lockstep_destroy(LockstepHandle* h) { delete h; }

}  // namespace rama
