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
DBuf<T> upload_vec(Ctx& ctx, const std::vector<T>& h) {
    DBuf<T> d = ctx.buf<T>(std::max<i64>((i64)h.size(), 1));
    ctx.upload(d.get(), h.data(), (i64)h.size());
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
std::vector<double> packed_dcoo_lb(Ctx& ctx, Pack& P, unsigned* tickets) {
    const int Q = P.Q();
    std::vector<double> out((size_t)Q, 0.0);
    const std::vector<i64> eb = edge_prefix(P);
    DBuf<i64> deb = upload_vec(ctx, eb);
    DBuf<double> d = ctx.buf<double>(Q);
    const float* W = P.A.val.get();
    seg_reduce<double>(
        ctx, [=] __device__(i64 e) { const float x = W[e]; return x < 0.0f ? (double)x : 0.0; }, eb.data(),
        deb.get(), Q, 0.0, OpSum(), d.get(), 1, tickets);
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
                              const std::vector<char>& part, unsigned* tickets) {
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
                        std::vector<double>& fin, std::vector<i64>& num_tri, unsigned* tickets) {
    const int Q = P.Q();
    PackedTriangles T = packed_cycles(ctx, P, c, factor, tol, on, tickets);
    bool need_lb = false, any_mp = false;
    std::vector<char> mp_on((size_t)Q, 0);
    for (int q = 0; q < Q; ++q) {
        if (!on[q]) continue;
        num_tri[q] = T.raw[q];
        if (T.raw[q] == 0) need_lb = true;
        else mp_on[q] = 1, any_mp = true;
    }
    if (need_lb) {
        const std::vector<double> lb = packed_dcoo_lb(ctx, P, tickets);
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
    std::vector<int> tblk((size_t)Q + 1, 0), eblk((size_t)Q + 1, 0), onv((size_t)Q, 0);
    std::vector<i64> pbeg((size_t)Q, 0);
    for (int q = 0; q < Q; ++q) {
        const int gt = (int)grid_for(tb[q + 1] - tb[q]), ge = (int)grid_for(eb[q + 1] - eb[q]);
        tblk[q + 1] = tblk[q] + gt;
        eblk[q + 1] = eblk[q] + ge;
        pbeg[q] = (q ? pbeg[q - 1] + (tblk[q] - tblk[q - 1]) + (eblk[q] - eblk[q - 1]) : 0);
        onv[q] = mp_on[q];
    }
    DBuf<int> dtblk = upload_vec(ctx, tblk), deblk = upload_vec(ctx, eblk), don = upload_vec(ctx, onv);
    DBuf<i64> dpbeg = upload_vec(ctx, pbeg);
    MPPack pk{Q, dtb.get(), deb.get(), dtblk.get(), deblk.get(), dpbeg.get()};
    std::vector<double> lb((size_t)Q, 0.0);
    mp_run_packed(ctx, S, pk, tblk[Q], eblk[Q], steps, tb.data(), eb.data(), don.get(), tickets, lb.data());
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
                                const std::vector<char>& part, unsigned* tickets) {
    const int Q = P.Q();
    std::vector<double> fin((size_t)Q, 0.0), prev((size_t)Q, 0.0);
    if (max_len < 3 || num_iter == 0 || num_outer == 0) {
        const std::vector<double> lb = packed_dcoo_lb(ctx, P, tickets);
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
        for (int c = 3; c <= max_len; ++c) packed_dual_update(ctx, P, c, num_iter, factor, tol, act, fin, num_tri, tickets);
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

static void set_status(PackedBatch& B, int id, int st, const std::string& msg) {
    if (B.status) B.status[id] = st;
    if (st) {
        std::lock_guard<std::mutex> l(*B.emu);
        if (B.first_error->empty()) *B.first_error = msg;
    }
}

static void write_labels(PackedBatch& B, int id, const int32_t* lab, i64 n) {
    B.n_labels[id] = n;
    const i64 cap = B.label_off[id + 1] - B.label_off[id];
    if (n > cap) throw Error(RAMA_ERR_INVALID, "label buffer too small for instance " + std::to_string(id));
    if (n > 0) std::copy(lab, lab + n, B.labels_out + B.label_off[id]);
}

// Solves B.k instances in lockstep on ctx; instances that switch to the
// forest contraction are continued by `workers` threads.  Per-instance
// failures go to status; a failure of the whole pack throws.
void solve_packed(Ctx& ctx, PackedBatch& B, int workers) {
    const Options& o = B.o;
    // RAMA_STAGE_TIMES: wall time of the packed phases on stderr (diagnostic)
    const bool times = std::getenv("RAMA_STAGE_TIMES") != nullptr;
    const auto tstart = std::chrono::steady_clock::now();
    std::map<std::string, double> stage_ms;
    auto tlast = tstart;
    auto mark = [&](const char* what, int rounds, int slots) {
        if (!times) return;
        ctx.sync();
        const auto now = std::chrono::steady_clock::now();
        stage_ms[what] += std::chrono::duration<double, std::milli>(now - tlast).count();
        tlast = now;
        if (rounds >= 0 && std::string(what) != "round") return;
        const double ms = std::chrono::duration<double, std::milli>(now - tstart).count();
        std::fprintf(stderr, "packed %-14s %9.3f ms  round %d  slots %d\n", what, ms, rounds, slots);
    };
    const int k = B.k;
    // the worker pool for handed-off instances
    std::mutex qmu;
    std::condition_variable qcv;
    std::deque<Handoff> tasks;
    bool closing = false;
    std::vector<std::thread> pool;
    auto worker = [&] {
        std::unique_ptr<Ctx> wctx;
        for (;;) {
            Handoff h;
            {
                std::unique_lock<std::mutex> l(qmu);
                qcv.wait(l, [&] { return closing || !tasks.empty(); });
                if (tasks.empty()) return;
                h = std::move(tasks.front());
                tasks.pop_front();
            }
            try {
                if (!wctx) wctx.reset(new Ctx(ctx.device, nullptr, false));
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
                solve_rounds(w, A, out, o, h.next_iter, false);
                std::vector<int32_t> lab((size_t)out.num_labels);
                if (out.num_labels > 0) w.d2h(lab.data(), out.labels.get(), out.num_labels);
                write_labels(B, h.id, lab.data(), out.num_labels);
                B.lbs[h.id] = h.lb;
                set_status(B, h.id, 0, "");
            } catch (const Error& e) {
                set_status(B, h.id, e.status, e.what());
            } catch (const std::exception& e) {
                set_status(B, h.id, RAMA_ERR_INTERNAL, e.what());
            }
        }
    };
    for (int w = 0; w < std::max(1, workers); ++w) pool.emplace_back(worker);
    auto hand_off = [&](Handoff&& h) {
        {
            std::lock_guard<std::mutex> l(qmu);
            tasks.push_back(std::move(h));
        }
        qcv.notify_one();
    };
    auto finish_pool = [&] {
        {
            std::lock_guard<std::mutex> l(qmu);
            closing = true;
        }
        qcv.notify_all();
        for (auto& t : pool) t.join();
    };

    try {
        // ---------------------------------------------- pack the inputs
        // raw edges up, then per-instance checks (max id, negative ids) and
        // the id offsets on the device
        Pack P;
        const i64 mall = B.edge_off[k];
        if (mall <= 0) {
            for (int q = 0; q < k; ++q) set_status(B, q, RAMA_ERR_INVALID, "the graph has no edges");
            finish_pool();
            return;
        }
        DBuf<int> ri = ctx.buf<int>(mall), rj = ctx.buf<int>(mall);
        DBuf<float> rc = ctx.buf<float>(mall);
        ctx.h2d(ri.get(), B.i, mall);
        ctx.h2d(rj.get(), B.j, mall);
        ctx.h2d(rc.get(), B.c, mall);
        std::vector<i64> eoff(B.edge_off, B.edge_off + k + 1);
        DBuf<i64> deoff = upload_vec(ctx, eoff);
        DBuf<int> chk = ctx.buf<int>(2 * (i64)k);  // [0, k) max id, [k, 2k) negative flag
        {
            int* CK = chk.get();
            for_each(ctx, 2 * (i64)k, [=] __device__(i64 x) { CK[x] = x < k ? -1 : 0; });
            const int* I = ri.get();
            const int* J = rj.get();
            range_reduce<CheckAcc>(
                ctx, deoff.get(), k,
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
        }
        std::vector<int> hchk((size_t)2 * k);
        ctx.d2h(hchk.data(), chk.get(), 2 * (i64)k);
        std::vector<int> ibase((size_t)k, -1);
        i64 m = 0, off = 0;
        for (int q = 0; q < k; ++q) {
            const i64 mq = B.edge_off[q + 1] - B.edge_off[q];
            if (mq <= 0) {
                set_status(B, q, RAMA_ERR_INVALID, "the graph has no edges");
                continue;
            }
            if (hchk[k + q]) {
                set_status(B, q, RAMA_ERR_INVALID, "node ids must be non-negative");
                continue;
            }
            Slot s{};
            s.id = q;
            s.span = hchk[q] + 1;
            ibase[q] = (int)off;
            off += s.span;
            m += mq;
            RAMA_REQUIRE(off < ((i64)1 << 31), "batch too large for 32-bit node ids");
            P.s.push_back(s);
        }
        if (P.Q() == 0) {
            finish_pool();
            return;
        }
        int Q = P.Q();
        {
            // the joint edge list with global ids (instances that failed the checks dropped)
            DBuf<int> dib = upload_vec(ctx, ibase);
            const int* IB = dib.get();
            const int* I = ri.get();
            const int* J = rj.get();
            const float* Cc = rc.get();
            const i64* EO = deoff.get();
            DBuf<int> di = ctx.buf<int>(m), dj = ctx.buf<int>(m);
            DBuf<float> dc = ctx.buf<float>(m);
            int* GI = di.get();
            int* GJ = dj.get();
            float* GC = dc.get();
            DBuf<int> cnt = ctx.buf<int>(1);
            select_into(
                ctx, mall, [=] __device__(i64 e) { return IB[seg_of(EO, k, e)] >= 0; },
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
            P.A = dcoo_from_edges(ctx, dj.get(), di.get(), dc.get(), m);
        }
        pack_index(ctx, P);
        {
            DBuf<SlotStats> st = ctx.buf<SlotStats>(Q);
            slot_stats_clear(ctx, st.get(), Q);
            slot_stats_edges(ctx, P.A.row.get(), P.A.col.get(), P.A.nnz, P.dbase.get(), Q, st.get());
            apply_stats(ctx, P, st.get(), std::vector<char>((size_t)Q, 1));
        }
        for (int q = 0; q < Q; ++q) P.s[q].n0 = std::max(P.s[q].rows, P.s[q].cols);
        DBuf<unsigned> tickets = ctx.buf<unsigned>(4096);
        ctx.zero(tickets.get(), 4096);
        RAMA_REQUIRE(k <= 4096, "at most 4096 instances per packed batch");

        mark("packed", 0, Q);
        // ------------------------------------------- initial dual (bound)
        {
            const std::vector<double> lb =
                packed_dual(ctx, P, o.max_cycle_length_lb, o.num_dual_itr_lb, o.tri_memory_factor,
                            o.num_outer_itr_dual, 1e-4f, std::vector<char>((size_t)Q, 1), tickets.get());
            for (int q = 0; q < Q; ++q) {
                P.s[q].lb = lb[q];
                B.lbs[P.s[q].id] = lb[q];
            }
        }
        mark("initial dual", 0, Q);
        if (o.only_compute_lb) {
            for (int q = 0; q < Q; ++q) {
                B.n_labels[P.s[q].id] = 0;
                set_status(B, P.s[q].id, 0, "");
            }
            finish_pool();
            return;
        }
        pack_index(ctx, P);  // label offsets from n0
        P.labels = ctx.buf<int>(std::max<i64>(P.lab[Q], 1));
        {
            int* L = P.labels.get();
            DBuf<i64> dlab = upload_vec(ctx, P.lab);
            const i64* DL = dlab.get();
            for_each(ctx, P.lab[Q], [=] __device__(i64 x) { L[x] = (int)(x - DL[seg_of(DL, Q, x)]); });
        }
        const bool use_mm = !(o.matching_thresh_crossover_ratio > 1.0f);

        // ------------------------------------------------- the rounds
        for (int iter = 0; P.Q() > 0; ++iter) {
            Q = P.Q();
            mark("round", iter, Q);
            std::vector<char> leave((size_t)Q, 0), handoff((size_t)Q, 0), contract_on((size_t)Q, 0);
            if (!use_mm) {
                for (int q = 0; q < Q; ++q) handoff[q] = 1;
            } else {
                if (iter > 0)
                    packed_dual(ctx, P, o.max_cycle_length_primal, o.num_dual_itr_primal, 1.0f, 1, 1e-4f,
                                std::vector<char>((size_t)Q, 1), tickets.get());
                mark("s:dual", iter, Q);
                // matching on the joint symmetric CSR
                Csr U = symmetric_csr(ctx, P.A.row.get(), P.A.col.get(), P.A.val.get(), P.A.nnz, P.span_total,
                                      false, 0.0f);
                DBuf<i64> dhb = ctx.buf<i64>(Q + 1);
                {
                    const int* OFF = U.off.get();
                    const int* BS = P.dbase.get();
                    i64* HB = dhb.get();
                    for_each(ctx, Q + 1, [=] __device__(i64 q) { HB[q] = OFF[BS[q]]; });
                }
                std::vector<i64> hb((size_t)Q + 1);
                ctx.d2h(hb.data(), dhb.get(), Q + 1);
                DBuf<int> map;
                std::vector<int> matched((size_t)Q, 0);
                matching_packed(ctx, U, P.nslot.get(), Q, hb.data(), dhb.get(), o.mean_multiplier_mm,
                                tickets.get(), map, matched.data());
                U = Csr();
                mark("s:matching", iter, Q);
                // per slot: finished, or contraction (and maybe the switch)
                std::vector<int> un((size_t)Q, 0), cont((size_t)Q, 0);
                bool any_done = false, any_cont = false;
                for (int q = 0; q < Q; ++q) {
                    if (matched[q] == 0) {
                        leave[q] = 1;
                        any_done = true;
                        continue;
                    }
                    contract_on[q] = 1;
                    cont[q] = 1;
                    any_cont = true;
                    un[q] = P.s[q].max_node + 1;
                    if ((float)matched[q] < (float)P.s[q].rows * o.matching_thresh_crossover_ratio) handoff[q] = 1;
                }
                if (any_cont) {
                    DBuf<int> dun = upload_vec(ctx, un), dcont = upload_vec(ctx, cont);
                    const int* UN = dun.get();
                    const int* CT = dcont.get();
                    const int* NS = P.nslot.get();
                    const int* BS = P.dbase.get();
                    const int* MP = map.get();
                    // compress_label_sequence of the matching map, per slot
                    const i64 n = P.span_total;
                    DBuf<int> rank = ctx.buf<int>(n + 1);
                    int* RK = rank.get();
                    scan_exclusive<int>(
                        ctx,
                        [=] __device__(i64 v) {
                            const int q = NS[v];
                            return (int)(CT[q] && (int)v - BS[q] < UN[q] && MP[v] == (int)v);
                        },
                        RK, n);
                    // new bases (rank at the slot's base) and spans, read back
                    DBuf<int> dnb = ctx.buf<int>(2 * (i64)Q);
                    int* NB = dnb.get();
                    for_each(ctx, Q, [=] __device__(i64 q) {
                        NB[q] = RK[BS[q]];
                        NB[Q + q] = RK[BS[q] + UN[q]] - RK[BS[q]];
                    });
                    std::vector<int> hnb((size_t)2 * Q);
                    ctx.d2h(hnb.data(), NB, 2 * (i64)Q);
                    // edges of finished slots leave before the contraction
                    if (any_done) {
                        DCoo A2;
                        const int* AR = P.A.row.get();
                        const int* AC = P.A.col.get();
                        const float* AW = P.A.val.get();
                        A2.row = ctx.buf<int>(P.A.nnz);
                        A2.col = ctx.buf<int>(P.A.nnz);
                        A2.val = ctx.buf<float>(P.A.nnz);
                        int* R2 = A2.row.get();
                        int* C2 = A2.col.get();
                        float* W2 = A2.val.get();
                        DBuf<int> cnt = ctx.buf<int>(1);
                        select_into(
                            ctx, P.A.nnz, [=] __device__(i64 e) { return CT[NS[AR[e]]] != 0; },
                            [=] __device__(i64 e, int p) {
                                R2[p] = AR[e];
                                C2[p] = AC[e];
                                W2[p] = AW[e];
                            },
                            cnt.get());
                        i64 m2 = 0;
                        for (int q = 0; q < Q; ++q)
                            if (cont[q]) m2 += P.s[q].nnz;
                        A2.nnz = m2;
                        P.A = std::move(A2);
                    }
                    // contraction of the continuing slots (new slot = rank among them)
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
                    DBuf<int> dnbn = upload_vec(ctx, nbase_new);
                    DBuf<int> nns = ctx.buf<int>(std::max(total_new, 1));
                    if (total_new > 0) {
                        k_node_slot<<<grid_for(total_new), kBlock, 0, ctx.stream>>>(dnbn.get(), Qc, total_new,
                                                                                     nns.get());
                        RAMA_LAUNCHED();
                    }
                    DBuf<int> nmap = ctx.buf<int>(std::max<i64>(n, 1));
                    {
                        int* NM = nmap.get();
                        for_each(ctx, n, [=] __device__(i64 v) { NM[v] = RK[MP[v]]; });
                    }
                    DCoo Bc;
                    DBuf<SlotStats> st = ctx.buf<SlotStats>(Qc);
                    contract_packed(ctx, P.A, nmap.get(), std::max(total_new - 1, 0), nns.get(), dnbn.get(), Qc,
                                    Bc, st.get());
                    std::vector<SlotStats> hst((size_t)Qc);
                    ctx.d2h(hst.data(), st.get(), Qc);
                    // labels (map_node_labels per slot: entries below Un move)
                    {
                        int* L = P.labels.get();
                        DBuf<i64> dlab = upload_vec(ctx, P.lab);
                        const i64* DL = dlab.get();
                        const i64 nl = P.lab[Q];
                        for_each(ctx, nl, [=] __device__(i64 x) {
                            const int q = seg_of(DL, Q, x);
                            if (!CT[q]) return;
                            const int m = L[x];
                            // cur[m] = compressed label of m's representative
                            if (m < UN[q]) L[x] = RK[MP[BS[q] + m]] - RK[BS[q]];
                        });
                    }
                    i64 nnz = 0;
                    for (int c = 0; c < Qc; ++c) {
                        Slot& s = P.s[cidx[c]];
                        if (hst[c].bad) {
                            set_status(B, s.id, RAMA_ERR_BAD_CONTRACTION, "Found bad contractions");
                            leave[cidx[c]] = 2;  // failed
                        }
                        s.cols = hst[c].max_col + 1;
                        s.rows = hst[c].max_row + 1;
                        s.max_node = hst[c].max_kept_col;
                        s.nnz = hst[c].kept;
                        s.span = hnb[Q + cidx[c]];
                        nnz += s.nnz;
                        if (s.nnz == 0) leave[cidx[c]] = 1;  // A.nnz == 0 ends the rounds
                    }
                    Bc.nnz = nnz;
                    // the pack now lives in the new id space of the continuing slots
                    std::vector<Slot> ns;
                    std::vector<char> nleave, nhand;
                    std::vector<i64> old_lab = P.lab;
                    DBuf<int> labels_old = std::move(P.labels);
                    // labels of slots that do not continue are final: copy them out
                    {
                        std::vector<int> fin;
                        for (int q = 0; q < Q; ++q)
                            if (!cont[q] && leave[q] == 1) fin.push_back(q);
                        P.labels = std::move(labels_old);
                        const std::vector<int32_t> hl = gather_labels(ctx, P, fin);
                        labels_old = std::move(P.labels);
                        i64 o = 0;
                        for (int q : fin) {
                            write_labels(B, P.s[q].id, hl.data() + o, P.s[q].n0);
                            set_status(B, P.s[q].id, 0, "");
                            o += P.s[q].n0;
                        }
                    }
                    P.A = std::move(Bc);
                    std::vector<Slot> keep_s;
                    for (int c = 0; c < Qc; ++c) keep_s.push_back(P.s[cidx[c]]);
                    std::vector<char> keep_leave, keep_hand;
                    for (int c = 0; c < Qc; ++c) {
                        keep_leave.push_back(leave[cidx[c]]);
                        keep_hand.push_back(handoff[cidx[c]]);
                    }
                    // rebuild the pack over the continuing slots
                    std::vector<i64> keep_lab_off;
                    for (int c = 0; c < Qc; ++c) keep_lab_off.push_back(old_lab[cidx[c]]);
                    P.s = keep_s;
                    pack_index(ctx, P);
                    // compact the labels of the continuing slots
                    {
                        std::vector<i64> n;
                        for (int c = 0; c < Qc; ++c) n.push_back(P.s[c].n0);
                        P.labels = gather_ranges(ctx, labels_old.get(), keep_lab_off, n);
                    }
                    leave.assign(keep_leave.begin(), keep_leave.end());
                    handoff.assign(keep_hand.begin(), keep_hand.end());
                } else {
                    // nothing contracts: every slot is finished
                    std::vector<int32_t> hl((size_t)P.lab[Q]);
                    if (P.lab[Q] > 0) ctx.d2h(hl.data(), P.labels.get(), P.lab[Q]);
                    for (int q = 0; q < Q; ++q) {
                        write_labels(B, P.s[q].id, hl.data() + P.lab[q], P.s[q].n0);
                        set_status(B, P.s[q].id, 0, "");
                    }
                    P.s.clear();
                    break;
                }
            }
            mark("s:contract", iter, Q);
            // ------------------------------ slots leaving after this round
            Q = P.Q();
            bool any_leave = false;
            for (int q = 0; q < Q; ++q) any_leave = any_leave || leave[q] || handoff[q];
            if (!any_leave) continue;
            std::vector<int> out_slots, hand_slots, stay;
            for (int q = 0; q < Q; ++q) {
                if (leave[q] == 1 || (handoff[q] && leave[q] != 2)) out_slots.push_back(q);
                if (handoff[q] && !leave[q]) hand_slots.push_back(q);
                if (!leave[q] && !handoff[q]) stay.push_back(q);
            }
            const std::vector<int32_t> hl = gather_labels(ctx, P, out_slots);
            // edges of the handed-off slots, gathered on the device in slot order
            std::vector<int32_t> er, ec;
            std::vector<float> ev;
            std::vector<i64> hoff;
            if (!hand_slots.empty()) {
                std::vector<int> isin((size_t)Q, 0);
                i64 mh = 0;
                for (int q : hand_slots) {
                    isin[q] = 1;
                    hoff.push_back(mh);
                    mh += P.s[q].nnz;
                }
                er.resize((size_t)mh);
                ec.resize((size_t)mh);
                ev.resize((size_t)mh);
                if (mh > 0) {
                    DBuf<int> din = upload_vec(ctx, isin);
                    const int* IN = din.get();
                    const int* NS = P.nslot.get();
                    const int* AR = P.A.row.get();
                    const int* AC = P.A.col.get();
                    const float* AW = P.A.val.get();
                    DBuf<int> gr = ctx.buf<int>(mh), gc = ctx.buf<int>(mh);
                    DBuf<float> gw = ctx.buf<float>(mh);
                    int* GR = gr.get();
                    int* GC = gc.get();
                    float* GW = gw.get();
                    DBuf<int> cnt = ctx.buf<int>(1);
                    select_into(
                        ctx, P.A.nnz, [=] __device__(i64 e) { return IN[NS[AR[e]]] != 0; },
                        [=] __device__(i64 e, int p) {
                            GR[p] = AR[e];
                            GC[p] = AC[e];
                            GW[p] = AW[e];
                        },
                        cnt.get());
                    ctx.d2h(er.data(), gr.get(), mh);
                    ctx.d2h(ec.data(), gc.get(), mh);
                    ctx.d2h(ev.data(), gw.get(), mh);
                }
            }
            i64 lo = 0;
            size_t hi_idx = 0;
            for (int q : out_slots) {
                const Slot& s = P.s[q];
                if (leave[q] == 1) {
                    write_labels(B, s.id, hl.data() + lo, s.n0);
                    set_status(B, s.id, 0, "");
                } else {
                    Handoff h;
                    h.id = s.id;
                    h.next_iter = use_mm ? iter + 1 : iter;
                    h.rows = s.rows;
                    h.cols = s.cols;
                    h.max_node = s.max_node;
                    h.lb = s.lb;
                    h.labels.assign(hl.begin() + lo, hl.begin() + lo + s.n0);
                    const i64 e0 = hoff[hi_idx++];
                    h.row.resize((size_t)s.nnz);
                    h.col.resize((size_t)s.nnz);
                    h.val.resize((size_t)s.nnz);
                    for (i64 e = 0; e < s.nnz; ++e) {
                        h.row[e] = er[e0 + e] - P.base[q];
                        h.col[e] = ec[e0 + e] - P.base[q];
                        h.val[e] = ev[e0 + e];
                    }
                    hand_off(std::move(h));
                }
                lo += s.n0;
            }
            if (stay.empty()) {
                P.s.clear();
                break;
            }
            // compact the pack: edges and labels of the staying slots, ids shifted down
            std::vector<int> keep((size_t)Q, 0), shift((size_t)Q, 0);
            int removed = 0;
            for (int q = 0; q < Q; ++q) {
                shift[q] = removed;
                const bool st = std::find(stay.begin(), stay.end(), q) != stay.end();
                keep[q] = st;
                if (!st) removed += P.s[q].span;
            }
            DBuf<int> dkeep = upload_vec(ctx, keep), dshift = upload_vec(ctx, shift);
            const int* KP = dkeep.get();
            const int* SH = dshift.get();
            const int* NS = P.nslot.get();
            DCoo A2;
            i64 m2 = 0;
            for (int q : stay) m2 += P.s[q].nnz;
            A2.nnz = m2;
            A2.row = ctx.buf<int>(std::max<i64>(m2, 1));
            A2.col = ctx.buf<int>(std::max<i64>(m2, 1));
            A2.val = ctx.buf<float>(std::max<i64>(m2, 1));
            {
                const int* AR = P.A.row.get();
                const int* AC = P.A.col.get();
                const float* AW = P.A.val.get();
                int* R2 = A2.row.get();
                int* C2 = A2.col.get();
                float* W2 = A2.val.get();
                DBuf<int> cnt = ctx.buf<int>(1);
                select_into(
                    ctx, P.A.nnz, [=] __device__(i64 e) { return KP[NS[AR[e]]] != 0; },
                    [=] __device__(i64 e, int p) {
                        const int q = NS[AR[e]];
                        R2[p] = AR[e] - SH[q];
                        C2[p] = AC[e] - SH[q];
                        W2[p] = AW[e];
                    },
                    cnt.get());
            }
            std::vector<i64> old_lab = P.lab;
            DBuf<int> labels_old = std::move(P.labels);
            std::vector<Slot> ns;
            for (int q : stay) ns.push_back(P.s[q]);
            P.s = ns;
            P.A = std::move(A2);
            pack_index(ctx, P);
            {
                std::vector<i64> off, n;
                for (size_t c = 0; c < stay.size(); ++c) {
                    off.push_back(old_lab[stay[c]]);
                    n.push_back(P.s[c].n0);
                }
                P.labels = gather_ranges(ctx, labels_old.get(), off, n);
            }
        }
        scan_check(ctx);
        mark("pack empty", -1, 0);
    } catch (...) {
        finish_pool();
        throw;
    }
    finish_pool();
    mark("workers done", -1, 0);
    if (times)
        for (auto& kv : stage_ms) std::fprintf(stderr, "packed total %-14s %9.3f ms\n", kv.first.c_str(), kv.second);
}

}  // namespace rama
