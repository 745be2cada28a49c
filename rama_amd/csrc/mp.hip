// mp.hip — triangle/edge min-marginal message passing
// (src/multicut_message_passing.cu:15-360).
//
// Construction sorts the 3T triangle edges and merges them into the graph's
// sorted edge list: graph edges keep their cost, new chords cost 0, and each
// edge's run of triangle items is its slot list in (triangle, slot) order.
// The edges in a triangle are numbered k in edge order; the iterations
// work in that numbering (their costs, shares and message runs) and leave
// the other edges alone.
//
// One iteration is two kernels:
//   k_mp_triangles  per triangle: gather the shares (cost / #triangles) of
//                   its three edges and add them to the triangle costs
//                   (send_messages_to_triplets), run the six min-marginal
//                   steps, store the three messages at their places in the
//                   edge-major message array and add the triangle's
//                   lower-bound term;
//   k_mp_edges      per edge in a triangle: cost = sum of its run of
//                   messages, read contiguously in (triangle, slot) order
//                   (replaces float atomicAdd), plus its lower-bound term.
// Lower-bound partials are per block in double; the edges in no triangle
// add a constant term.  In the device-driven loop the edge kernel's last
// block sums them and applies the stall test, so the host reads the outcome
// once per loop.
#include <cstdlib>
#include <cstring>

#include "rama_internal.hpp"
#include "tuning.hpp"

namespace rama {

// What an edge in `cnt` triangles passes to each of them
// (send_messages_to_triplets, src/multicut_message_passing.cu:218-229).
__device__ __forceinline__ float edge_share(float c, int cnt) { return c / float(cnt); }

// Whether the iterations of a message-passing loop replay one captured HIP
// graph (every iteration launches the same kernels with the same arguments;
// the stall test on the device turns the rest into no-ops).
// RAMA_MP_GRAPH=0 launches them one by one; the kernel profiler's events
// are not captured, so timing these kernels also launches them directly.
static bool mp_graphs() {
    const char* e = std::getenv("RAMA_MP_GRAPH");
    if (e && std::strcmp(e, "0") == 0) return false;
    return !prof::wants("mp_triangles") && !prof::wants("mp_edges");
}


// out[i] = lower_bound(a[0..n), q[i]) for ascending queries q[0..m).  With
// kRepUpper a query equal to its predecessor takes upper_bound instead (the
// answers stay ascending).  Each block bounds its 256 answers by the answers
// of its first and last query (two full searches); a window of at most
// kLbWin keys is staged in LDS with coalesced loads and searched there, a
// wider one in place.
constexpr int kLbWin = 2048;
template <class K, bool kRepUpper>
__device__ __forceinline__ bool lb_upper(const K* __restrict__ q, i64 i) {
    return kRepUpper && i > 0 && q[i - 1] == q[i];
}
template <class K>
__device__ __forceinline__ i64 bound_dev(const K* a, i64 lo, i64 hi, K x, bool upper) {
    while (lo < hi) {
        const i64 mid = (lo + hi) >> 1;
        if (a[mid] < x || (upper && a[mid] == x)) lo = mid + 1; else hi = mid;
    }
    return lo;
}
// The answers of each block's first and last query, one thread per search
// (all blocks' searches run side by side instead of two per block ahead of
// its own work).
template <class K, bool kRepUpper>
__global__ __launch_bounds__(kBlock) void k_sorted_lb_ends(const K* __restrict__ a, i64 n, const K* __restrict__ q,
                                                           i64 m, i64* __restrict__ ends) {
    const i64 t = (i64)blockIdx.x * kBlock + threadIdx.x;
    const i64 nb = (m + kBlock - 1) / kBlock;
    if (t >= 2 * nb) return;
    const i64 b0 = (t >> 1) * kBlock;
    const i64 qi = (t & 1) ? (b0 + kBlock < m ? b0 + kBlock : m) - 1 : b0;
    ends[t] = bound_dev(a, 0, n, q[qi], lb_upper<K, kRepUpper>(q, qi));
}

template <class K, bool kRepUpper>
__global__ __launch_bounds__(kBlock) void k_sorted_lb(const K* __restrict__ a, i64 n, const K* __restrict__ q, i64 m,
                                                      const i64* __restrict__ ends, int* __restrict__ out) {
    __shared__ K sa[kLbWin];
    const i64 b0 = (i64)blockIdx.x * kBlock;
    const i64 b1 = (b0 + kBlock < m ? b0 + kBlock : m) - 1;
    const i64 r0 = ends[2 * (i64)blockIdx.x], r1 = ends[2 * (i64)blockIdx.x + 1];
    const i64 i = b0 + threadIdx.x;
    if (r1 - r0 <= kLbWin) {  // uniform
        const int w = (int)(r1 - r0);
        for (int k = threadIdx.x; k < w; k += kBlock) sa[k] = a[r0 + k];
        __syncthreads();
        if (i <= b1) {
            const K x = q[i];
            const bool up = lb_upper<K, kRepUpper>(q, i);
            int lo = 0, hi = w;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (sa[mid] < x || (up && sa[mid] == x)) lo = mid + 1; else hi = mid;
            }
            out[i] = (int)(r0 + lo);
        }
    } else if (i <= b1) {
        out[i] = (int)bound_dev(a, r0, r1, q[i], lb_upper<K, kRepUpper>(q, i));
    }
}

template <class K, bool kRepUpper = false>
static void sorted_lb(Ctx& ctx, const K* a, i64 n, const K* q, i64 m, int* out) {
    if (m <= 0) return;
    const unsigned nb = grid_for(m);
    DBuf<i64> ends = ctx.buf<i64>(2 * (i64)nb);
    k_sorted_lb_ends<K, kRepUpper><<<grid_for(2 * (i64)nb), kBlock, 0, ctx.stream>>>(a, n, q, m, ends.get());
    RAMA_LAUNCHED();
    k_sorted_lb<K, kRepUpper><<<nb, kBlock, 0, ctx.stream>>>(a, n, q, m, ends.get(), out);
    RAMA_LAUNCHED();
}

// Builds the message-passing problem.  The graph's edges are already sorted
// (a (row, col) pair may repeat); only the 3T triangle-edge items (key, slot
// 3t+k) are sorted.  Triangle keys absent from the graph are the chords.  In
// the merged edge list a graph entry's position is (#graph entries before it)
// + (#chord keys below it), a chord's (#graph entries below it) + (#chords
// before it); an edge's slots are the sorted triangle items carrying its key,
// which are contiguous and already in (triangle, slot) order (the sort is
// stable).  Every graph entry is kept with its own cost and a repeated pair's
// triangles all go to its first entry (set_union_by_key and
// set_intersection_by_key in src/multicut_message_passing.cu:46-50,131-156).
template <class K>
struct RunHead {  // item q starts a run of equal sorted keys
    const K* kt;
    __device__ __forceinline__ bool operator()(i64 q) const { return q == 0 || kt[q] != kt[q - 1]; }
};

template <class K>
static void mp_build_impl(Ctx& ctx, const DCoo& A, Triangles& tri, MPState& S, int kb) {
    const i64 E = A.nnz, T = tri.n, N3 = 3 * T;
    S.T = T;
    const int* AR = A.row.get();
    const int* AC = A.col.get();
    const float* AV = A.val.get();
    DBuf<K> ka = ctx.buf<K>(E);
    K* KA = ka.get();
    for_each(ctx, E, [=] __device__(i64 e) { KA[e] = ((K)(u32)AR[e] << kb) | (K)(u32)AC[e]; });
    DBuf<K> kt = ctx.buf<K>(N3);
    DBuf<u32> vt = ctx.buf<u32>(N3);
    K* KT = kt.get();
    u32* VT = vt.get();
    const int* T1 = tri.t1.get();
    const int* T2 = tri.t2.get();
    const int* T3 = tri.t3.get();
    // triangle costs start at zero, cleared in the same launch
    S.t12 = ctx.buf<float>(T);
    S.t13 = ctx.buf<float>(T);
    S.t23 = ctx.buf<float>(T);
    float* TC12 = S.t12.get();
    float* TC13 = S.t13.get();
    float* TC23 = S.t23.get();
    for_each(ctx, T, [=] __device__(i64 t) {
        TC12[t] = 0.0f;
        TC13[t] = 0.0f;
        TC23[t] = 0.0f;
        const K a = (u32)T1[t], b = (u32)T2[t], c = (u32)T3[t];
        KT[3 * t + 0] = (a << kb) | b;
        KT[3 * t + 1] = (a << kb) | c;
        KT[3 * t + 2] = (b << kb) | c;
        VT[3 * t + 0] = (u32)(3 * t + 0);
        VT[3 * t + 1] = (u32)(3 * t + 1);
        VT[3 * t + 2] = (u32)(3 * t + 2);
    });
    prof::Site site_("mp_build");
    radix_sort_pairs<K>(ctx, kt, vt, N3, 2 * kb);
    KT = kt.get();  // an odd pass count hands back the scratch buffers
    VT = vt.get();
    // position of every sorted triangle item among the graph keys
    DBuf<int> pa = ctx.buf<int>(N3 > 0 ? N3 : 1);
    int* PA = pa.get();
    sorted_lb<K>(ctx, KA, E, KT, N3, PA);
    // chords: first item of a key run whose key is not a graph edge
    auto is_chord = [=] __device__(i64 q) {
        if (q > 0 && KT[q] == KT[q - 1]) return 0;
        const i64 p = PA[q];
        return (int)(p == E || KA[p] != KT[q]);
    };
    DBuf<int> cpos;
    const i64 Nc = select_positions(ctx, is_chord, N3, cpos);
    DBuf<K> ck = ctx.buf<K>(Nc);
    DBuf<int> chead = ctx.buf<int>(Nc);
    K* CK = ck.get();
    int* CH = chead.get();
    const int* CP = cpos.get();
    if (Nc > 0)
        for_each(ctx, N3, [=] __device__(i64 q) {
            if (is_chord(q)) {
                CK[CP[q]] = KT[q];
                CH[CP[q]] = (int)q;
            }
        });
    // chords below every graph key (none to search when there are none)
    DBuf<int> cbelow;
    if (Nc > 0) {
        cbelow = ctx.buf<int>(E > 0 ? E : 1);
        sorted_lb<K>(ctx, CK, Nc, KA, E, cbelow.get());
    }
    const int* CB = Nc > 0 ? cbelow.get() : nullptr;
    const i64 Em = E + Nc;
    S.E = Em;
    S.ei = ctx.buf<int>(Em);
    S.ej = ctx.buf<int>(Em);
    S.ec = ctx.buf<float>(Em);
    S.inact = ctx.buf<unsigned char>(Em > 0 ? Em : 1);
    S.c12 = ctx.buf<int>(T);
    S.c13 = ctx.buf<int>(T);
    S.c23 = ctx.buf<int>(T);
    S.delta = ctx.buf<float>(N3 > 0 ? N3 : 1);
    S.pos = ctx.buf<u32>(N3 > 0 ? N3 : 1);
    int* EI = S.ei.get();
    int* EJ = S.ej.get();
    float* EC = S.ec.get();
    unsigned char* INACT = S.inact.get();
    const u64 mask = ((u64)1 << kb) - 1;
    // Places in the merged list: a graph edge's index plus the chords below
    // its key; a chord's index among the chords plus the graph keys below
    // it (PA of its head).  Every place starts out in no triangle (the
    // numbering scan below marks those in one), chords excepted.
    for_each(ctx, E, [=] __device__(i64 e) {
        const i64 u = e + (CB ? CB[e] : 0);
        EI[u] = AR[e];
        EJ[u] = AC[e];
        EC[u] = AV[e];
        INACT[u] = 1;
    });
    // chords (cost 0)
    for_each(ctx, Nc, [=] __device__(i64 c) {
        const i64 u = c + PA[CH[c]];
        EI[u] = (int)((u64)CK[c] >> kb);
        EJ[u] = (int)((u64)CK[c] & mask);
        EC[u] = 0.0f;
        INACT[u] = 0;
    });
    // largest col of the merged list: A's (its max_node) or a chord's; all
    // cols are read when A's is not known
    S.dims = ctx.buf<int>(2);
    if (A.max_node >= 0 || E == 0)
        reduce_to<int>(ctx, [=] __device__(i64 c) { return (int)((u64)CK[c] & mask); }, Nc, A.max_node, OpMax(),
                       S.dims.get());
    else
        reduce_to<int>(ctx, [=] __device__(i64 u) { return EJ[u]; }, Em, -1, OpMax(), S.dims.get());
    // The edges in a triangle, ascending, are numbered k = 0 .. Ea-1, one
    // per run of equal sorted items: act[k] is the run head's place in the
    // merged list (as above; a head is never a repeat) and its messages are
    // the places hq[k] .. hq[k+1]-1.  One scan over the items numbers the
    // runs and gives every triangle slot its edge number k (the triangle
    // sweep gathers share[k]) and its place q (where it sends its message).
    S.act = ctx.buf<int>(N3 > 0 ? N3 : 1);
    S.hq = ctx.buf<int>(N3 + 1);
    int* C12 = S.c12.get();
    int* C13 = S.c13.get();
    int* C23 = S.c23.get();
    u32* POS = S.pos.get();
    int* ACT = S.act.get();
    int* HQ = S.hq.get();
    {
        DBuf<int> na = ctx.buf<int>(1);
        int* am = ctx.mirror<int>();
        scan_apply<int>(
            ctx, FlagFn<RunHead<K>>{RunHead<K>{KT}}, N3,
            [=] __device__(i64 q, int before, int head) {
                const int k = before + head - 1;
                if (head) {
                    const int u = (int)(PA[q] + CP[q]);
                    ACT[k] = u;
                    HQ[k] = (int)q;
                    INACT[u] = 0;
                }
                const u32 v = VT[q], t = v / 3, s = v % 3;
                (s == 0 ? C12 : (s == 1 ? C13 : C23))[t] = k;
                POS[v] = (u32)q;  // slot 3t+s sends to place q, inside its edge's run
            },
            na.get(), am);
        int h = 0;
        ctx.read_mirror(na.get(), am, &h);
        S.Ea = h;
    }
    // edge k's cost while the iterations run, and what each triangle
    // absorbs from it (cost / #triangles)
    const i64 Ea = S.Ea;
    S.eca = ctx.buf<float>(Ea > 0 ? Ea : 1);
    S.share = ctx.buf<float>(Ea > 0 ? Ea : 1);
    float* ECA = S.eca.get();
    float* SH = S.share.get();
    for_each(ctx, Ea + 1, [=] __device__(i64 k) {
        if (k == Ea) {
            HQ[Ea] = (int)N3;
            return;
        }
        const int q0 = HQ[k], q1 = k + 1 < Ea ? HQ[k + 1] : (int)N3;
        const float c = EC[ACT[k]];
        ECA[k] = c;
        SH[k] = edge_share(c, q1 - q0);
    });
}

void mp_build(Ctx& ctx, const DCoo& A, Triangles& tri, MPState& S) {
    int maxv = A.max_dim() - 1;
    if (tri.n > 0)
        maxv = std::max(maxv, tri.max_id >= 0 ? tri.max_id : max_id(ctx, tri.t3.get(), tri.t3.get(), tri.n));
    const int kb = bits_for(maxv < 1 ? 1 : maxv);
    if (2 * kb <= 32)
        mp_build_impl<u32>(ctx, A, tri, S, kb);
    else
        mp_build_impl<u64>(ctx, A, tri, S, kb);
    const i64 nb = (std::max(S.T, S.E) + kBlock - 1) / kBlock + 2;
    S.lbparts = ctx.buf<double>(2 * nb + 2);
    S.max_col = S.last_row = -1;  // read back with the final bound
}

// Device-side state of the dual loop (src/dual_solver.cu:17-26): the
// current and previous lower bound and whether iterations still run.
struct MPControl {
    double lb;
    double prev;
    int active;
    int iterations;
    int max_col;   // copied from MPState::dims so one read returns both
    int last_row;
    double linact;  // bound term of the edges in no triangle, added to every iteration's sum
};

// Thread t's share of a fixed-order sum: p[t] + p[t + kBlock] + ... in
// increasing index, with eight loads in flight at a time (the additions keep
// their order, so the result does not depend on the batching).
__device__ __forceinline__ double strided_sum(const double* __restrict__ p, i64 n) {
    double s = 0.0;
    i64 k = threadIdx.x;
    for (; k + 7 * kBlock < n; k += 8 * kBlock) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = p[k + (i64)u * kBlock];
#pragma unroll
        for (int u = 0; u < 8; ++u) s += v[u];
    }
    for (; k < n; k += kBlock) s += p[k];
    return s;
}

// After an iteration on a large grid: sum the lower-bound partials in the
// same fixed order as the fused path and apply the stall test.
__global__ __launch_bounds__(kBlock) void k_mp_finalize(const double* __restrict__ parts, i64 n, MPControl* ctl) {
    __shared__ double ws[kWaves];
    if (!ctl->active) return;
    const double sum = block_reduce(strided_sum(parts, n), OpSum(), ws);
    if (threadIdx.x == 0) {
        const double lb = sum + ctl->linact;
        ctl->prev = ctl->lb;
        ctl->lb = lb;
        ctl->iterations += 1;
        if (lb - ctl->prev < 1e-3) ctl->active = 0;
    }
}

// Triangle lower-bound term (triangle_lb_func, src/multicut_message_passing.cu:186-199).
__device__ __forceinline__ float tri_lb(float a, float b, float d) {
    float lb = 0.0f;
    lb = fminf(lb, a + b);
    lb = fminf(lb, a + d);
    lb = fminf(lb, b + d);
    lb = fminf(lb, a + b + d);
    return lb;
}

// min_marginal (src/multicut_message_passing.cu:274-280)
__device__ __forceinline__ float min_marginal(float x, float y, float z) {
    const float mm1 = fminf(x + y + z, fminf(x + y, x + z));
    const float mm0 = (float)fmin(0.0, (double)(y + z));
    return mm1 - mm0;
}

// send_messages_to_triplets (:218-229) for one triangle: each edge passes
// cost / #triangles (edge_share); the edge kernel stores that share next to
// the cost, so a triangle gathers one value per edge.

// One triangle's update (send_messages_to_triplets + the six min-marginal
// steps); returns its lower-bound term.
template <bool kSweep>
__device__ __forceinline__ double mp_triangle(i64 t, const int* __restrict__ c12, const int* __restrict__ c13,
                                              const int* __restrict__ c23, const float* __restrict__ share,
                                              float* __restrict__ t12, float* __restrict__ t13,
                                              float* __restrict__ t23, const u32* __restrict__ pos,
                                              float* __restrict__ delta) {
    // the edge ids and message places are streamed once per sweep:
    // non-temporal loads (measured faster)
    float a = t12[t], b = t13[t], d = t23[t];
    a += share[__builtin_nontemporal_load(&c12[t])];
    b += share[__builtin_nontemporal_load(&c13[t])];
    d += share[__builtin_nontemporal_load(&c23[t])];
    if (kSweep) {
        // decrease_triangle_costs_func (:282-332): steps with weights
        // 1/3, 1/2, 1, 1/2, 1, 1 over slots 12, 13, 23, 12, 13, 12.
        float e12 = 0.0f, e13 = 0.0f, e23 = 0.0f;
        {
            const float mm = min_marginal(a, b, d);
            a = (float)((double)a - 1.0 / 3.0 * (double)mm);
            e12 = (float)((double)e12 + 1.0 / 3.0 * (double)mm);
        }
        {
            const float mm = min_marginal(b, a, d);
            b = (float)((double)b - 1.0 / 2.0 * (double)mm);
            e13 = (float)((double)e13 + 1.0 / 2.0 * (double)mm);
        }
        {
            const float mm = min_marginal(d, a, b);
            d -= mm;
            e23 += mm;
        }
        {
            const float mm = min_marginal(a, b, d);
            a = (float)((double)a - 1.0 / 2.0 * (double)mm);
            e12 = (float)((double)e12 + 1.0 / 2.0 * (double)mm);
        }
        {
            const float mm = min_marginal(b, a, d);
            b -= mm;
            e13 += mm;
        }
        {
            const float mm = min_marginal(a, b, d);
            a -= mm;
            e12 += mm;
        }
        delta[__builtin_nontemporal_load(&pos[3 * t + 0])] = e12;
        delta[__builtin_nontemporal_load(&pos[3 * t + 1])] = e13;
        delta[__builtin_nontemporal_load(&pos[3 * t + 2])] = e23;
    }
    t12[t] = a;
    t13[t] = b;
    t23[t] = d;
    return (double)tri_lb(a, b, d);
}

// The edge update (send_messages_to_edges' sums, decrease_edge_costs_func,
// :231-239) for the edges in a triangle (act, ascending; the others keep
// their cost, and their constant lower-bound term is added by the
// finaliser): with kSweep, an edge's cost becomes the sum of its slot
// messages; without, it is zeroed.  Its share cost / #triangles is stored
// beside it.
//
// Active edges k and k + kBlock (those below end) of one thread: offsets,
// costs and the first four messages of both are loaded before either is
// summed; each edge's sum keeps its slot order.  Returns the two
// lower-bound terms added (k's first).
template <bool kSweep>
__device__ __forceinline__ double mp_edge_pair(i64 k, i64 end, const int* __restrict__ hq,
                                               const float* __restrict__ delta, float* __restrict__ eca,
                                               float* __restrict__ share) {
    constexpr int kPre = 4;
    i64 kk[2];
    bool has[2];
    int s0[2], s1[2];
    float c[2];
    float x[2][kPre];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        kk[h] = k + (i64)h * kBlock;
        has[h] = kk[h] < end;
        s0[h] = has[h] ? hq[kk[h]] : 0;
        s1[h] = has[h] ? hq[kk[h] + 1] : 0;
        c[h] = has[h] ? eca[kk[h]] : 0.0f;
    }
    if (kSweep) {
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int u = 0; u < kPre; ++u) x[h][u] = s0[h] + u < s1[h] ? delta[s0[h] + u] : 0.0f;
    }
    double lb = 0.0;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        if (!has[h]) continue;
        float cc = c[h];
        if (s1[h] > s0[h]) {
            if (kSweep) {
                // in slot order
                double sum = 0.0;
#pragma unroll
                for (int u = 0; u < kPre; ++u)
                    if (s0[h] + u < s1[h]) sum += (double)x[h][u];
                int k = s0[h] + kPre;
                for (; k + 4 <= s1[h]; k += 4) {
                    const float y0 = delta[k], y1 = delta[k + 1], y2 = delta[k + 2], y3 = delta[k + 3];
                    sum += (double)y0;
                    sum += (double)y1;
                    sum += (double)y2;
                    sum += (double)y3;
                }
                for (; k < s1[h]; ++k) sum += (double)delta[k];
                cc = (float)sum;
            } else {
                cc = 0.0f;
            }
            eca[kk[h]] = cc;
            share[kk[h]] = edge_share(cc, s1[h] - s0[h]);
        }
        lb += cc < 0.0f ? (double)cc : 0.0;
    }
    return lb;
}

template <bool kSweep>
__global__ __launch_bounds__(kBlock) void k_mp_triangles(i64 T, const int* __restrict__ c12, const int* __restrict__ c13,
                                                         const int* __restrict__ c23, const float* __restrict__ share,
                                                         float* __restrict__ t12, float* __restrict__ t13,
                                                         float* __restrict__ t23, const u32* __restrict__ pos,
                                                         float* __restrict__ delta, double* __restrict__ lbpart,
                                                         const MPControl* __restrict__ ctl) {
    __shared__ double ws[kWaves];
    if (ctl && !ctl->active) return;  // uniform: the loop has stopped
    const unsigned blk = xcd_block();
    const i64 t = (i64)blk * kBlock + threadIdx.x;
    const double lb = t < T ? mp_triangle<kSweep>(t, c12, c13, c23, share, t12, t13, t23, pos, delta) : 0.0;
    const double r = block_reduce(lb, OpSum(), ws);
    if (threadIdx.x == 0) lbpart[blk] = r;
}

// Edge side (mp_edge per edge), adding the edge lower-bound terms per block.
//
// With `parts` set (the device-driven loop on small grids), the last block to
// finish (atomic ticket, reset afterwards) also sums all nparts lower-bound
// partials of the iteration in a fixed order and applies the stall test
// `lb - prev < 1e-3` (src/dual_solver.cu:17-26) to ctl, as k_mp_finalize does.
// Each block's ticket carries a release, whose cost grows with the grid, so
// large grids finalise in a separate one-block launch instead.
template <bool kSweep>
__global__ __launch_bounds__(kBlock) void k_mp_edges(i64 Ea, const int* __restrict__ hq,
                                                     const float* __restrict__ delta, float* __restrict__ eca,
                                                     float* __restrict__ share, double* __restrict__ lbpart,
                                                     MPControl* ctl, const double* parts, i64 nparts,
                                                     unsigned* ticket) {
    __shared__ double ws[kWaves];
    __shared__ bool last;
    if (ctl && !ctl->active) return;
    const unsigned blk = xcd_block();
    const i64 k = (i64)blk * kMpEdgesPerBlock + threadIdx.x;
    const double lb = mp_edge_pair<kSweep>(k, Ea, hq, delta, eca, share);
    const double r = block_reduce(lb, OpSum(), ws);
    if (threadIdx.x == 0) lbpart[blk] = r;
    if (!parts) return;
    if (threadIdx.x == 0) {
        const unsigned t = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        last = t == gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    __threadfence();  // acquire: the other blocks' partials are visible to plain loads
    const double sum = block_reduce(strided_sum(parts, nparts), OpSum(), ws);
    if (threadIdx.x == 0) {
        const double tot = sum + ctl->linact;
        ctl->prev = ctl->lb;
        ctl->lb = tot;
        ctl->iterations += 1;
        if (tot - ctl->prev < 1e-3) ctl->active = 0;
        *ticket = 0;
    }
}

static double sum_parts(Ctx& ctx, MPState& S, i64 n) {
    double* P = S.lbparts.get();
    const i64 slot = S.lbparts.size() - 1;
    reduce_to<double>(ctx, [=] __device__(i64 k) { return P[k]; }, n, 0.0, OpSum(), P + slot);
    return ctx.read1(P + slot);
}

static void launch_round(Ctx& ctx, MPState& S, bool sweep, MPControl* ctl = nullptr) {
    const i64 T = S.T, Ea = S.Ea;
    const unsigned gt = grid_for(T), ge = mp_edge_blocks(Ea);
    double* P = S.lbparts.get();
    {
        // per triangle: 3 edge ids, 3 share gathers, 3 costs in and out,
        // 3 message places, 3 messages out = 72 B
        prof::Scope ps(ctx, sweep ? "mp_triangles" : "mp_triangles_absorb", (sweep ? 72.0 : 48.0) * T);
        if (sweep)
            k_mp_triangles<true><<<gt, kBlock, 0, ctx.stream>>>(T, S.c12.get(), S.c13.get(), S.c23.get(),
                                                                 S.share.get(), S.t12.get(), S.t13.get(),
                                                                 S.t23.get(), S.pos.get(), S.delta.get(), P, ctl);
        else
            k_mp_triangles<false><<<gt, kBlock, 0, ctx.stream>>>(T, S.c12.get(), S.c13.get(), S.c23.get(),
                                                                  S.share.get(), S.t12.get(), S.t13.get(),
                                                                  S.t23.get(), S.pos.get(), S.delta.get(), P, ctl);
        RAMA_LAUNCHED();
    }
    // the device-driven loop (ctl set) finalises each iteration in the edge
    // kernel on small grids, in a separate launch on large ones
    const bool fused = ctl && ge <= (unsigned)tune::get(tune::MpFusedMaxBlocks, Ea);
    const double* parts = fused ? P : nullptr;
    const i64 nparts = (i64)gt + (i64)ge;
    unsigned* ticket = fused ? ctx.reduce_ticket() : nullptr;
    {
        // per edge in a triangle: its message range, cost in/out, share
        // out; per slot: its message
        prof::Scope ps(ctx, sweep ? "mp_edges" : "mp_edges_zero", 16.0 * Ea + (sweep ? 12.0 * T : 0.0));
        if (sweep)
            k_mp_edges<true><<<ge, kBlock, 0, ctx.stream>>>(Ea, S.hq.get(), S.delta.get(), S.eca.get(),
                                                             S.share.get(), P + gt, ctl, parts, nparts, ticket);
        else
            k_mp_edges<false><<<ge, kBlock, 0, ctx.stream>>>(Ea, S.hq.get(), S.delta.get(), S.eca.get(),
                                                              S.share.get(), P + gt, ctl, parts, nparts, ticket);
        RAMA_LAUNCHED();
    }
    if (ctl && !fused) {
        k_mp_finalize<<<1, kBlock, 0, ctx.stream>>>(P, nparts, ctl);
        RAMA_LAUNCHED();
    }
}

// ------------------------------------------------------------------
// Several instances in one message-passing state (packed batch).  Instance
// q owns triangles [tbeg[q], tbeg[q+1]), edges [ebeg[q], ebeg[q+1]) and,
// of those, the edges in a triangle act[abeg[q] .. abeg[q+1]); its blocks
// are [tblk[q], tblk[q+1]) of the triangle launch and [eblk[q], eblk[q+1])
// of the edge launch, each covering consecutive elements from the
// instance's start exactly as its own launches would.  Its lower-bound
// partials sit at parts + pbeg[q] (triangle blocks, then edge blocks) and
// are summed in the same order as k_mp_finalize sums them, plus its own
// bound term of the edges in no triangle, so every instance's bound and
// stall test equal those of a solve on its own.
template <bool kSweep>
__global__ __launch_bounds__(kBlock) void k_mp_triangles_packed(MPPack pk, const int* __restrict__ c12,
                                                                const int* __restrict__ c13,
                                                                const int* __restrict__ c23,
                                                                const float* __restrict__ share,
                                                                float* __restrict__ t12, float* __restrict__ t13,
                                                                float* __restrict__ t23,
                                                                const u32* __restrict__ pos,
                                                                float* __restrict__ delta,
                                                                double* __restrict__ parts,
                                                                const MPControl* __restrict__ ctl) {
    __shared__ double ws[kWaves];
    const int q = block_owner(pk.tblk, pk.Q, (int)blockIdx.x);
    if (!ctl[q].active) return;  // uniform per block
    const int lb = (int)blockIdx.x - pk.tblk[q];
    const i64 t = pk.tbeg[q] + (i64)lb * kBlock + threadIdx.x;
    const double v =
        t < pk.tbeg[q + 1] ? mp_triangle<kSweep>(t, c12, c13, c23, share, t12, t13, t23, pos, delta) : 0.0;
    const double r = block_reduce(v, OpSum(), ws);
    if (threadIdx.x == 0) parts[pk.pbeg[q] + lb] = r;
}

template <bool kSweep>
__global__ __launch_bounds__(kBlock) void k_mp_edges_packed(MPPack pk, const int* __restrict__ hq,
                                                            const float* __restrict__ delta, float* __restrict__ eca,
                                                            float* __restrict__ share, double* __restrict__ parts,
                                                            const MPControl* __restrict__ ctl) {
    __shared__ double ws[kWaves];
    const int q = block_owner(pk.eblk, pk.Q, (int)blockIdx.x);
    if (!ctl[q].active) return;
    const int lb = (int)blockIdx.x - pk.eblk[q];
    const i64 k = pk.abeg[q] + (i64)lb * kMpEdgesPerBlock + threadIdx.x;
    const double v = mp_edge_pair<kSweep>(k, pk.abeg[q + 1], hq, delta, eca, share);
    const double r = block_reduce(v, OpSum(), ws);
    if (threadIdx.x == 0) parts[pk.pbeg[q] + (pk.tblk[q + 1] - pk.tblk[q]) + lb] = r;
}

// One block per instance: k_mp_finalize on the instance's partials.
__global__ __launch_bounds__(kBlock) void k_mp_finalize_packed(MPPack pk, const double* __restrict__ parts,
                                                               MPControl* ctl) {
    __shared__ double ws[kWaves];
    const int q = blockIdx.x;
    MPControl* c = ctl + q;
    if (!c->active) return;
    const i64 n = (i64)(pk.tblk[q + 1] - pk.tblk[q]) + (pk.eblk[q + 1] - pk.eblk[q]);
    const double sum = block_reduce(strided_sum(parts + pk.pbeg[q], n), OpSum(), ws);
    if (threadIdx.x == 0) {
        const double lb = sum + c->linact;
        c->prev = c->lb;
        c->lb = lb;
        c->iterations += 1;
        if (lb - c->prev < 1e-3) c->active = 0;
    }
}

// Per instance: initial bound = edge part + triangle part (as mp_run's
// first two reductions), the constant part of the edges in no triangle,
// loop state cleared; `on` marks the instances whose loop runs.
__global__ void k_mp_init_packed(int Q, const double* __restrict__ p_edges, const double* __restrict__ p_tris,
                                 const double* __restrict__ p_rest, const int* __restrict__ on, MPControl* ctl) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= Q) return;
    ctl[q].linact = p_rest[q];
    ctl[q].lb = p_edges[q] + p_tris[q];
    ctl[q].prev = 0.0;
    ctl[q].active = on[q];
    ctl[q].iterations = 0;
    ctl[q].max_col = -1;
    ctl[q].last_row = -1;
}

void mp_run_packed(Ctx& ctx, MPState& S, const MPPack& pk, int total_tblocks, int total_eblocks, int steps,
                   const i64* seg_tbeg_host, const i64* seg_ebeg_host, const int* d_on,
                   double* lb_out) {
    const int Q = pk.Q;
    // initial bound per instance, reduced as reduce_to reduces each part
    // alone; the edges in no triangle as mp_run's pair reduction sums them
    DBuf<double> pe = ctx.buf<double>(3 * (i64)Q);
    const float* EC = S.ec.get();
    const unsigned char* NI = S.inact.get();
    const float* A = S.t12.get();
    const float* B = S.t13.get();
    const float* D = S.t23.get();
    seg_reduce<double>(ctx, [=] __device__(i64 e) { const float c = EC[e]; return c < 0.0f ? (double)c : 0.0; },
                       seg_ebeg_host, pk.ebeg, Q, 0.0, OpSum(), pe.get(), 1);
    seg_reduce<double>(ctx, [=] __device__(i64 t) { return (double)tri_lb(A[t], B[t], D[t]); }, seg_tbeg_host,
                       pk.tbeg, Q, 0.0, OpSum(), pe.get() + Q, 1);
    seg_reduce<double>(
        ctx, [=] __device__(i64 e) { const float c = EC[e]; return NI[e] && c < 0.0f ? (double)c : 0.0; },
        seg_ebeg_host, pk.ebeg, Q, 0.0, OpSum(), pe.get() + 2 * (i64)Q, 1);
    DBuf<MPControl> ctl = ctx.buf<MPControl>(Q);
    k_mp_init_packed<<<grid_for(Q), kBlock, 0, ctx.stream>>>(Q, pe.get(), pe.get() + Q, pe.get() + 2 * (i64)Q, d_on,
                                                             ctl.get());
    RAMA_LAUNCHED();
    const i64 nparts = (i64)total_tblocks + total_eblocks;
    DBuf<double> parts = ctx.buf<double>(nparts);
    ctx.repeat(steps, mp_graphs(), [&] {
        k_mp_triangles_packed<true><<<(unsigned)total_tblocks, kBlock, 0, ctx.stream>>>(
            pk, S.c12.get(), S.c13.get(), S.c23.get(), S.share.get(), S.t12.get(), S.t13.get(), S.t23.get(),
            S.pos.get(), S.delta.get(), parts.get(), ctl.get());
        RAMA_LAUNCHED();
        k_mp_edges_packed<true><<<(unsigned)total_eblocks, kBlock, 0, ctx.stream>>>(
            pk, S.hq.get(), S.delta.get(), S.eca.get(), S.share.get(), parts.get(), ctl.get());
        RAMA_LAUNCHED();
        k_mp_finalize_packed<<<(unsigned)Q, kBlock, 0, ctx.stream>>>(pk, parts.get(), ctl.get());
        RAMA_LAUNCHED();
    });
    mp_store_costs(ctx, S);
    std::vector<MPControl> h((size_t)Q);
    ctx.d2h(h.data(), ctl.get(), Q);
    for (int q = 0; q < Q; ++q) lb_out[q] = h[q].lb;
}

__global__ void k_active_ranges(const int* __restrict__ act, i64 n, const i64* __restrict__ ebeg, int Q,
                                i64* __restrict__ abeg) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q > Q) return;
    const i64 x = ebeg[q];
    i64 lo = 0, hi = n;
    while (lo < hi) {
        const i64 mid = (lo + hi) >> 1;
        if ((i64)act[mid] < x) lo = mid + 1; else hi = mid;
    }
    abeg[q] = lo;
}

void mp_active_ranges(Ctx& ctx, const MPState& S, const i64* ebeg, int Q, i64* abeg) {
    k_active_ranges<<<grid_for((i64)Q + 1), kBlock, 0, ctx.stream>>>(S.act.get(), S.Ea, ebeg, Q, abeg);
    RAMA_LAUNCHED();
}

void mp_store_costs(Ctx& ctx, MPState& S) {
    float* EC = S.ec.get();
    const float* ECA = S.eca.get();
    const int* ACT = S.act.get();
    for_each(ctx, S.Ea, [=] __device__(i64 k) { EC[ACT[k]] = ECA[k]; });
}

// lower_bound() (:201-216) of the current state.
double mp_lower_bound(Ctx& ctx, MPState& S) {
    mp_store_costs(ctx, S);
    const float* EC = S.ec.get();
    const float* A = S.t12.get();
    const float* B = S.t13.get();
    const float* D = S.t23.get();
    const i64 E = S.E;
    double* P = S.lbparts.get();
    reduce_to<double>(ctx, [=] __device__(i64 e) { const float c = EC[e]; return c < 0.0f ? (double)c : 0.0; }, E, 0.0,
                      OpSum(), P);
    reduce_to<double>(ctx, [=] __device__(i64 t) { return (double)tri_lb(A[t], B[t], D[t]); }, S.T, 0.0, OpSum(),
                      P + 1);
    double h[2];
    ctx.read(P, h, 2);
    return h[0] + h[1];
}

void mp_send_to_triplets(Ctx& ctx, MPState& S) { launch_round(ctx, S, false); }

// The whole iteration loop of dual_update_cycle_length (src/dual_solver.cu:17-29)
// queued without host round trips: the stall test runs on the device and
// turns the remaining iterations into no-ops.  Returns the final lower bound.
double mp_run(Ctx& ctx, MPState& S, int steps) {
    DBuf<MPControl> ctl = ctx.buf<MPControl>(1);
    double* P = S.lbparts.get();
    const float* EC = S.ec.get();
    const float* A = S.t12.get();
    const float* B = S.t13.get();
    const float* D = S.t23.get();
    const i64 E = S.E;
    const unsigned char* NI = S.inact.get();
    // the edges' bound term, and that of the edges in no triangle (constant
    // over the iterations), each summed as reduce_to sums it
    reduce_pair_to(
        ctx,
        [=] __device__(i64 e, double& a, double& b) {
            const float c = EC[e];
            a = c < 0.0f ? (double)c : 0.0;
            b = NI[e] ? a : 0.0;
        },
        E, P, P + 2);
    reduce_to<double>(ctx, [=] __device__(i64 t) { return (double)tri_lb(A[t], B[t], D[t]); }, S.T, 0.0, OpSum(),
                      P + 1);
    MPControl* C = ctl.get();
    const int* DM = S.E > 0 ? S.dims.get() : nullptr;
    const int* LR = S.E > 0 ? S.ei.get() + (S.E - 1) : nullptr;
    for_each(ctx, 1, [=] __device__(i64) {
        C->lb = P[0] + P[1];
        C->prev = 0.0;
        C->active = 1;
        C->iterations = 0;
        C->max_col = DM ? DM[0] : -1;
        C->last_row = LR ? *LR : -1;
        C->linact = P[2];
    });
    ctx.reduce_ticket();  // created (and cleared) before any capture
    ctx.repeat(steps, mp_graphs(), [&] { launch_round(ctx, S, true, C); });
    MPControl h;
    ctx.read(C, &h, 1);
    S.max_col = h.max_col;
    S.last_row = h.last_row;
    return h.lb;
}

void mp_iteration(Ctx& ctx, MPState& S) { launch_round(ctx, S, true); }

// Lower bound right after mp_iteration / mp_send_to_triplets, from the
// fused per-block partials.
double mp_fused_lower_bound(Ctx& ctx, MPState& S) {
    const i64 n = (i64)grid_for(S.T) + (i64)mp_edge_blocks(S.Ea);
    const double sum = sum_parts(ctx, S, n);
    // plus the bound term of the edges in no triangle
    const float* EC = S.ec.get();
    const unsigned char* NI = S.inact.get();
    const double rest = reduce<double>(
        ctx, [=] __device__(i64 e) { const float c = EC[e]; return NI[e] && c < 0.0f ? (double)c : 0.0; }, S.E, 0.0,
        OpSum());
    return sum + rest;
}

// reparametrized_edge_costs (:356-360) as a sorted dCOO.
DCoo mp_reparametrized(Ctx& ctx, MPState& S) {
    mp_store_costs(ctx, S);
    if (S.E > 0 && S.max_col >= 0) {
        DCoo A;
        A.row = std::move(S.ei);
        A.col = std::move(S.ej);
        A.val = std::move(S.ec);
        A.nnz = S.E;
        A.cols = S.max_col + 1;
        A.rows = S.last_row + 1;
        A.max_node = S.max_col;
        return A;
    }
    return dcoo_from_sorted(ctx, std::move(S.ei), std::move(S.ej), std::move(S.ec), S.E);
}

}  // namespace rama
