// mp.hip — triangle/edge min-marginal message passing
// (src/multicut_message_passing.cu:15-360).
//
// Construction merges the 3T triangle edges with the graph's edges in one
// stable radix sort (graph edges first within a key), so each distinct edge
// is a run of the sorted sequence: the run head gives the edge id, an
// included graph edge gives its cost (new chords cost 0), and the run's
// triangle items are the edge's slot list in (triangle, slot) order.
//
// One iteration is two kernels:
//   k_mp_triangles  per triangle: gather the three edge costs / counts, add
//                   c/count to the triangle costs (send_messages_to_triplets),
//                   run the six min-marginal steps, store the three messages
//                   per slot and the triangle's lower-bound term;
//   k_mp_edges      per edge: cost = sum of its slot messages in fixed order
//                   (replaces float atomicAdd), plus the edge's
//                   lower-bound term.
// Lower-bound partials are per block in double; the host reads one scalar
// per iteration for the stall test.
#include "rama_internal.hpp"

namespace rama {

constexpr u32 kGraphItem = 0x80000000u;

void mp_build(Ctx& ctx, const DCoo& A, Triangles& tri, MPState& S) {
    const i64 E = A.nnz, T = tri.n;
    S.T = T;
    const i64 N = E + 3 * T;
    int maxv = A.max_dim() - 1;
    if (T > 0) maxv = std::max(maxv, max_id(ctx, tri.t3.get(), tri.t3.get(), T));
    const int kb = bits_for(maxv < 1 ? 1 : maxv);
    DBuf<u64> keys = ctx.buf<u64>(N);
    DBuf<u32> vals = ctx.buf<u32>(N);
    u64* K = keys.get();
    u32* V = vals.get();
    {
        const int* R = A.row.get();
        const int* C = A.col.get();
        for_each(ctx, E, [=] __device__(i64 e) {
            K[e] = ((u64)(u32)R[e] << kb) | (u32)C[e];
            V[e] = kGraphItem | (u32)e;
        });
        const int* T1 = tri.t1.get();
        const int* T2 = tri.t2.get();
        const int* T3 = tri.t3.get();
        for_each(ctx, T, [=] __device__(i64 t) {
            const u64 a = (u32)T1[t], b = (u32)T2[t], c = (u32)T3[t];
            K[E + 3 * t + 0] = (a << kb) | b;
            K[E + 3 * t + 1] = (a << kb) | c;
            K[E + 3 * t + 2] = (b << kb) | c;
            V[E + 3 * t + 0] = (u32)(3 * t + 0);
            V[E + 3 * t + 1] = (u32)(3 * t + 1);
            V[E + 3 * t + 2] = (u32)(3 * t + 2);
        });
    }
    if (T > 0) radix_sort_pairs(ctx, keys, vals, N, 2 * kb);  // graph part is already sorted when T == 0
    DBuf<int> uid;
    const i64 Em = select_positions(
        ctx, [=] __device__(i64 q) { return (int)(q == 0 || K[q] != K[q - 1]); }, N, uid);
    DBuf<int> tpos;
    select_positions(ctx, [=] __device__(i64 q) { return (int)((V[q] & kGraphItem) == 0); }, N, tpos);
    S.E = Em;
    S.ei = ctx.buf<int>(Em);
    S.ej = ctx.buf<int>(Em);
    S.ec = ctx.buf<float>(Em);
    S.ecnt = ctx.buf<int>(Em);
    S.soff = ctx.buf<int>(Em + 1);
    S.slots = ctx.buf<int>(3 * T > 0 ? 3 * T : 1);
    S.c12 = ctx.buf<int>(T);
    S.c13 = ctx.buf<int>(T);
    S.c23 = ctx.buf<int>(T);
    S.t12 = ctx.buf<float>(T);
    S.t13 = ctx.buf<float>(T);
    S.t23 = ctx.buf<float>(T);
    S.delta = ctx.buf<float>(3 * T > 0 ? 3 * T : 1);
    ctx.zero(S.t12.get(), T);
    ctx.zero(S.t13.get(), T);
    ctx.zero(S.t23.get(), T);
    ctx.zero(S.ec.get(), Em);
    ctx.zero(S.ecnt.get(), Em);
    int* EI = S.ei.get();
    int* EJ = S.ej.get();
    float* EC = S.ec.get();
    int* ECNT = S.ecnt.get();
    int* SOFF = S.soff.get();
    int* SL = S.slots.get();
    int* CR[3] = {S.c12.get(), S.c13.get(), S.c23.get()};
    int* C12 = CR[0];
    int* C13 = CR[1];
    int* C23 = CR[2];
    const int* U = uid.get();
    const int* TP = tpos.get();
    const float* AV = A.val.get();
    const u64 mask = ((u64)1 << kb) - 1;
    for_each(ctx, N, [=] __device__(i64 q) {
        const u32 v = V[q];
        const bool head = q == 0 || K[q] != K[q - 1];
        const int u = U[q] - (head ? 0 : 1);  // uid holds the exclusive count of heads
        if (head) {
            EI[u] = (int)(K[q] >> kb);
            EJ[u] = (int)(K[q] & mask);
            SOFF[u] = TP[q];
        }
        if (v & kGraphItem) {
            EC[u] = AV[v & ~kGraphItem];
        } else {
            SL[TP[q]] = (int)v;
            const u32 t = v / 3, k = v % 3;
            (k == 0 ? C12 : (k == 1 ? C13 : C23))[t] = u;
            atomicAdd(&ECNT[u], 1);
        }
    });
    const int* TPe = tpos.get();
    int* SO = S.soff.get();
    for_each(ctx, 1, [=] __device__(i64) { SO[Em] = TPe[N]; });
    const i64 nb = (std::max(T, Em) + kBlock - 1) / kBlock + 2;
    S.lbparts = ctx.buf<double>(2 * nb + 2);
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

// send_messages_to_triplets (:218-229) for one triangle.
__device__ __forceinline__ void absorb(const float* __restrict__ ec, const int* __restrict__ ecnt, int e, float& t) {
    t += ec[e] / float(ecnt[e]);
}

template <bool kSweep>
__global__ __launch_bounds__(kBlock) void k_mp_triangles(i64 T, const int* __restrict__ c12, const int* __restrict__ c13,
                                                         const int* __restrict__ c23, const float* __restrict__ ec,
                                                         const int* __restrict__ ecnt, float* __restrict__ t12,
                                                         float* __restrict__ t13, float* __restrict__ t23,
                                                         float* __restrict__ delta, double* __restrict__ lbpart) {
    __shared__ double ws[kWaves];
    const i64 t = (i64)blockIdx.x * kBlock + threadIdx.x;
    double lb = 0.0;
    if (t < T) {
        float a = t12[t], b = t13[t], d = t23[t];
        absorb(ec, ecnt, c12[t], a);
        absorb(ec, ecnt, c13[t], b);
        absorb(ec, ecnt, c23[t], d);
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
            delta[3 * t + 0] = e12;
            delta[3 * t + 1] = e13;
            delta[3 * t + 2] = e23;
        }
        t12[t] = a;
        t13[t] = b;
        t23[t] = d;
        lb = (double)tri_lb(a, b, d);
    }
    const double r = block_reduce(lb, OpSum(), ws);
    if (threadIdx.x == 0) lbpart[blockIdx.x] = r;
}

// Edge side: with kSweep, cost = sum of the slot messages (edges in no
// triangle keep their cost); without, edges in a triangle are zeroed
// (decrease_edge_costs_func, :231-239).  Adds the edge lower-bound term.
template <bool kSweep>
__global__ __launch_bounds__(kBlock) void k_mp_edges(i64 E, const int* __restrict__ ecnt, const int* __restrict__ soff,
                                                     const int* __restrict__ slots, const float* __restrict__ delta,
                                                     float* __restrict__ ec, double* __restrict__ lbpart) {
    __shared__ double ws[kWaves];
    const i64 e = (i64)blockIdx.x * kBlock + threadIdx.x;
    double lb = 0.0;
    if (e < E) {
        float c = ec[e];
        if (ecnt[e] > 0) {
            if (kSweep) {
                double s = 0.0;
                for (int k = soff[e]; k < soff[e + 1]; ++k) s += (double)delta[slots[k]];
                c = (float)s;
            } else {
                c = 0.0f;
            }
            ec[e] = c;
        }
        lb = c < 0.0f ? (double)c : 0.0;
    }
    const double r = block_reduce(lb, OpSum(), ws);
    if (threadIdx.x == 0) lbpart[blockIdx.x] = r;
}

static double sum_parts(Ctx& ctx, MPState& S, i64 n) {
    double* P = S.lbparts.get();
    const i64 slot = S.lbparts.size() - 1;
    reduce_to<double>(ctx, [=] __device__(i64 k) { return P[k]; }, n, 0.0, OpSum(), P + slot);
    return ctx.read1(P + slot);
}

static void launch_round(Ctx& ctx, MPState& S, bool sweep) {
    const i64 T = S.T, E = S.E;
    const unsigned gt = grid_for(T), ge = grid_for(E);
    double* P = S.lbparts.get();
    // per triangle: 3 edge ids, 3 x (cost, count) gathers, 3 costs in and
    // out, 3 messages out = 72 B
    prof::Scope ps1(ctx, sweep ? "mp_triangles" : "mp_triangles_absorb", 72.0 * T);
    if (sweep)
        k_mp_triangles<true><<<gt, kBlock, 0, ctx.stream>>>(T, S.c12.get(), S.c13.get(), S.c23.get(), S.ec.get(),
                                                             S.ecnt.get(), S.t12.get(), S.t13.get(), S.t23.get(),
                                                             S.delta.get(), P);
    else
        k_mp_triangles<false><<<gt, kBlock, 0, ctx.stream>>>(T, S.c12.get(), S.c13.get(), S.c23.get(), S.ec.get(),
                                                              S.ecnt.get(), S.t12.get(), S.t13.get(), S.t23.get(),
                                                              S.delta.get(), P);
    RAMA_LAUNCHED();
    // per edge: count, offsets, cost in/out; per slot: slot id + message
    prof::Scope ps2(ctx, sweep ? "mp_edges" : "mp_edges_zero", 16.0 * E + 24.0 * T);
    if (sweep)
        k_mp_edges<true><<<ge, kBlock, 0, ctx.stream>>>(E, S.ecnt.get(), S.soff.get(), S.slots.get(), S.delta.get(),
                                                         S.ec.get(), P + gt);
    else
        k_mp_edges<false><<<ge, kBlock, 0, ctx.stream>>>(E, S.ecnt.get(), S.soff.get(), S.slots.get(), S.delta.get(),
                                                          S.ec.get(), P + gt);
    RAMA_LAUNCHED();
}

// lower_bound() (:201-216) of the current state.
double mp_lower_bound(Ctx& ctx, MPState& S) {
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

void mp_iteration(Ctx& ctx, MPState& S) { launch_round(ctx, S, true); }

// Lower bound right after mp_iteration / mp_send_to_triplets, from the
// fused per-block partials.
double mp_fused_lower_bound(Ctx& ctx, MPState& S) {
    const i64 n = (i64)grid_for(S.T) + (i64)grid_for(S.E);
    return sum_parts(ctx, S, n);
}

// reparametrized_edge_costs (:356-360) as a sorted dCOO.
DCoo mp_reparametrized(Ctx& ctx, MPState& S) {
    return dcoo_from_sorted(ctx, std::move(S.ei), std::move(S.ej), std::move(S.ec), S.E);
}

}  // namespace rama
