// forest.hip — connected components, maximum spanning forest and the
// conflict-free contraction mapping of the MST branch
// (src/edge_contractions_woc.cu, src/edge_contractions_woc_thrust.cu,
//  semantics of external/ECL-CC and external/cudaMST/gpuMST).
//
// CC: lock-free union-find that always links the larger root under the
// smaller, so every component ends labelled by its minimum node id — the
// canonical labels ECL-CC produces — independent of scheduling.
//
// Spanning forest: Boruvka with a strict total order on edges (cost
// descending, edge index ascending) packed into one u64 for atomicMax, so
// the forest is the unique maximum spanning forest under that order.
//
// Contraction mapping: forest kept as row-sorted symmetric arrays
// (frow, fadj, fw) + offsets; conflicted 3-/4-cycles along the forest, then
// the bidirectional frontier search of the reference, all on device.
#include "rama_internal.hpp"

namespace rama {

__device__ __forceinline__ int cc_find(int* p, int x) {
    int y = __atomic_load_n(&p[x], __ATOMIC_RELAXED);
    while (y != x) {
        const int z = __atomic_load_n(&p[y], __ATOMIC_RELAXED);
        if (z != y) __atomic_store_n(&p[x], z, __ATOMIC_RELAXED);  // path halving (monotone)
        x = y;
        y = z;
    }
    return x;
}

__global__ void k_cc_hook(int n, const int* __restrict__ off, const int* __restrict__ adj, int* p) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    for (int k = off[v]; k < off[v + 1]; ++k) {
        const int u = adj[k];
        if (u <= v) continue;  // each undirected edge once
        while (true) {
            int a = cc_find(p, v), b = cc_find(p, u);
            if (a == b) break;
            if (a > b) { const int t = a; a = b; b = t; }
            if (atomicCAS(&p[b], b, a) == b) break;
        }
    }
}

// Each thread writes only its own entry and chases without path halving: a
// halving store could overwrite another node's freshly written root with a
// mere ancestor.
__global__ void k_cc_flatten(int n, int* p) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    int r = v;
    int q = __atomic_load_n(&p[r], __ATOMIC_RELAXED);
    while (q != r) {
        r = q;
        q = __atomic_load_n(&p[r], __ATOMIC_RELAXED);
    }
    __atomic_store_n(&p[v], r, __ATOMIC_RELAXED);
}

void connected_components(Ctx& ctx, const int* off, const int* adj, int n, int* labels) {
    if (n <= 0) return;
    iota(ctx, labels, n);
    k_cc_hook<<<grid_for(n), kBlock, 0, ctx.stream>>>(n, off, adj, labels);
    RAMA_LAUNCHED();
    k_cc_flatten<<<grid_for(n), kBlock, 0, ctx.stream>>>(n, labels);
    RAMA_LAUNCHED();
}

// ------------------------------------------------------------- Boruvka
__device__ __forceinline__ u64 edge_key(float w, u32 e) {
    u32 b = __float_as_uint(w);
    b = (b & 0x80000000u) ? ~b : (b | 0x80000000u);  // order-preserving float -> uint
    return ((u64)b << 32) | (u64)(0xFFFFFFFFu - e);
}

i64 max_spanning_forest(Ctx& ctx, const int* r, const int* c, const float* w, i64 m, int n, DBuf<int>& ids) {
    ids = ctx.buf<int>(m > 0 ? m : 1);
    if (m == 0 || n == 0) return 0;
    DBuf<int> comp = ctx.buf<int>(n), parent = ctx.buf<int>(n), alive = ctx.buf<int>(m), alive2 = ctx.buf<int>(m);
    DBuf<u64> best = ctx.buf<u64>(n);
    DBuf<int> in_forest = ctx.buf<int>(m);
    DBuf<int> flag = ctx.buf<int>(1);
    iota(ctx, comp.get(), n);
    iota(ctx, alive.get(), m);
    ctx.zero(in_forest.get(), m);
    int* CMP = comp.get();
    int* PAR = parent.get();
    u64* BEST = best.get();
    int* INF = in_forest.get();
    int* FL = flag.get();
    i64 na = m;
    for (int round = 0; round < 64 && na > 0; ++round) {
        ctx.zero(BEST, n);
        const int* AL = alive.get();
        for_each(ctx, na, [=] __device__(i64 q) {
            const int e = AL[q];
            const int cu = CMP[r[e]], cv = CMP[c[e]];
            if (cu == cv) return;
            const u64 k = edge_key(w[e], (u32)e);
            atomicMax((unsigned long long*)&BEST[cu], (unsigned long long)k);
            atomicMax((unsigned long long*)&BEST[cv], (unsigned long long)k);
        });
        int* fl = ctx.flag_slot(FL);
        for_each(ctx, n, [=] __device__(i64 v) {
            PAR[v] = (int)v;
            if (CMP[v] != (int)v) return;  // only component roots act
            const u64 k = BEST[v];
            if (k == 0) return;
            const int e = (int)(0xFFFFFFFFu - (u32)(k & 0xFFFFFFFFu));
            const int cu = CMP[r[e]], cv = CMP[c[e]];
            const int other = cu == (int)v ? cv : cu;
            const bool mutual = BEST[other] == k;
            if (!(mutual && (int)v < other)) PAR[v] = other;
            INF[e] = 1;
            fl[0] = 1;
        });
        if (ctx.flag_value(FL, fl) == 0) break;
        // pointer doubling on the component forest until every root is found
        for (int it = 0; it < 64; ++it) {
            int* fl2 = ctx.flag_slot(FL);
            for_each(ctx, n, [=] __device__(i64 v) {
                const int p = PAR[v], pp = PAR[p];
                if (p != pp) {
                    PAR[v] = pp;
                    fl2[0] = 1;
                }
            });
            if (ctx.flag_value(FL, fl2) == 0) break;
        }
        for_each(ctx, n, [=] __device__(i64 v) { CMP[v] = PAR[CMP[v]]; });
        // keep only edges that still join two components
        DBuf<int> pos;
        const i64 na2 = select_positions(
            ctx, [=] __device__(i64 q) { const int e = AL[q]; return (int)(CMP[r[e]] != CMP[c[e]]); }, na, pos);
        int* A2 = alive2.get();
        const int* P = pos.get();
        for_each(ctx, na, [=] __device__(i64 q) {
            const int e = AL[q];
            if (CMP[r[e]] != CMP[c[e]]) A2[P[q]] = e;
        });
        std::swap(alive, alive2);
        na = na2;
    }
    DBuf<int> pos;
    const i64 k = select_positions(ctx, [=] __device__(i64 e) { return INF[e]; }, m, pos);
    int* IDS = ids.get();
    const int* P = pos.get();
    for_each(ctx, m, [=] __device__(i64 e) {
        if (INF[e]) IDS[P[e]] = (int)e;
    });
    return k;
}

// -------------------------------------------------- contraction mapping

// Lowest common neighbour walk on the forest (include/rama_utils.h:62-84).
__device__ __forceinline__ int lcn(int v1, int v2, const int* off, const int* adj, int& i1, int& i2) {
    const int e1 = off[v1 + 1], e2 = off[v2 + 1];
    while (i1 < e1 && i2 < e2) {
        const int a = adj[i1], b = adj[i2];
        if (a == b) {
            ++i1;
            ++i2;
            return a;
        }
        if (a < b) ++i1;
        if (a > b) ++i2;
    }
    return -1;
}

// Index of the forest entry (row, col) in row-sorted arrays, -1 if absent.
__device__ __forceinline__ int find_entry(const int* off, const int* adj, int row, int col) {
    int lo = off[row], hi = off[row + 1] - 1;
    while (lo <= hi) {
        const int mid = (lo + hi) >> 1;
        const int a = adj[mid];
        if (a == col) return mid;
        if (a < col) lo = mid + 1; else hi = mid - 1;
    }
    return -1;
}

struct WocState {
    int n = 0;  // num_nodes = A.max_dim()
    DBuf<int> frow, fadj, off;
    DBuf<float> fw;
    i64 F = 0;  // forest entries (both directions)
    DBuf<int> rr, rc;
    i64 R = 0;
    DBuf<int> cc;
    // frontiers
    struct Frontier {
        DBuf<int> node, parent, rep, bidx;
        DBuf<float> bval;
        i64 size = 0;
    } rowf, colf;
};

static void forest_offsets(Ctx& ctx, WocState& S) { compute_offsets(ctx, S.frow.get(), S.F, S.n - 1, S.off.get()); }

// Remove the forest entries flagged in rm (both directions already flagged).
static void forest_compact(Ctx& ctx, WocState& S, const int* rm) {
    DBuf<int> pos;
    const i64 k = select_positions(ctx, [=] __device__(i64 e) { return (int)(rm[e] == 0); }, S.F, pos);
    DBuf<int> nr = ctx.buf<int>(k), na = ctx.buf<int>(k);
    DBuf<float> nw = ctx.buf<float>(k);
    int* NR = nr.get();
    int* NA = na.get();
    float* NW = nw.get();
    const int* FR = S.frow.get();
    const int* FA = S.fadj.get();
    const float* FW = S.fw.get();
    const int* P = pos.get();
    for_each(ctx, S.F, [=] __device__(i64 e) {
        if (rm[e] == 0) {
            NR[P[e]] = FR[e];
            NA[P[e]] = FA[e];
            NW[P[e]] = FW[e];
        }
    });
    S.frow = std::move(nr);
    S.fadj = std::move(na);
    S.fw = std::move(nw);
    S.F = k;
    forest_offsets(ctx, S);
}

// Flag entry k and its reverse for removal.
__device__ __forceinline__ void flag_pair(int* rm, const int* off, const int* frow, const int* fadj, int k) {
    rm[k] = 1;
    const int rev = find_entry(off, fadj, fadj[k], frow[k]);
    if (rev >= 0) rm[rev] = 1;
}

static void frontier_reset(Ctx& ctx, WocState::Frontier& f, const int* seeds, i64 R) {
    f.size = R;
    f.node = ctx.buf<int>(R);
    f.parent = ctx.buf<int>(R);
    f.rep = ctx.buf<int>(R);
    f.bidx = ctx.buf<int>(R);
    f.bval = ctx.buf<float>(R);
    ctx.d2d(f.node.get(), seeds, R);
    fill(ctx, f.parent.get(), R, -1);
    iota(ctx, f.rep.get(), R);
    ctx.zero(f.bidx.get(), R);
    fill(ctx, f.bval.get(), R, 3.402823466e38f);
}

// filter_by_cc (src/edge_contractions_woc.cu:257-276; _thrust.cu:95-133).
static i64 filter_by_cc(Ctx& ctx, WocState& S, bool reset_frontiers) {
    forest_offsets(ctx, S);
    connected_components(ctx, S.off.get(), S.fadj.get(), S.n, S.cc.get());
    const int* CC = S.cc.get();
    const int* RR = S.rr.get();
    const int* RC = S.rc.get();
    DBuf<int> pos;
    const i64 k = select_positions(ctx, [=] __device__(i64 e) { return (int)(CC[RR[e]] == CC[RC[e]]); }, S.R, pos);
    DBuf<int> nr = ctx.buf<int>(k), nc = ctx.buf<int>(k);
    int* NR = nr.get();
    int* NC = nc.get();
    const int* P = pos.get();
    for_each(ctx, S.R, [=] __device__(i64 e) {
        if (CC[RR[e]] == CC[RC[e]]) {
            NR[P[e]] = RR[e];
            NC[P[e]] = RC[e];
        }
    });
    S.rr = std::move(nr);
    S.rc = std::move(nc);
    S.R = k;
    if (reset_frontiers) {
        frontier_reset(ctx, S.rowf, S.rr.get(), k);
        frontier_reset(ctx, S.colf, S.rc.get(), k);
    }
    return k;
}

// check_triangles (src/edge_contractions_woc.cu:10-37,141-161)
static bool check_triangles(Ctx& ctx, WocState& S) {
    if (S.R == 0) return false;
    DBuf<int> rm = ctx.buf<int>(S.F), any = ctx.buf<int>(1);
    ctx.zero(rm.get(), S.F);
    int* RM = rm.get();
    int* ANY = any.get();
    const int* OFF = S.off.get();
    const int* ADJ = S.fadj.get();
    const int* FR = S.frow.get();
    const float* FW = S.fw.get();
    const int* RR = S.rr.get();
    const int* RC = S.rc.get();
    int* anyf = ctx.flag_slot(ANY);  // raised by the kernel below
    for_each(ctx, S.R, [=] __device__(i64 r) {
        const int v1 = RR[r], v2 = RC[r];
        int i1 = OFF[v1], i2 = OFF[v2];
        const int mid = lcn(v1, v2, OFF, ADJ, i1, i2);
        if (mid >= 0) {
            const int k = FW[i1 - 1] < FW[i2 - 1] ? i1 - 1 : i2 - 1;
            flag_pair(RM, OFF, FR, ADJ, k);
            anyf[0] = 1;
        }
    });
    if (ctx.flag_value(ANY, anyf) == 0) return false;
    forest_compact(ctx, S, RM);
    return true;
}

// check_quadrangles (src/edge_contractions_woc.cu:39-79,163-193)
static bool check_quadrangles(Ctx& ctx, WocState& S) {
    if (S.R == 0) return false;
    const int* OFF = S.off.get();
    const int* ADJ = S.fadj.get();
    const int* FR = S.frow.get();
    const float* FW = S.fw.get();
    const int* RR = S.rr.get();
    const int* RC = S.rc.get();
    DBuf<i64> xoff = ctx.buf<i64>(S.R + 1);
    i64* xm = ctx.mirror<i64>();
    scan_exclusive<i64>(ctx, [=] __device__(i64 r) { const int v = RR[r]; return (i64)(OFF[v + 1] - OFF[v]); },
                        xoff.get(), S.R, xm);
    i64 X;
    ctx.read_mirror(xoff.get() + S.R, xm, &X);
    if (X == 0) return false;
    DBuf<int> rm = ctx.buf<int>(S.F), any = ctx.buf<int>(1);
    ctx.zero(rm.get(), S.F);
    int* RM = rm.get();
    int* ANY = any.get();
    const i64* XO = xoff.get();
    const i64 R = S.R;
    int* anyf = ctx.flag_slot(ANY);  // raised by the kernel below
    for_each(ctx, X, [=] __device__(i64 x) {
        i64 lo = 0, hi = R - 1;  // largest r with XO[r] <= x
        while (lo < hi) {
            const i64 mid = (lo + hi + 1) >> 1;
            if (XO[mid] <= x) lo = mid; else hi = mid - 1;
        }
        const int v1 = RR[lo], v2 = RC[lo];
        const int e1 = OFF[v1] + (int)(x - XO[lo]);
        const int n1 = ADJ[e1];
        int i1 = OFF[n1], i2 = OFF[v2];
        const int mid = lcn(n1, v2, OFF, ADJ, i1, i2);
        if (mid >= 0) {
            const float a = FW[e1], b = FW[i1 - 1], c = FW[i2 - 1];
            const int k = (a < b && a < c) ? e1 : (b < c ? i1 - 1 : i2 - 1);
            flag_pair(RM, OFF, FR, ADJ, k);
            anyf[0] = 1;
        }
    });
    if (ctx.flag_value(ANY, anyf) == 0) return false;
    forest_compact(ctx, S, RM);
    return true;
}

// filter_by_thresholding (src/edge_contractions_woc_thrust.cu:513-542)
static bool filter_by_thresholding(Ctx& ctx, WocState& S, float mult) {
    const float* FW = S.fw.get();
    DBuf<double> s = ctx.buf<double>(2);
    double* sm = ctx.mirror<double>();
    reduce_to<double>(ctx, [=] __device__(i64 e) { const float x = FW[e]; return x >= 0.0f ? (double)x : 0.0; }, S.F,
                      0.0, OpSum(), s.get(), sm);
    reduce_to<double>(ctx, [=] __device__(i64 e) { return FW[e] >= 0.0f ? 1.0 : 0.0; }, S.F, 0.0, OpSum(),
                      s.get() + 1, sm ? sm + 1 : nullptr);
    double h[2];
    ctx.read_mirror(s.get(), sm, h, 2);
    const float sf = (float)h[0];
    const float thr = mult * sf / (float)(i64)h[1];
    DBuf<int> rm = ctx.buf<int>(S.F);
    int* RM = rm.get();
    for_each(ctx, S.F, [=] __device__(i64 e) { RM[e] = FW[e] < thr ? 1 : 0; });
    const i64 cnt = reduce<i64>(ctx, [=] __device__(i64 e) { return (i64)RM[e]; }, S.F, (i64)0, OpSum());
    if (cnt == 0) return false;
    forest_compact(ctx, S, RM);
    return true;
}

// expand_frontier + expand_cuda (src/edge_contractions_woc_thrust.cu:149-199,412-469)
static void expand(Ctx& ctx, WocState& S, WocState::Frontier& f) {
    const int* OFF = S.off.get();
    const int* ADJ = S.fadj.get();
    const float* FW = S.fw.get();
    const int* ND = f.node.get();
    const int* PA = f.parent.get();
    DBuf<int> xo = ctx.buf<int>(f.size + 1);
    int* xm = ctx.mirror<int>();
    scan_exclusive<int>(
        ctx, [=] __device__(i64 q) { const int v = ND[q]; return OFF[v + 1] - OFF[v] - (PA[q] != -1 ? 1 : 0); },
        xo.get(), f.size, xm);
    int xh;
    ctx.read_mirror(xo.get() + f.size, xm, &xh);
    const i64 X = xh;
    WocState::Frontier g;
    g.size = X;
    g.node = ctx.buf<int>(X);
    g.parent = ctx.buf<int>(X);
    g.rep = ctx.buf<int>(X);
    g.bidx = ctx.buf<int>(X);
    g.bval = ctx.buf<float>(X);
    int* GN = g.node.get();
    int* GP = g.parent.get();
    int* GR = g.rep.get();
    int* GB = g.bidx.get();
    float* GV = g.bval.get();
    const int* RP = f.rep.get();
    const int* BI = f.bidx.get();
    const float* BV = f.bval.get();
    const int* XO = xo.get();
    for_each(ctx, f.size, [=] __device__(i64 q) {
        const int src = ND[q], par = PA[q];
        int o = XO[q];
        const int oend = XO[q + 1];
        for (int k = OFF[src]; k < OFF[src + 1] && o < oend; ++k) {
            const int dst = ADJ[k];
            if (dst == par) continue;
            GN[o] = dst;
            GR[o] = RP[q];
            GP[o] = src;
            if (FW[k] < BV[q]) {
                GB[o] = k;
                GV[o] = FW[k];
            } else {
                GB[o] = BI[q];
                GV[o] = BV[q];
            }
            ++o;
        }
    });
    f = std::move(g);
}

// Sort a frontier by (node, rep); returns the sorted keys and the order.
static void frontier_sorted(Ctx& ctx, const WocState::Frontier& f, int rb, DBuf<u64>& keys, DBuf<u32>& order,
                            int kb_total) {
    keys = ctx.buf<u64>(f.size);
    order = ctx.buf<u32>(f.size);
    u64* K = keys.get();
    u32* O = order.get();
    const int* ND = f.node.get();
    const int* RP = f.rep.get();
    for_each(ctx, f.size, [=] __device__(i64 q) {
        K[q] = ((u64)(u32)ND[q] << rb) | (u32)RP[q];
        O[q] = (u32)q;
    });
    prof::Site site_("forest");
    radix_sort_pairs(ctx, keys, order, f.size, kb_total);
}

// filter_cycles (src/edge_contractions_woc_thrust.cu:273-410): a (node, rep)
// key present in both frontiers closes a conflicted cycle; the bottleneck
// with the smaller cost is removed (column side on ties).
static bool filter_cycles(Ctx& ctx, WocState& S) {
    if (S.rowf.size == 0 || S.colf.size == 0) return false;
    const int rb = bits_for(S.R > 1 ? S.R - 1 : 1);
    const int kb = bits_for(S.n > 1 ? S.n - 1 : 1) + rb;
    DBuf<u64> kr, kc;
    DBuf<u32> orr, oc;
    frontier_sorted(ctx, S.rowf, rb, kr, orr, kb);
    frontier_sorted(ctx, S.colf, rb, kc, oc, kb);
    DBuf<int> rm = ctx.buf<int>(S.F), any = ctx.buf<int>(1);
    ctx.zero(rm.get(), S.F);
    int* RM = rm.get();
    int* ANY = any.get();
    const u64* KR = kr.get();
    const u64* KC = kc.get();
    const u32* OR = orr.get();
    const u32* OC = oc.get();
    const i64 nc = S.colf.size;
    const int* RBI = S.rowf.bidx.get();
    const float* RBV = S.rowf.bval.get();
    const int* CBI = S.colf.bidx.get();
    const float* CBV = S.colf.bval.get();
    const int* OFF = S.off.get();
    const int* FR = S.frow.get();
    const int* ADJ = S.fadj.get();
    int* anyf = ctx.flag_slot(ANY);  // raised by the kernel below
    for_each(ctx, S.rowf.size, [=] __device__(i64 q) {
        const u64 k = KR[q];
        i64 lo = 0, hi = nc - 1;
        while (lo < hi) {
            const i64 mid = (lo + hi) >> 1;
            if (KC[mid] < k) lo = mid + 1; else hi = mid;
        }
        if (KC[lo] != k) return;
        const u32 a = OR[q], b = OC[lo];
        const int idx = RBV[a] < CBV[b] ? RBI[a] : CBI[b];
        flag_pair(RM, OFF, FR, ADJ, idx);
        anyf[0] = 1;
    });
    if (ctx.flag_value(ANY, anyf) == 0) return false;
    forest_compact(ctx, S, RM);
    return true;
}

// edge_contractions_woc(A) + find_contraction_mapping
// (src/edge_contractions_woc.cu:341-423, src/edge_contractions_woc_thrust.cu:544-599).
int woc_mapping(Ctx& ctx, const DCoo& A, DBuf<int>& map, i64& map_size) {
    WocState S;
    S.n = A.max_dim();
    map_size = 0;
    const float* W = A.val.get();
    const int* AR = A.row.get();
    const int* AC = A.col.get();
    // partition_copy: attractive c > 0, repulsive the rest (stable)
    DBuf<int> ppos;
    const i64 P = select_positions(ctx, [=] __device__(i64 e) { return (int)(W[e] > 0.0f); }, A.nnz, ppos);
    if (P == 0) return 0;
    S.R = A.nnz - P;
    DBuf<int> pr = ctx.buf<int>(P), pc = ctx.buf<int>(P);
    DBuf<float> pw = ctx.buf<float>(P);
    S.rr = ctx.buf<int>(S.R);
    S.rc = ctx.buf<int>(S.R);
    {
        int* PR = pr.get();
        int* PC = pc.get();
        float* PW = pw.get();
        int* RR = S.rr.get();
        int* RC = S.rc.get();
        const int* PP = ppos.get();
        for_each(ctx, A.nnz, [=] __device__(i64 e) {
            const int q = PP[e];
            if (W[e] > 0.0f) {
                PR[q] = AR[e];
                PC[q] = AC[e];
                PW[q] = W[e];
            } else {
                RR[e - q] = AR[e];
                RC[e - q] = AC[e];
            }
        });
    }
    // maximum spanning forest, symmetrised and sorted
    DBuf<int> ids;
    const i64 k = max_spanning_forest(ctx, pr.get(), pc.get(), pw.get(), P, S.n, ids);
    {
        DBuf<int> fr = ctx.buf<int>(k), fc = ctx.buf<int>(k);
        DBuf<float> fw = ctx.buf<float>(k);
        int* FR = fr.get();
        int* FC = fc.get();
        float* FWp = fw.get();
        const int* ID = ids.get();
        const int* PR = pr.get();
        const int* PC = pc.get();
        const float* PW = pw.get();
        for_each(ctx, k, [=] __device__(i64 q) {
            FR[q] = PR[ID[q]];
            FC[q] = PC[ID[q]];
            FWp[q] = PW[ID[q]];
        });
        Csr F = symmetric_csr_forest(ctx, fr.get(), fc.get(), fw.get(), k, S.n);
        S.F = F.nnz;
        S.fadj = std::move(F.adj);
        S.fw = std::move(F.w);
        S.off = std::move(F.off);
        S.frow = ctx.buf<int>(S.F);
        int* FRW = S.frow.get();
        const int* OFF = S.off.get();
        for_each(ctx, S.n, [=] __device__(i64 v) {
            for (int q = OFF[v]; q < OFF[v + 1]; ++q) FRW[q] = (int)v;
        });
    }
    S.cc = ctx.buf<int>(S.n);
    if (S.F == 0) return 0;
    if (ctx.verbose) std::printf("# MST edges %lld, # Repulsive edges %lld\n", (long long)S.F, (long long)S.R);
    filter_by_cc(ctx, S, false);
    if (check_triangles(ctx, S)) filter_by_cc(ctx, S, false);
    check_quadrangles(ctx, S);
    if (S.F == 0) return 0;
    // frontier mapper
    i64 num_rep = filter_by_cc(ctx, S, true);
    int itr = 0;
    while (num_rep > 0 && S.F > 0) {
        if (itr % 5 == 0)
            if (filter_by_thresholding(ctx, S, 0.9f)) num_rep = filter_by_cc(ctx, S, true);
        if (num_rep == 0 || S.F == 0) break;
        if (S.rowf.size == 0 && S.colf.size == 0) break;  // both searches exhausted: nothing left to find
        expand(ctx, S, S.rowf);
        if (filter_cycles(ctx, S)) num_rep = filter_by_cc(ctx, S, true);
        if (num_rep == 0 || S.F == 0) break;
        expand(ctx, S, S.colf);
        if (filter_cycles(ctx, S)) num_rep = filter_by_cc(ctx, S, true);
        itr++;
    }
    map = ctx.buf<int>(S.n);
    compress_labels(ctx, S.cc.get(), S.n, S.n - 1, map.get(), false);
    map_size = S.n;
    if (ctx.verbose) std::printf("Found conflict-free contraction mapping (%lld forest entries)\n", (long long)S.F);
    return (int)S.F;
}

}  // namespace rama
