// rama_oracle.cpp — TEST INFRASTRUCTURE ONLY.
//
// Serial, deterministic CPU restatement of RAMA's primal-dual multicut solver
// (pawelswoboda/RAMA).  It is the checker for the HIP library: only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
// product path never links or calls it.
//
// Every function cites the reference file:line it restates.  Where the
// reference is nondeterministic (float atomics, budget races, reduction
// association) the oracle follows the deterministic policy the HIP library
// implements, stated at the function:
//   P1  sorts are stable (thrust's zip-key sort_by_key is a stable merge sort);
//   P2  sums that feed a decision or a stored cost accumulate in double in a
//       fixed order and are rounded to float once;
//   P3  the triangle budget keeps the first `budget` triangle writes in
//       enumeration order (rep edge, expansion, common neighbour);
//   P4  message-passing edge updates sum the triangle deltas in (triangle,
//       slot) order instead of float atomicAdd;
//   P5  the spanning forest is the unique maximum spanning forest under the
//       total order (cost descending, edge index ascending).
//
// Built by oracle/build.py with g++ -O2 -ffp-contract=off (no FMA) so its
// float arithmetic rounds exactly as the HIP kernels compiled with
// -ffp-contract=off do.

#include <algorithm>
#include <array>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <stdexcept>
#include <string>
#include <tuple>
#include <unordered_set>
#include <vector>

#include "../include/rama_hip.h"

namespace {

thread_local std::string g_err;

struct Coo {
    std::vector<int> row, col;
    std::vector<float> val;
    int rows_ = 0, cols_ = 0;
    bool directed = true;
    size_t nnz() const { return val.size(); }
    int max_dim() const { return std::max(rows_, cols_); }
};

// Stable sort of (row, col, val) by (row, col)  [P1]
// include/rama_utils.h:287-294 (coo_sorting with float payload)
void coo_sort(std::vector<int>& r, std::vector<int>& c, std::vector<float>& v) {
    const size_t n = r.size();
    std::vector<size_t> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) {
        return r[a] != r[b] ? r[a] < r[b] : c[a] < c[b];
    });
    std::vector<int> r2(n), c2(n);
    std::vector<float> v2(n);
    for (size_t k = 0; k < n; ++k) { r2[k] = r[idx[k]]; c2[k] = c[idx[k]]; v2[k] = v[idx[k]]; }
    r.swap(r2); c.swap(c2); v.swap(v2);
}

// dCOO::init (src/dCOO.cu:28-55) with sort_edge_nodes (include/rama_utils.h:231-254)
void coo_init(Coo& A, bool is_sorted) {
    if (!is_sorted) {
        if (A.directed)
            for (size_t e = 0; e < A.nnz(); ++e) {
                int a = A.row[e], b = A.col[e];
                A.row[e] = std::min(a, b);
                A.col[e] = std::max(a, b);
            }
        coo_sort(A.row, A.col, A.val);
    }
    if (A.nnz() == 0) return;
    if (A.cols_ == 0) A.cols_ = *std::max_element(A.col.begin(), A.col.end()) + 1;
    if (A.rows_ == 0) A.rows_ = A.row.back() + 1;
}

// compute_offsets (include/rama_utils.h:403-413): offsets of size max_value+2
std::vector<int> compute_offsets(const std::vector<int>& ids, int max_value) {
    std::vector<int> off(max_value + 2, 0);
    for (int v : ids) off[v + 1]++;
    for (size_t k = 1; k < off.size(); ++k) off[k] += off[k - 1];
    return off;
}

// compress_label_sequence (include/rama_utils.h:160-177)
std::vector<int> compress_label_sequence(const std::vector<int>& data, int max_label) {
    std::vector<int> mask(max_label + 1, 0);
    for (int d : data) mask[d] = 1;
    std::vector<int> excl(max_label + 1, 0);
    int run = 0;
    for (int k = 0; k <= max_label; ++k) { excl[k] = run; run += mask[k]; }
    std::vector<int> out(data.size());
    for (size_t k = 0; k < data.size(); ++k) out[k] = excl[data[k]];
    return out;
}

// get_lb (include/rama_utils.h:179-190): sum of negative parts in double
double get_lb(const std::vector<float>& c) {
    double s = 0.0;
    for (float x : c) s += (x < 0.0f ? (double)x : 0.0);
    return s;
}

// map_nodes + contract_cuda (src/dCOO.cu:14-26,57-95).  Duplicate (row,col)
// costs are summed in double in sorted (= input, P1) order, rounded once [P2].
Coo contract(const Coo& A, const std::vector<int>& map) {
    const size_t n = A.nnz();
    std::vector<int> r(n), c(n);
    std::vector<float> v(A.val);
    for (size_t e = 0; e < n; ++e) {
        int a = map[A.row[e]], b = map[A.col[e]];
        r[e] = std::min(a, b);
        c[e] = std::max(a, b);
    }
    coo_sort(r, c, v);
    Coo B;
    B.directed = true;
    for (size_t e = 0; e < n;) {
        size_t f = e;
        double s = 0.0;
        while (f < n && r[f] == r[e] && c[f] == c[e]) { s += (double)v[f]; ++f; }
        B.row.push_back(r[e]);
        B.col.push_back(c[e]);
        B.val.push_back((float)s);
        e = f;
    }
    B.rows_ = B.row.back() + 1;
    B.cols_ = *std::max_element(B.col.begin(), B.col.end()) + 1;
    return B;
}

// dCOO::diagonal + has_bad_contractions (src/dCOO.cu:120-142, src/rama_cuda.cu:18-30)
bool has_bad_contractions(const Coo& A, double* diag_sum) {
    bool bad = false;
    double s = 0.0;
    for (size_t e = 0; e < A.nnz(); ++e)
        if (A.row[e] == A.col[e]) {
            s += A.val[e];
            if (A.val[e] < 0.0f) bad = true;
        }
    if (diag_sum) *diag_sum = s;
    return bad;
}

// dCOO::remove_diagonal (src/dCOO.cu:97-115): stable, keeps rows_/cols_
void remove_diagonal(Coo& A) {
    size_t w = 0;
    for (size_t e = 0; e < A.nnz(); ++e)
        if (A.row[e] != A.col[e]) {
            A.row[w] = A.row[e]; A.col[w] = A.col[e]; A.val[w] = A.val[e]; ++w;
        }
    A.row.resize(w); A.col.resize(w); A.val.resize(w);
}

// export_undirected + to_undirected (src/dCOO.cu:164-172, include/rama_utils.h:107-131)
Coo export_undirected(const Coo& A) {
    Coo U;
    U.directed = false;
    const size_t n = A.nnz();
    U.row.resize(2 * n); U.col.resize(2 * n); U.val.resize(2 * n);
    for (size_t e = 0; e < n; ++e) {
        U.row[e] = A.row[e]; U.col[e] = A.col[e]; U.val[e] = A.val[e];
        U.row[n + e] = A.col[e]; U.col[n + e] = A.row[e]; U.val[n + e] = A.val[e];
    }
    coo_init(U, false);
    return U;
}

// Mean of the non-negative entries times a multiplier, as
// determine_matching_threshold (src/maximum_matching_vertex_based.cu:76-86) and
// filter_by_thresholding (src/edge_contractions_woc_thrust.cu:522-528): the
// sum accumulates in double and is rounded to float [P2]; the rest is float.
float positive_mean_threshold(const std::vector<float>& data, float mult, bool* any) {
    double s = 0.0;
    int cnt = 0;
    for (float x : data)
        if (x >= 0.0f) { s += (double)x; ++cnt; }
    *any = cnt > 0;
    const float sf = (float)s;
    return mult * sf / (float)cnt;
}

// filter_edges_by_matching_vertex_based (src/maximum_matching_vertex_based.cu:10-137)
std::pair<std::vector<int>, int> matching_vertex_based(const Coo& U, float mm) {
    const int n = U.rows_;
    std::vector<int> best(n, -1), matched(n, 0), map(n);
    std::iota(map.begin(), map.end(), 0);
    const std::vector<int> off = compute_offsets(U.row, n - 1);
    bool any;
    const float thr = positive_mean_threshold(U.val, mm, &any);
    if (!any || thr < 0) return {map, 0};
    int prev = 0;
    for (int t = 0; t < 10; ++t) {
        bool running = false;
        // pick_best_neighbour / get_best_neighbour (:10-47)
        for (int v = 0; v < n; ++v) {
            if (matched[v]) continue;
            float mx = thr;
            int nid = -1;
            for (int l = off[v]; l < off[v + 1]; ++l) {
                const float cv = U.val[l];
                const int cn = U.col[l];
                if (cv <= mx || matched[cn] || cn == v) continue;
                mx = cv;
                nid = cn;
            }
            best[v] = nid;
        }
        // match_neighbours (:49-74)
        for (int v = 0; v < n; ++v) {
            if (matched[v]) continue;
            const int vb = best[v];
            if (vb == -1 || matched[vb]) continue;
            if (best[vb] == v) {
                map[std::max(vb, v)] = std::min(vb, v);
                matched[v] = 1;
                matched[vb] = 1;
                running = true;
            }
        }
        int cur = 0;
        for (int v = 0; v < n; ++v) cur += matched[v];
        const float rel = (cur - prev) / (prev + 1.0f);
        prev = cur;
        if (!running || rel < 0.1f) break;
    }
    return {map, prev};
}

// ---------------------------------------------------------------- cycles

struct Csr {
    std::vector<int> off, adj;
    std::vector<float> w;
};

// Lowest common neighbour walk (include/rama_utils.h:62-84).
inline int lcn(int v1, int v2, const Csr& G, int& i1, int& i2) {
    while (i1 < G.off[v1 + 1] && i2 < G.off[v2 + 1]) {
        const int a = G.adj[i1], b = G.adj[i2];
        if (a == b) { ++i1; ++i2; return a; }
        if (a < b) ++i1;
        if (a > b) ++i2;
    }
    return -1;
}

struct TriBuf {
    std::vector<int> t1, t2, t3;
    int64_t budget = 0;
    int64_t count = 0;  // attempted writes (may exceed budget)
    // write_triangle (src/conflicted_cycles_cuda.cu:13-29) [P3]
    bool write(int a, int b, int c) {
        const int64_t k = count++;
        if (k >= budget) return true;
        t1[k] = std::min(a, std::min(b, c));
        t2[k] = std::max(std::min(a, b), std::min(std::max(a, b), c));
        t3[k] = std::max(a, std::max(b, c));
        return false;
    }
    bool full() const { return count >= budget; }
};

// rearrange_triangles (include/rama_utils.h:352-363): drop degenerate, sort
// vertices, sort triples, unique; returns the new count.
int64_t rearrange(TriBuf& T, int64_t valid) {
    std::vector<std::array<int, 3>> v;
    v.reserve(valid);
    for (int64_t k = 0; k < valid; ++k) {
        int a = T.t1[k], b = T.t2[k], c = T.t3[k];
        if (a == b || a == c || b == c) continue;
        int lo = std::min(a, std::min(b, c)), hi = std::max(a, std::max(b, c));
        int mid = std::max(std::min(a, b), std::min(std::max(a, b), c));
        v.push_back({lo, mid, hi});
    }
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
    for (size_t k = 0; k < v.size(); ++k) { T.t1[k] = v[k][0]; T.t2[k] = v[k][1]; T.t3[k] = v[k][2]; }
    return (int64_t)v.size();
}

// conflicted_cycles_cuda (src/conflicted_cycles_cuda.cu:170-343).  Returns the
// raw triangle slots (reference triangles_v1..3 after the final resize,
// truncated at the budget).
void conflicted_cycles(const Coo& A, int max_len, float factor, float tol, std::vector<int>& o1,
                       std::vector<int>& o2, std::vector<int>& o3) {
    o1.clear(); o2.clear(); o3.clear();
    if (max_len > 5) throw std::runtime_error("max_cycle_length should be <= 5. Received: " + std::to_string(max_len));
    if (max_len < 3 || A.nnz() == 0) return;
    // determine_threshold (:170-174)
    const float mn = *std::min_element(A.val.begin(), A.val.end());
    const float thr = mn * tol;
    // create_matrices (:176-219)
    std::vector<int> pr, pc, rr, rc;
    std::vector<float> pv;
    for (size_t e = 0; e < A.nnz(); ++e) {
        if (A.val[e] > -1.0f * thr) { pr.push_back(A.row[e]); pc.push_back(A.col[e]); pv.push_back(A.val[e]); }
        if (A.val[e] < thr) { rr.push_back(A.row[e]); rc.push_back(A.col[e]); }
    }
    if (pr.empty() || rr.empty()) return;
    const int N = A.max_dim();
    Csr G;
    {
        const size_t p = pr.size();
        std::vector<int> r(2 * p), c(2 * p);
        std::vector<float> v(2 * p);
        for (size_t e = 0; e < p; ++e) {
            r[e] = pr[e]; c[e] = pc[e]; v[e] = pv[e];
            r[p + e] = pc[e]; c[p + e] = pr[e]; v[p + e] = pv[e];
        }
        coo_sort(r, c, v);
        G.off = compute_offsets(r, N - 1);
        G.adj = c;
        G.w = v;
    }
    const int64_t R = (int64_t)rr.size();
    TriBuf T;
    T.budget = (int64_t)(int)(factor * (float)R);
    T.t1.assign(T.budget, 0); T.t2.assign(T.budget, 0); T.t3.assign(T.budget, 0);
    // find_triangles_parallel (:31-61)
    for (int64_t r = 0; r < R && !T.full(); ++r) {
        const int v1 = rr[r], v2 = rc[r];
        int i1 = G.off[v1], i2 = G.off[v2];
        bool filled = false;
        while (!filled) {
            const int mid = lcn(v1, v2, G, i1, i2);
            if (mid == -1) break;
            filled = T.write(v1, v2, mid);
        }
    }
    if (max_len >= 4 && T.count < T.budget) {
        T.count = rearrange(T, T.count);
        // find_quadrangles_parallel (:63-101); expansion order (rep, local)
        for (int64_t r = 0; r < R && !T.full(); ++r) {
            const int v1 = rr[r], v2 = rc[r];
            const int d1 = G.off[v1 + 1] - G.off[v1];
            for (int l = 0; l < d1 && !T.full(); ++l) {
                const int n1 = G.adj[G.off[v1] + l];
                int i1 = G.off[n1], i2 = G.off[v2];
                bool filled = false;
                while (!filled) {
                    const int mid = lcn(n1, v2, G, i1, i2);
                    if (mid == -1) break;
                    T.write(v1, v2, n1);
                    filled = T.write(v2, n1, mid);
                }
            }
        }
    }
    if (max_len >= 5 && T.count < T.budget) {
        T.count = rearrange(T, T.count);
        // find_pentagons_parallel (:103-148); local = l1 + l2 * deg(v1)
        for (int64_t r = 0; r < R && !T.full(); ++r) {
            const int v1 = rr[r], v2 = rc[r];
            const int d1 = G.off[v1 + 1] - G.off[v1];
            const int d2 = G.off[v2 + 1] - G.off[v2];
            const int64_t X = (int64_t)d1 * d2;
            for (int64_t loc = 0; loc < X && !T.full(); ++loc) {
                const int l1 = (int)(loc % d1), l2 = (int)(loc / d1);
                const int n1 = G.adj[G.off[v1] + l1];
                const int n2 = G.adj[G.off[v2] + l2];
                if (n1 == n2) continue;
                int i1 = G.off[n1], i2 = G.off[n2];
                bool filled = false;
                while (!filled) {
                    const int mid = lcn(n1, n2, G, i1, i2);
                    if (mid == -1) break;
                    T.write(v1, v2, n1);
                    T.write(v2, n1, mid);
                    filled = T.write(v2, mid, n2);
                }
            }
        }
    }
    const int64_t n = std::min(T.count, T.budget);
    o1.assign(T.t1.begin(), T.t1.begin() + n);
    o2.assign(T.t2.begin(), T.t2.begin() + n);
    o3.assign(T.t3.begin(), T.t3.begin() + n);
}

// normalize_triangles (include/rama_utils.h:325-350)
void normalize_triangles(std::vector<int>& t1, std::vector<int>& t2, std::vector<int>& t3) {
    TriBuf T;
    T.t1 = t1; T.t2 = t2; T.t3 = t3;
    const int64_t n = rearrange(T, (int64_t)t1.size());
    t1.assign(T.t1.begin(), T.t1.begin() + n);
    t2.assign(T.t2.begin(), T.t2.begin() + n);
    t3.assign(T.t3.begin(), T.t3.begin() + n);
}

// ------------------------------------------------------- message passing

inline uint64_t key2(int a, int b) { return ((uint64_t)(uint32_t)a << 32) | (uint32_t)b; }

struct MP {
    std::vector<int> i, j, t1, t2, t3, c12, c13, c23, cnt;
    std::vector<float> c, t12, t13, t23;
    // slot lists per edge in (triangle, slot) order [P4]
    std::vector<int64_t> soff;
    std::vector<int64_t> slots;  // slot id = 3*t + k
    std::vector<float> delta;

    // constructor + compute_triangle_edge_correspondence (src/multicut_message_passing.cu:15-176)
    MP(const Coo& A, std::vector<int> a1, std::vector<int> a2, std::vector<int> a3) {
        normalize_triangles(a1, a2, a3);
        t1 = a1; t2 = a2; t3 = a3;
        const int64_t T = (int64_t)t1.size();
        // triangle edges, sorted and unique (:81-99)
        std::vector<uint64_t> tk;
        tk.reserve(3 * T);
        for (int64_t t = 0; t < T; ++t) {
            tk.push_back(key2(t1[t], t2[t]));
            tk.push_back(key2(t1[t], t3[t]));
            tk.push_back(key2(t2[t], t3[t]));
        }
        std::sort(tk.begin(), tk.end());
        tk.erase(std::unique(tk.begin(), tk.end()), tk.end());
        // set_union_by_key(A, triangle edges) (:131-156): every entry of A is
        // kept with its own cost, repeated (row, col) pairs included, and the
        // triangle edges absent from A are appended with cost 0; A is sorted
        // by (row, col) and the union's stable re-sort keeps repeats in A's order.
        std::vector<uint64_t> keys;
        keys.reserve(A.nnz() + tk.size());
        size_t a = 0, b = 0;
        const size_t na = A.nnz(), nt = tk.size();
        while (a < na || b < nt) {
            const uint64_t ka = a < na ? key2(A.row[a], A.col[a]) : ~0ull;
            if (a < na && (b == nt || ka <= tk[b])) {
                if (b < nt && ka == tk[b]) ++b;
                keys.push_back(ka);
                c.push_back(A.val[a]);
                ++a;
            } else {
                keys.push_back(tk[b++]);
                c.push_back(0.0f);
            }
        }
        const size_t E = keys.size();
        i.resize(E); j.resize(E);
        for (size_t e = 0; e < E; ++e) { i[e] = (int)(keys[e] >> 32); j[e] = (int)(keys[e] & 0xffffffffu); }
        // a triangle edge corresponds to the first entry carrying its key
        // (set_intersection_by_key in compute_triangle_edge_correspondence, :46-50)
        auto find = [&](uint64_t k) { return (int)(std::lower_bound(keys.begin(), keys.end(), k) - keys.begin()); };
        c12.resize(T); c13.resize(T); c23.resize(T);
        cnt.assign(E, 0);
        for (int64_t t = 0; t < T; ++t) {
            c12[t] = find(key2(t1[t], t2[t])); cnt[c12[t]]++;
            c13[t] = find(key2(t1[t], t3[t])); cnt[c13[t]]++;
            c23[t] = find(key2(t2[t], t3[t])); cnt[c23[t]]++;
        }
        soff.assign(E + 1, 0);
        for (size_t e = 0; e < E; ++e) soff[e + 1] = soff[e] + cnt[e];
        slots.resize(3 * T);
        std::vector<int64_t> cur(soff.begin(), soff.end() - 1);
        for (int64_t t = 0; t < T; ++t) {
            slots[cur[c12[t]]++] = 3 * t + 0;
            slots[cur[c13[t]]++] = 3 * t + 1;
            slots[cur[c23[t]]++] = 3 * t + 2;
        }
        t12.assign(T, 0.0f); t13.assign(T, 0.0f); t23.assign(T, 0.0f);
        delta.assign(3 * T, 0.0f);
    }

    // lower_bound = edge_lower_bound + triangle_lower_bound (:178-216), summed in double [P2]
    double lower_bound() const {
        double s = 0.0;
        for (float x : c) s += (x < 0.0f ? (double)x : 0.0);
        for (size_t t = 0; t < t12.size(); ++t) {
            const float a = t12[t], b = t13[t], d = t23[t];
            float lb = 0.0f;
            lb = std::min(lb, a + b);
            lb = std::min(lb, a + d);
            lb = std::min(lb, b + d);
            lb = std::min(lb, a + b + d);
            s += (double)lb;
        }
        return s;
    }

    // send_messages_to_triplets (:218-269)
    void send_to_triplets() {
        for (size_t t = 0; t < t12.size(); ++t) {
            t12[t] += c[c12[t]] / float(cnt[c12[t]]);
            t13[t] += c[c13[t]] / float(cnt[c13[t]]);
            t23[t] += c[c23[t]] / float(cnt[c23[t]]);
        }
        for (size_t e = 0; e < c.size(); ++e)
            if (cnt[e] > 0) c[e] = 0.0f;
    }

    static float min_marginal(float x, float y, float z) {
        const float mm1 = std::min(x + y + z, std::min(x + y, x + z));
        const float mm0 = (float)std::min(0.0, (double)(y + z));
        return mm1 - mm0;
    }

    // send_messages_to_edges / decrease_triangle_costs_func (:271-348);
    // deltas land in edge costs in (triangle, slot) order [P4]
    void send_to_edges() {
        for (size_t t = 0; t < t12.size(); ++t) {
            float a = t12[t], b = t13[t], d = t23[t];
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
            t12[t] = a; t13[t] = b; t23[t] = d;
            delta[3 * t + 0] = e12; delta[3 * t + 1] = e13; delta[3 * t + 2] = e23;
        }
        for (size_t e = 0; e < c.size(); ++e) {
            if (cnt[e] == 0) continue;
            double s = (double)c[e];
            for (int64_t k = soff[e]; k < soff[e + 1]; ++k) s += (double)delta[slots[k]];
            c[e] = (float)s;
        }
    }

    void iteration() { send_to_triplets(); send_to_edges(); }
};

// dual_update_cycle_length (src/dual_solver.cu:7-35)
std::tuple<Coo, double, int64_t> dual_update_cycle_length(const Coo& A, int c, int steps, float factor, float tol) {
    std::vector<int> a1, a2, a3;
    conflicted_cycles(A, c, factor, tol, a1, a2, a3);
    const int64_t num_tri = (int64_t)a1.size();
    if (num_tri == 0) return {A, get_lb(A.val), 0};
    MP mp(A, a1, a2, a3);
    double prev = 0.0;
    for (int it = 0; it < steps; ++it) {
        const double lb = mp.lower_bound();
        if (it > 0 && (lb - prev) < 1e-3) break;
        mp.iteration();
        prev = lb;
    }
    const double fin = mp.lower_bound();
    Coo R;
    R.directed = true;
    R.row = mp.i; R.col = mp.j; R.val = mp.c;
    coo_init(R, true);
    return {R, fin, num_tri};
}

// dual_solver (src/dual_solver.cu:37-63)
double dual_solver(Coo& A, int max_len, int num_iter, float factor, int num_outer, float tol) {
    if (max_len < 3 || num_iter == 0 || num_outer == 0) return get_lb(A.val);
    double fin = 0.0, prev = 0.0;
    int64_t num_tri = 0;
    for (int o = 0; o < num_outer; ++o) {
        for (int c = 3; c <= max_len; ++c) std::tie(A, fin, num_tri) = dual_update_cycle_length(A, c, num_iter, factor, tol);
        if (num_tri == 0 || (o > 0 && (fin - prev) < 1e-3)) break;
        prev = fin;
    }
    return fin;
}

// ------------------------------------------------- components and forest

// Connected components with canonical minimum-id labels (ECL-CC semantics,
// external/ECL-CC/ECL-CC_11.cu:315-350).
std::vector<int> connected_components(int n, const std::vector<int>& off, const std::vector<int>& adj) {
    std::vector<int> p(n);
    std::iota(p.begin(), p.end(), 0);
    auto find = [&](int x) {
        while (p[x] != x) { p[x] = p[p[x]]; x = p[x]; }
        return x;
    };
    for (int v = 0; v < n; ++v)
        for (int k = off[v]; k < off[v + 1]; ++k) {
            int a = find(v), b = find(adj[k]);
            if (a == b) continue;
            if (a < b) p[b] = a; else p[a] = b;
        }
    for (int v = 0; v < n; ++v) p[v] = find(v);
    return p;
}

// Maximum spanning forest of (r, c, w) [P5] (replaces
// external/cudaMST/gpuMST/gpuMST.cu:364-388).  Returns forest edge indices
// in ascending order (maximum_spanning_tree sorts and uniques its ids).
std::vector<int> max_spanning_forest(int n, const std::vector<int>& r, const std::vector<int>& c,
                                     const std::vector<float>& w) {
    std::vector<int> idx(r.size());
    std::iota(idx.begin(), idx.end(), 0);
    std::sort(idx.begin(), idx.end(), [&](int a, int b) { return w[a] != w[b] ? w[a] > w[b] : a < b; });
    std::vector<int> p(n);
    std::iota(p.begin(), p.end(), 0);
    auto find = [&](int x) {
        while (p[x] != x) { p[x] = p[p[x]]; x = p[x]; }
        return x;
    };
    std::vector<int> out;
    for (int e : idx) {
        int a = find(r[e]), b = find(c[e]);
        if (a == b) continue;
        p[std::max(a, b)] = std::min(a, b);
        out.push_back(e);
    }
    std::sort(out.begin(), out.end());
    return out;
}

struct Frontier {
    std::vector<int> node, parent, rep, bidx;
    std::vector<float> bval;
    // frontier(seeds) (include/edge_contractions_woc_thrust.h:11-19)
    void reset(const std::vector<int>& seeds) {
        node = seeds;
        parent.assign(seeds.size(), -1);
        rep.resize(seeds.size());
        std::iota(rep.begin(), rep.end(), 0);
        bidx.assign(seeds.size(), 0);
        bval.assign(seeds.size(), FLT_MAX);
    }
};

struct Woc {
    int num_nodes;
    std::vector<int> mr, mc, rr, rc, cc;
    std::vector<float> mw;
    std::vector<int> moff;
    Frontier rowf, colf;

    // edge_contractions_woc ctor (src/edge_contractions_woc.cu:341-383)
    explicit Woc(const Coo& A) : num_nodes(A.max_dim()) {
        cc.assign(num_nodes, 0);
        std::vector<int> pr, pc;
        std::vector<float> pw;
        for (size_t e = 0; e < A.nnz(); ++e) {
            if (A.val[e] > 0.0f) { pr.push_back(A.row[e]); pc.push_back(A.col[e]); pw.push_back(A.val[e]); }
            else { rr.push_back(A.row[e]); rc.push_back(A.col[e]); }
        }
        if (pr.empty()) { rr.clear(); rc.clear(); return; }
        const std::vector<int> f = max_spanning_forest(num_nodes, pr, pc, pw);
        const size_t k = f.size();
        mr.resize(2 * k); mc.resize(2 * k); mw.resize(2 * k);
        for (size_t q = 0; q < k; ++q) {
            mr[q] = pr[f[q]]; mc[q] = pc[f[q]]; mw[q] = pw[f[q]];
            mr[k + q] = pc[f[q]]; mc[k + q] = pr[f[q]]; mw[k + q] = pw[f[q]];
        }
        coo_sort(mr, mc, mw);
        moff = compute_offsets(mr, num_nodes - 1);
    }

    Csr forest() const { Csr G; G.off = moff; G.adj = mc; G.w = mw; return G; }

    // filter_by_cc (src/edge_contractions_woc.cu:257-276 / _thrust.cu:95-133)
    int filter_by_cc(bool reset_frontiers) {
        moff = compute_offsets(mr, num_nodes - 1);
        cc = connected_components(num_nodes, moff, mc);
        size_t w = 0;
        for (size_t e = 0; e < rr.size(); ++e)
            if (cc[rr[e]] == cc[rc[e]]) { rr[w] = rr[e]; rc[w] = rc[e]; ++w; }
        rr.resize(w); rc.resize(w);
        if (reset_frontiers) { rowf.reset(rr); colf.reset(rc); }
        return (int)w;
    }

    // Remove forest entries whose key (i,j) or (j,i) is listed (set_difference_by_key
    // after to_undirected, src/edge_contractions_woc.cu:287-339, _thrust.cu:357-403).
    void remove_keys(const std::vector<std::pair<int, int>>& rm) {
        std::unordered_set<uint64_t> s;
        for (auto& q : rm) { s.insert(key2(q.first, q.second)); s.insert(key2(q.second, q.first)); }
        size_t w = 0;
        for (size_t e = 0; e < mr.size(); ++e)
            if (!s.count(key2(mr[e], mc[e]))) { mr[w] = mr[e]; mc[w] = mc[e]; mw[w] = mw[e]; ++w; }
        mr.resize(w); mc.resize(w); mw.resize(w);
        moff = compute_offsets(mr, num_nodes - 1);
    }

    bool remove_by_mask(const std::vector<char>& valid) {
        std::vector<std::pair<int, int>> rm;
        for (size_t e = 0; e < valid.size(); ++e)
            if (!valid[e]) rm.push_back({mr[e], mc[e]});
        if (rm.empty()) return false;
        remove_keys(rm);
        return true;
    }

    // check_triangles (src/edge_contractions_woc.cu:10-37,141-161)
    bool check_triangles() {
        if (rr.empty()) return false;
        std::vector<char> valid(mw.size(), 1);
        Csr G = forest();
        for (size_t r = 0; r < rr.size(); ++r) {
            int i1 = G.off[rr[r]], i2 = G.off[rc[r]];
            const int mid = lcn(rr[r], rc[r], G, i1, i2);
            if (mid >= 0) {
                const float a = mw[i1 - 1], b = mw[i2 - 1];
                if (a < b) valid[i1 - 1] = 0; else valid[i2 - 1] = 0;
            }
        }
        return remove_by_mask(valid);
    }

    // check_quadrangles (src/edge_contractions_woc.cu:39-79,163-193)
    bool check_quadrangles() {
        if (rr.empty()) return false;
        std::vector<char> valid(mw.size(), 1);
        Csr G = forest();
        for (size_t r = 0; r < rr.size(); ++r) {
            const int v1 = rr[r], v2 = rc[r];
            for (int e1 = G.off[v1]; e1 < G.off[v1 + 1]; ++e1) {
                const int n1 = G.adj[e1];
                int i1 = G.off[n1], i2 = G.off[v2];
                const int mid = lcn(n1, v2, G, i1, i2);
                if (mid >= 0) {
                    const float a = mw[e1], b = mw[i1 - 1], c = mw[i2 - 1];
                    if (a < b && a < c) valid[e1] = 0;
                    else if (b < c) valid[i1 - 1] = 0;
                    else valid[i2 - 1] = 0;
                }
            }
        }
        return remove_by_mask(valid);
    }

    // filter_by_thresholding (src/edge_contractions_woc_thrust.cu:513-542)
    bool filter_by_thresholding(float mult) {
        bool any;
        const float thr = positive_mean_threshold(mw, mult, &any);
        size_t w = 0;
        for (size_t e = 0; e < mw.size(); ++e)
            if (!(mw[e] < thr)) { mr[w] = mr[e]; mc[w] = mc[e]; mw[w] = mw[e]; ++w; }
        if (w == mw.size()) return false;
        mr.resize(w); mc.resize(w); mw.resize(w);
        return true;
    }

    // expand_frontier + expand_cuda (src/edge_contractions_woc_thrust.cu:149-199,412-469)
    void expand(Frontier& f) {
        const std::vector<int> off = compute_offsets(mr, num_nodes - 1);
        Frontier g;
        for (size_t q = 0; q < f.node.size(); ++q) {
            const int src = f.node[q];
            for (int k = off[src]; k < off[src + 1]; ++k) {
                const int dst = mc[k];
                if (dst == f.parent[q]) continue;
                g.node.push_back(dst);
                g.rep.push_back(f.rep[q]);
                g.parent.push_back(src);
                if (mw[k] < f.bval[q]) { g.bidx.push_back(k); g.bval.push_back(mw[k]); }
                else { g.bidx.push_back(f.bidx[q]); g.bval.push_back(f.bval[q]); }
            }
        }
        f = std::move(g);
    }

    // filter_cycles (src/edge_contractions_woc_thrust.cu:273-410): entries keyed
    // (node, rep) present in both frontiers close a conflicted cycle; the
    // smaller bottleneck wins, the column side on ties (reduce_intersecting_paths).
    bool filter_cycles() {
        auto order = [](const Frontier& f) {
            std::vector<size_t> o(f.node.size());
            std::iota(o.begin(), o.end(), 0);
            std::stable_sort(o.begin(), o.end(), [&](size_t a, size_t b) {
                return f.node[a] != f.node[b] ? f.node[a] < f.node[b] : f.rep[a] < f.rep[b];
            });
            return o;
        };
        const std::vector<size_t> ro = order(rowf), co = order(colf);
        // merge, reduce groups of equal keys in merged order (row before col)
        struct Item { int node, rep, bidx; float bval; };
        std::vector<Item> merged;
        merged.reserve(ro.size() + co.size());
        size_t a = 0, b = 0;
        auto less = [](const Item& x, const Item& y) { return x.node != y.node ? x.node < y.node : x.rep < y.rep; };
        while (a < ro.size() || b < co.size()) {
            Item x{}, y{};
            if (a < ro.size()) x = {rowf.node[ro[a]], rowf.rep[ro[a]], rowf.bidx[ro[a]], rowf.bval[ro[a]]};
            if (b < co.size()) y = {colf.node[co[b]], colf.rep[co[b]], colf.bidx[co[b]], colf.bval[co[b]]};
            if (b >= co.size() || (a < ro.size() && !less(y, x))) { merged.push_back(x); ++a; }
            else { merged.push_back(y); ++b; }
        }
        std::vector<int> to_remove;
        for (size_t s = 0; s < merged.size();) {
            size_t t = s;
            int bi = merged[s].bidx;
            float bv = merged[s].bval;
            int count = 0;
            while (t < merged.size() && merged[t].node == merged[s].node && merged[t].rep == merged[s].rep) {
                if (t > s) {
                    if (!(bv < merged[t].bval)) { bi = merged[t].bidx; bv = merged[t].bval; }
                }
                ++count;
                ++t;
            }
            if (count != 1) to_remove.push_back(bi);
            s = t;
        }
        if (to_remove.empty()) return false;
        std::sort(to_remove.begin(), to_remove.end());
        to_remove.erase(std::unique(to_remove.begin(), to_remove.end()), to_remove.end());
        std::vector<std::pair<int, int>> rm;
        for (int k : to_remove) rm.push_back({mr[k], mc[k]});
        remove_keys(rm);
        return true;
    }

    // find_contraction_mapping (src/edge_contractions_woc.cu:387-423 and
    // src/edge_contractions_woc_thrust.cu:544-599)
    std::pair<std::vector<int>, int> find_contraction_mapping() {
        if (mr.empty()) return {{}, 0};
        filter_by_cc(false);
        if (check_triangles()) filter_by_cc(false);
        check_quadrangles();
        // hand-off to the frontier mapper
        if (mr.empty()) return {{}, 0};
        rowf.reset(rr);
        colf.reset(rc);
        int num_rep = filter_by_cc(true);
        int itr = 0;
        while (num_rep > 0 && !mr.empty()) {
            if (itr % 5 == 0)
                if (filter_by_thresholding(0.9f)) num_rep = filter_by_cc(true);
            if (num_rep == 0 || mr.empty()) break;
            expand(rowf);
            if (filter_cycles()) num_rep = filter_by_cc(true);
            if (num_rep == 0 || mr.empty()) break;
            expand(colf);
            if (filter_cycles()) num_rep = filter_by_cc(true);
            itr++;
        }
        std::vector<int> map = compress_label_sequence(cc, (int)cc.size() - 1);
        return {map, (int)mr.size()};
    }
};

// map_node_labels (src/rama_cuda.cu:32-52)
void map_node_labels(const std::vector<int>& cur, std::vector<int>& orig) {
    for (auto& m : orig)
        if (m < (int)cur.size()) m = cur[m];
}

// compute_sanitized_graph (include/rama_utils.h:429-466)
std::vector<int> sanitize(std::vector<int>& i, std::vector<int>& j, std::vector<float>& c) {
    for (size_t e = 0; e < i.size(); ++e) {
        int a = i[e], b = j[e];
        i[e] = std::min(a, b); j[e] = std::max(a, b);
    }
    coo_sort(i, j, c);
    size_t w = 0;
    for (size_t e = 0; e < i.size(); ++e)
        if (w == 0 || i[e] != i[w - 1] || j[e] != j[w - 1]) { i[w] = i[e]; j[w] = j[e]; c[w] = c[e]; ++w; }
    i.resize(w); j.resize(w); c.resize(w);
    int mx = 0;
    for (size_t e = 0; e < w; ++e) mx = std::max(mx, std::max(i[e], j[e]));
    std::vector<int> used(mx + 1, 0);
    for (size_t e = 0; e < w; ++e) { used[i[e]] = 1; used[j[e]] = 1; }
    std::vector<int> ids(mx + 1, 0);
    int run = 0;
    for (int k = 0; k <= mx; ++k) { ids[k] = run; run += used[k]; }
    for (size_t e = 0; e < w; ++e) { i[e] = ids[i[e]]; j[e] = ids[j[e]]; }
    return ids;
}

// desanitize_node_labels (include/rama_utils.h:468-493)
std::vector<int> desanitize(const std::vector<int>& labels, const std::vector<int>& ids) {
    std::vector<int> out(ids.size(), -1);
    for (size_t n = 0; n < ids.size(); ++n) {
        const int cur = ids[n];
        if (n == ids.size() - 1 || cur != ids[n + 1]) out[n] = labels[cur];
    }
    return out;
}

struct SolveOut {
    std::vector<int> labels;
    double lb = 0.0;
    int iterations = 0;
    std::vector<std::vector<int>> timeline;
};

// rama_cuda(dCOO&, opts) (src/rama_cuda.cu:54-164)
SolveOut solve_dcoo(Coo& A, const rama_options_c& o) {
    SolveOut out;
    auto begin = std::chrono::steady_clock::now();
    out.lb = dual_solver(A, o.max_cycle_length_lb, o.num_dual_itr_lb, o.tri_memory_factor, o.num_outer_itr_dual, 1e-4f);
    std::vector<int> labels(A.max_dim());
    std::iota(labels.begin(), labels.end(), 0);
    if (o.only_compute_lb) return out;
    bool use_mm = !(o.matching_thresh_crossover_ratio > 1.0f);
    for (size_t iter = 0; A.nnz() > 0; ++iter) {
        if (iter > 0) dual_solver(A, o.max_cycle_length_primal, o.num_dual_itr_primal, 1.0f, 1, 1e-4f);
        std::vector<int> cur;
        int n_contract = 0;
        if (use_mm) {
            Coo U = export_undirected(A);
            std::vector<int> m;
            std::tie(m, n_contract) = matching_vertex_based(U, o.mean_multiplier_mm);
            cur = compress_label_sequence(m, (int)m.size() - 1);
            if ((float)n_contract < (float)A.rows_ * o.matching_thresh_crossover_ratio) use_mm = false;
        } else {
            Woc w(A);
            std::tie(cur, n_contract) = w.find_contraction_mapping();
        }
        if (n_contract == 0) break;
        Coo B = contract(A, cur);
        if (has_bad_contractions(B, nullptr)) throw std::runtime_error("Found bad contractions");
        A = std::move(B);
        remove_diagonal(A);
        map_node_labels(cur, labels);
        out.iterations = (int)iter + 1;
        if (o.dump_timeline) out.timeline.push_back(labels);
        if (o.max_time_sec >= 0) {
            auto t = std::chrono::duration_cast<std::chrono::seconds>(std::chrono::steady_clock::now() - begin).count();
            // RAMA_CLOCK_OFFSET_SEC: whole seconds added to the elapsed time,
            // the same test hook as the GPU solver's
            if (const char* off = std::getenv("RAMA_CLOCK_OFFSET_SEC")) t += std::atol(off);
            if (t > o.max_time_sec) break;
        }
    }
    out.labels = labels;
    return out;
}

Coo from_arrays(const int32_t* r, const int32_t* c, const float* v, int64_t m, int rows, int cols) {
    Coo A;
    A.row.assign(r, r + m); A.col.assign(c, c + m); A.val.assign(v, v + m);
    A.rows_ = rows; A.cols_ = cols;
    A.directed = true;
    return A;
}

template <class F>
int guarded(F&& f) {
    try {
        g_err.clear();
        return f();
    } catch (const std::exception& e) {
        g_err = e.what();
        return std::string(e.what()) == "Found bad contractions" ? RAMA_ERR_BAD_CONTRACTION : RAMA_ERR_INVALID;
    }
}

}  // namespace

// ------------------------------------------------------------------ C ABI
extern "C" {

const char* oracle_last_error(void) { return g_err.c_str(); }

int oracle_solve(const int32_t* i, const int32_t* j, const float* c, int64_t m, const rama_options_c* o,
                 int32_t* labels_out, int64_t cap, int64_t* n_labels, double* lb_out, int32_t* iters_out) {
    return guarded([&] {
        std::vector<int> vi(i, i + m), vj(j, j + m);
        std::vector<float> vc(c, c + m);
        std::vector<int> ids;
        if (o->sanitize_graph) ids = sanitize(vi, vj, vc);
        Coo A;
        A.directed = true;
        A.row = vj; A.col = vi; A.val = vc;  // dCOO(col=i, row=j) (src/rama_cuda.cu:177)
        coo_init(A, false);
        SolveOut s = solve_dcoo(A, *o);
        std::vector<int> lab = s.labels;
        if (o->sanitize_graph && !o->only_compute_lb) lab = desanitize(lab, ids);
        *n_labels = (int64_t)lab.size();
        *lb_out = s.lb;
        if (iters_out) *iters_out = s.iterations;
        if ((int64_t)lab.size() > cap) throw std::runtime_error("label buffer too small");
        std::copy(lab.begin(), lab.end(), labels_out);
        return 0;
    });
}

int oracle_dcoo_init(const int32_t* i, const int32_t* j, const float* c, int64_t m, int32_t* out_row,
                     int32_t* out_col, float* out_val, int32_t* out_dims) {
    return guarded([&] {
        Coo A;
        A.row.assign(j, j + m); A.col.assign(i, i + m); A.val.assign(c, c + m);
        coo_init(A, false);
        std::copy(A.row.begin(), A.row.end(), out_row);
        std::copy(A.col.begin(), A.col.end(), out_col);
        std::copy(A.val.begin(), A.val.end(), out_val);
        out_dims[0] = A.rows_; out_dims[1] = A.cols_;
        return 0;
    });
}

int oracle_contract(const int32_t* row, const int32_t* col, const float* val, int64_t m, const int32_t* map,
                    int64_t map_size, int32_t* out_row, int32_t* out_col, float* out_val, int64_t* out_m,
                    int32_t* out_dims, int32_t* out_bad, double* out_diag_sum) {
    return guarded([&] {
        Coo A = from_arrays(row, col, val, m, 0, 0);
        std::vector<int> mp(map, map + map_size);
        Coo B = contract(A, mp);
        *out_bad = has_bad_contractions(B, out_diag_sum) ? 1 : 0;
        out_dims[0] = B.rows_; out_dims[1] = B.cols_;
        remove_diagonal(B);
        *out_m = (int64_t)B.nnz();
        std::copy(B.row.begin(), B.row.end(), out_row);
        std::copy(B.col.begin(), B.col.end(), out_col);
        std::copy(B.val.begin(), B.val.end(), out_val);
        return 0;
    });
}

int oracle_compute_offsets(const int32_t* ids, int64_t n, int32_t max_value, int32_t* out) {
    return guarded([&] {
        std::vector<int> v(ids, ids + n);
        std::vector<int> o = compute_offsets(v, max_value);
        std::copy(o.begin(), o.end(), out);
        return 0;
    });
}

int oracle_compress_labels(const int32_t* data, int64_t n, int32_t max_label, int32_t* out) {
    return guarded([&] {
        std::vector<int> v(data, data + n);
        std::vector<int> o = compress_label_sequence(v, max_label);
        std::copy(o.begin(), o.end(), out);
        return 0;
    });
}

int oracle_matching(const int32_t* row, const int32_t* col, const float* val, int64_t m, int32_t rows,
                    int32_t cols, float mm, int32_t* out_map, int64_t* out_map_size, int32_t* out_matched) {
    return guarded([&] {
        Coo A = from_arrays(row, col, val, m, rows, cols);
        Coo U = export_undirected(A);
        auto r = matching_vertex_based(U, mm);
        *out_map_size = (int64_t)r.first.size();
        std::copy(r.first.begin(), r.first.end(), out_map);
        *out_matched = r.second;
        return 0;
    });
}

int oracle_conflicted_cycles(const int32_t* row, const int32_t* col, const float* val, int64_t m, int32_t rows,
                             int32_t cols, int32_t max_len, float factor, float tol, int32_t* o1, int32_t* o2,
                             int32_t* o3, int64_t cap, int64_t* out_n, int64_t* out_raw) {
    return guarded([&] {
        Coo A = from_arrays(row, col, val, m, rows, cols);
        std::vector<int> a, b, c;
        conflicted_cycles(A, max_len, factor, tol, a, b, c);
        *out_raw = (int64_t)a.size();
        normalize_triangles(a, b, c);
        *out_n = (int64_t)a.size();
        if ((int64_t)a.size() > cap) throw std::runtime_error("triangle buffer too small");
        std::copy(a.begin(), a.end(), o1);
        std::copy(b.begin(), b.end(), o2);
        std::copy(c.begin(), c.end(), o3);
        return 0;
    });
}

int oracle_message_passing(const int32_t* row, const int32_t* col, const float* val, int64_t m, const int32_t* t1,
                           const int32_t* t2, const int32_t* t3, int64_t T, int32_t pre, int32_t iters,
                           int32_t* out_i, int32_t* out_j, float* out_c, int64_t cap, int64_t* out_m,
                           double* out_lb) {
    return guarded([&] {
        Coo A = from_arrays(row, col, val, m, 0, 0);
        MP mp(A, std::vector<int>(t1, t1 + T), std::vector<int>(t2, t2 + T), std::vector<int>(t3, t3 + T));
        for (int k = 0; k < pre; ++k) mp.send_to_triplets();
        for (int k = 0; k < iters; ++k) {
            out_lb[k] = mp.lower_bound();
            mp.iteration();
        }
        out_lb[iters] = mp.lower_bound();
        *out_m = (int64_t)mp.i.size();
        if ((int64_t)mp.i.size() > cap) throw std::runtime_error("edge buffer too small");
        std::copy(mp.i.begin(), mp.i.end(), out_i);
        std::copy(mp.j.begin(), mp.j.end(), out_j);
        std::copy(mp.c.begin(), mp.c.end(), out_c);
        return 0;
    });
}

int oracle_dual_solver(const int32_t* row, const int32_t* col, const float* val, int64_t m, int32_t rows,
                       int32_t cols, int32_t max_len, int32_t num_iter, float factor, int32_t num_outer, float tol,
                       int32_t* out_row, int32_t* out_col, float* out_val, int64_t cap, int64_t* out_m,
                       int32_t* out_dims, double* out_lb) {
    return guarded([&] {
        Coo A = from_arrays(row, col, val, m, rows, cols);
        *out_lb = dual_solver(A, max_len, num_iter, factor, num_outer, tol);
        *out_m = (int64_t)A.nnz();
        out_dims[0] = A.rows_; out_dims[1] = A.cols_;
        if ((int64_t)A.nnz() > cap) throw std::runtime_error("edge buffer too small");
        std::copy(A.row.begin(), A.row.end(), out_row);
        std::copy(A.col.begin(), A.col.end(), out_col);
        std::copy(A.val.begin(), A.val.end(), out_val);
        return 0;
    });
}

int oracle_connected_components(const int32_t* off, const int32_t* cols, int32_t n, int64_t nnz, int32_t* out) {
    return guarded([&] {
        std::vector<int> o(off, off + n + 1), a(cols, cols + nnz);
        std::vector<int> l = connected_components(n, o, a);
        std::copy(l.begin(), l.end(), out);
        return 0;
    });
}

int oracle_max_spanning_forest(const int32_t* row, const int32_t* col, const float* val, int64_t m,
                               int32_t* out_row, int32_t* out_col, float* out_val, int64_t* out_m) {
    return guarded([&] {
        std::vector<int> pr, pc;
        std::vector<float> pw;
        int n = 0;
        for (int64_t e = 0; e < m; ++e) {
            n = std::max(n, std::max(row[e], col[e]) + 1);
            if (val[e] > 0.0f) { pr.push_back(row[e]); pc.push_back(col[e]); pw.push_back(val[e]); }
        }
        std::vector<int> f = max_spanning_forest(n, pr, pc, pw);
        *out_m = (int64_t)f.size();
        for (size_t q = 0; q < f.size(); ++q) { out_row[q] = pr[f[q]]; out_col[q] = pc[f[q]]; out_val[q] = pw[f[q]]; }
        return 0;
    });
}

int oracle_woc_mapping(const int32_t* row, const int32_t* col, const float* val, int64_t m, int32_t rows,
                       int32_t cols, int32_t* out_map, int64_t* out_map_size, int32_t* out_count) {
    return guarded([&] {
        Coo A = from_arrays(row, col, val, m, rows, cols);
        Woc w(A);
        auto r = w.find_contraction_mapping();
        *out_map_size = (int64_t)r.first.size();
        std::copy(r.first.begin(), r.first.end(), out_map);
        *out_count = r.second;
        return 0;
    });
}

double oracle_objective(const int32_t* labels, int64_t n, const int32_t* i, const int32_t* j, const float* c,
                        int64_t m) {
    double s = 0.0;
    for (int64_t e = 0; e < m; ++e)
        if (labels[i[e]] != labels[j[e]]) s += c[e];
    (void)n;
    return s;
}

}  // extern "C"
