// capi.hip — the extern "C" boundary of librama_hip.so (include/rama_hip.h).
// Converts exceptions to status codes + a thread-local message; owns H2D/D2H
// copies for the host-buffer entry points.
#include <chrono>
#include <cmath>
#include <memory>

#include "rama_internal.hpp"

struct rama_timeline {
    std::vector<std::vector<int>> snaps;
};

namespace {

thread_local std::string g_last_error;

template <class F>
int guarded(F&& f) {
    try {
        g_last_error.clear();
        f();
        return RAMA_OK;
    } catch (const rama::Error& e) {
        g_last_error = e.what();
        return e.status;
    } catch (const std::bad_alloc&) {
        g_last_error = "host allocation failed";
        return RAMA_ERR_OOM;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return RAMA_ERR_INTERNAL;
    }
}

using namespace rama;

int pick_device() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        (void)hipGetLastError();
        throw Error(RAMA_ERR_NO_DEVICE, "no HIP device available");
    }
    return 0;  // the reference always uses the first visible device (include/rama_utils.h:15-18)
}

template <class T>
DBuf<T> upload(Ctx& ctx, const T* h, i64 n) {
    DBuf<T> d = ctx.buf<T>(n);
    ctx.h2d(d.get(), h, n);
    return d;
}

template <class T>
void download(Ctx& ctx, T* h, const T* d, i64 n) {
    ctx.d2h(h, d, n);
}

}  // namespace

// The device memory of rama_solve_device_noalloc: one copy per GPU, held
// from the code object's load until the process ends.  In host code the
// name stands for a host-side shadow, not a device address; only the address
// from hipGetSymbolAddress (noalloc_scratch) is handed to kernels and copies.
static __device__ __align__(256) unsigned char scratch[RAMA_NOALLOC_SCRATCH_BYTES];

namespace {

std::mutex g_scratch_mu;
std::map<int, void*> g_scratch_addr;  // device -> its copy of `scratch`

// Device address of `scratch` on `device`, which must be current: looked up
// on the first call for the device, then reused.
void* noalloc_scratch(int device) {
    std::lock_guard<std::mutex> l(g_scratch_mu);
    auto it = g_scratch_addr.find(device);
    if (it != g_scratch_addr.end()) return it->second;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
        (void)hipGetLastError();
        throw Error(RAMA_ERR_HIP, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    }
    g_scratch_addr.emplace(device, p);
    return p;
}

}  // namespace

namespace rama {

void set_last_error(const std::string& s) { g_last_error = s; }

static void solve_graph(Ctx& ctx, DCoo& A, const Options& o, const Sanitized& san, DBuf<int>& labels, i64& n_labels,
                        double* lb_out, int32_t* time_ms, rama_timeline** tl);

// Shared body of the solve entry points; `i`, `j`, `c` are device buffers
// owned by this call.
void solve_common(Ctx& ctx, DBuf<int>& i, DBuf<int>& j, DBuf<float>& c, i64 m, const rama_options_c* copts,
                  DBuf<int>& labels, i64& n_labels, double* lb_out, int32_t* time_ms, rama_timeline** tl) {
    RAMA_REQUIRE(copts != nullptr, "options must not be NULL");
    RAMA_REQUIRE(m > 0, "the graph has no edges");
    const Options o = to_options(*copts);
    Sanitized san;
    if (o.sanitize_graph) san = sanitize_graph(ctx, i, j, c, m);
    DCoo A = dcoo_from_edges(ctx, j.get(), i.get(), c.get(), m);  // dCOO(col=i, row=j) (src/rama_cuda.cu:177)
    i.reset();
    j.reset();
    c.reset();
    solve_graph(ctx, A, o, san, labels, n_labels, lb_out, time_ms, tl);
}

// The device entry points: the caller's arrays are read in place (building
// the graph only reads them); sanitising edits its input, so it works on
// copies.
static void solve_view(Ctx& ctx, const int32_t* d_i, const int32_t* d_j, const float* d_c, i64 m,
                       const rama_options_c* copts, DBuf<int>& labels, i64& n_labels, double* lb_out,
                       int32_t* time_ms, rama_timeline** tl) {
    RAMA_REQUIRE(copts != nullptr, "options must not be NULL");
    RAMA_REQUIRE(m > 0, "the graph has no edges");
    if (copts->sanitize_graph) {
        DBuf<int> di = ctx.buf<int>(m), dj = ctx.buf<int>(m);
        DBuf<float> dc = ctx.buf<float>(m);
        ctx.d2d(di.get(), d_i, m);
        ctx.d2d(dj.get(), d_j, m);
        ctx.d2d(dc.get(), d_c, m);
        solve_common(ctx, di, dj, dc, m, copts, labels, n_labels, lb_out, time_ms, tl);
        return;
    }
    const Options o = to_options(*copts);
    // the host entry's argument order: dcoo_from_edges orients every edge as
    // (min, max), so the reference's dCOO(col=j, row=i) (src/rama_cuda.cu:202)
    // is the same graph
    DCoo A = dcoo_from_edges(ctx, d_j, d_i, d_c, m);
    solve_graph(ctx, A, o, Sanitized{}, labels, n_labels, lb_out, time_ms, tl);
}

static void solve_graph(Ctx& ctx, DCoo& A, const Options& o, const Sanitized& san, DBuf<int>& labels, i64& n_labels,
                        double* lb_out, int32_t* time_ms, rama_timeline** tl) {
    const auto t0 = std::chrono::steady_clock::now();
    SolveResult r = solve(ctx, A, o);
    ctx.sync();
    const auto t1 = std::chrono::steady_clock::now();
    if (time_ms) *time_ms = (int32_t)std::chrono::duration_cast<std::chrono::milliseconds>(t1 - t0).count();
    *lb_out = r.lb;
    if (o.sanitize_graph && !o.only_compute_lb) {
        labels = ctx.buf<int>(san.size);
        desanitize_labels(ctx, san, r.labels.get(), labels.get());
        n_labels = san.size;
    } else {
        labels = std::move(r.labels);
        n_labels = r.num_labels;
    }
    // the caller hands the timeline out once every later check has passed
    if (tl) *tl = new rama_timeline{std::move(r.timeline)};
}

}  // namespace rama

extern "C" {

const char* rama_last_error(void) { return g_last_error.c_str(); }

const char* rama_version(void) { return "rama_amd 0.1.0 (gfx950)"; }

void rama_options_default(rama_options_c* o) {
    if (!o) return;
    *o = rama_options_c{5, 10, 3, 5, 1, 0.05f, 0.05f, 2.0f, 0, -1, 0, 1, 0};
}

int rama_options_preset(const char* name, rama_options_c* o) {
    if (!o) return RAMA_ERR_INVALID;
    rama_options_default(o);
    const std::string s = name ? name : "";
    if (s == "PD") return 0;
    if (s == "P") {
        o->max_cycle_length_lb = 0;
        o->num_dual_itr_lb = 0;
        o->max_cycle_length_primal = 0;
        o->num_dual_itr_primal = 0;
        return 0;
    }
    if (s == "PD+") {
        o->max_cycle_length_lb = 5;
        o->num_dual_itr_lb = 10;
        o->max_cycle_length_primal = 5;
        o->num_dual_itr_primal = 10;
        return 0;
    }
    if (s == "D") {
        o->max_cycle_length_lb = 5;
        o->num_dual_itr_lb = 10;
        o->num_outer_itr_dual = 5;
        o->only_compute_lb = 1;
        return 0;
    }
    return 1;  // unknown: PD defaults, as the reference's never-thrown error (:51-52)
}

int rama_solve_host(const int32_t* i, const int32_t* j, const float* costs, int64_t m, const rama_options_c* opts,
                    int32_t* labels_out, int64_t cap, int64_t* n_labels_out, double* lb_out, int32_t* time_ms_out,
                    rama_timeline** tl) {
    return guarded([&] {
        RAMA_REQUIRE(i && j && costs && n_labels_out && lb_out, "NULL argument");
        if (tl) *tl = nullptr;
        Ctx ctx(pick_device(), nullptr, opts && opts->verbose);
        DBuf<int> di = upload(ctx, i, m), dj = upload(ctx, j, m);
        DBuf<float> dc = upload(ctx, costs, m);
        DBuf<int> labels;
        i64 n = 0;
        rama_timeline* t = nullptr;
        std::unique_ptr<rama_timeline> own;
        solve_common(ctx, di, dj, dc, m, opts, labels, n, lb_out, time_ms_out, tl ? &t : nullptr);
        own.reset(t);
        *n_labels_out = n;
        if (n > cap) throw Error(RAMA_ERR_INVALID, "label buffer too small: need " + std::to_string(n));
        if (n > 0) download(ctx, labels_out, labels.get(), n);
        if (tl) *tl = own.release();
    });
}

int rama_solve_device(const int32_t* d_i, const int32_t* d_j, const float* d_c, int64_t m, const rama_options_c* opts,
                      int device, void* stream, int32_t* d_labels_out, int64_t cap, int64_t* n_labels_out,
                      double* lb_out, int32_t* time_ms_out, rama_timeline** tl) {
    return guarded([&] {
        RAMA_REQUIRE(d_i && d_j && d_c && n_labels_out && lb_out, "NULL argument");
        if (tl) *tl = nullptr;
        Ctx ctx(device, static_cast<hipStream_t>(stream), opts && opts->verbose);
        DBuf<int> labels;
        i64 n = 0;
        rama_timeline* t = nullptr;
        std::unique_ptr<rama_timeline> own;
        solve_view(ctx, d_i, d_j, d_c, m, opts, labels, n, lb_out, time_ms_out, tl ? &t : nullptr);
        own.reset(t);
        *n_labels_out = n;
        if (n > cap) throw Error(RAMA_ERR_INVALID, "label buffer too small: need " + std::to_string(n));
        if (n > 0) ctx.d2d(d_labels_out, labels.get(), n);
        ctx.sync();
        if (tl) *tl = own.release();
    });
}

int rama_solve_device_noalloc(const int32_t* d_i, const int32_t* d_j, const float* d_c, int64_t m,
                              const rama_options_c* opts, int device, void* stream, int32_t* d_labels_out,
                              int64_t cap, int64_t* n_labels_out, double* lb_out, int32_t* time_ms_out,
                              int64_t* scratch_peak_out) {
    return guarded([&] {
        RAMA_REQUIRE(d_i && d_j && d_c && n_labels_out && lb_out, "NULL argument");
        RAMA_REQUIRE(m <= RAMA_NOALLOC_MAX_EDGES, "rama_solve_device_noalloc: " + std::to_string(m) +
                                                      " edges exceed the limit of " +
                                                      std::to_string((long long)RAMA_NOALLOC_MAX_EDGES));
        RAMA_HIP(hipSetDevice(device));
        void* region = noalloc_scratch(device);
        Ctx ctx(device, static_cast<hipStream_t>(stream), opts && opts->verbose, region,
                (size_t)RAMA_NOALLOC_SCRATCH_BYTES);
        DBuf<int> labels;
        i64 n = 0;
        solve_view(ctx, d_i, d_j, d_c, m, opts, labels, n, lb_out, time_ms_out, nullptr);
        *n_labels_out = n;
        if (n > cap) throw Error(RAMA_ERR_INVALID, "label buffer too small: need " + std::to_string(n));
        if (n > 0) ctx.d2d(d_labels_out, labels.get(), n);
        ctx.sync();
        if (scratch_peak_out) *scratch_peak_out = (int64_t)ctx.arena.peak();
    });
}

int64_t rama_timeline_count(const rama_timeline* t) { return t ? (int64_t)t->snaps.size() : 0; }
int64_t rama_timeline_length(const rama_timeline* t, int64_t k) {
    return (t && k >= 0 && k < (int64_t)t->snaps.size()) ? (int64_t)t->snaps[k].size() : 0;
}
const int32_t* rama_timeline_labels(const rama_timeline* t, int64_t k) {
    return (t && k >= 0 && k < (int64_t)t->snaps.size()) ? t->snaps[k].data() : nullptr;
}
void rama_timeline_free(rama_timeline* t) { delete t; }

double rama_objective(const int32_t* labels, int64_t n, const int32_t* i, const int32_t* j, const float* c,
                      int64_t m) {
    double s = 0.0;
    for (int64_t e = 0; e < m; ++e) {
        const int32_t a = i[e], b = j[e];
        if (a < 0 || b < 0 || a >= n || b >= n) return std::nan("");
        if (labels[a] != labels[b]) s += c[e];
    }
    return s;
}

// ------------------------------------------------------------ stage entries

int rama_stage_dcoo_init(const int32_t* i, const int32_t* j, const float* c, int64_t m, int32_t* out_row,
                         int32_t* out_col, float* out_val, int32_t* out_dims) {
    return guarded([&] {
        Ctx ctx(pick_device(), nullptr, false);
        DBuf<int> di = upload(ctx, i, m), dj = upload(ctx, j, m);
        DBuf<float> dc = upload(ctx, c, m);
        DCoo A = dcoo_from_edges(ctx, dj.get(), di.get(), dc.get(), m);
        download(ctx, out_row, A.row.get(), m);
        download(ctx, out_col, A.col.get(), m);
        download(ctx, out_val, A.val.get(), m);
        out_dims[0] = A.rows;
        out_dims[1] = A.cols;
    });
}

static DCoo upload_dcoo(Ctx& ctx, const int32_t* r, const int32_t* c, const float* v, i64 m, int rows, int cols) {
    DCoo A;
    A.row = upload(ctx, r, m);
    A.col = upload(ctx, c, m);
    A.val = upload(ctx, v, m);
    A.nnz = m;
    A.rows = rows;
    A.cols = cols;
    return A;
}

int rama_stage_contract(const int32_t* row, const int32_t* col, const float* val, int64_t m, const int32_t* node_map,
                        int64_t map_size, int32_t* out_row, int32_t* out_col, float* out_val, int64_t* out_m,
                        int32_t* out_dims, int32_t* out_bad, double* out_diag_sum) {
    return guarded([&] {
        Ctx ctx(pick_device(), nullptr, false);
        DCoo A = upload_dcoo(ctx, row, col, val, m, 0, 0);
        DBuf<int> mp = upload(ctx, node_map, map_size);
        int mx = 0;
        for (int64_t k = 0; k < map_size; ++k) mx = std::max(mx, node_map[k]);
        bool bad = false;
        DCoo B = contract(ctx, A, mp.get(), mx, &bad, out_diag_sum);
        *out_m = B.nnz;
        *out_bad = bad ? 1 : 0;
        out_dims[0] = B.rows;
        out_dims[1] = B.cols;
        download(ctx, out_row, B.row.get(), B.nnz);
        download(ctx, out_col, B.col.get(), B.nnz);
        download(ctx, out_val, B.val.get(), B.nnz);
    });
}

int rama_stage_contract_symmetric(const int32_t* row, const int32_t* col, const float* val, int64_t m,
                                  const int32_t* node_map, int64_t map_size, int32_t* out_n, int32_t* out_off,
                                  int64_t off_cap, int32_t* out_adj, float* out_w, int64_t cap, int64_t* out_nnz) {
    return guarded([&] {
        if (m < 0 || map_size < 0) throw Error(RAMA_ERR_INVALID, "bad size");
        Ctx ctx(pick_device(), nullptr, false);
        DCoo A = upload_dcoo(ctx, row, col, val, m, 0, 0);
        DBuf<int> mp = upload(ctx, node_map, map_size);
        int mx = 0;
        for (int64_t k = 0; k < map_size; ++k) mx = std::max(mx, node_map[k]);
        bool bad = false;
        DCoo B = contract(ctx, A, mp.get(), mx, &bad, nullptr);
        const int n = B.max_dim();
        // the CSR the contraction built, else the one the cycle search builds
        Csr G = B.sym ? std::move(*B.sym)
                      : symmetric_csr(ctx, B.row.get(), B.col.get(), B.val.get(), B.nnz, n, false, 0.0f);
        *out_n = n;
        *out_nnz = G.nnz;
        if ((int64_t)n + 1 > off_cap || G.nnz > cap) throw Error(RAMA_ERR_INVALID, "output buffer too small");
        download(ctx, out_off, G.off.get(), (i64)n + 1);
        download(ctx, out_adj, G.adj.get(), G.nnz);
        download(ctx, out_w, G.w.get(), G.nnz);
    });
}

int rama_stage_compute_offsets(const int32_t* ids, int64_t n, int32_t max_value, int32_t* out) {
    return guarded([&] {
        Ctx ctx(pick_device(), nullptr, false);
        DBuf<int> d = upload(ctx, ids, n);
        DBuf<int> o = ctx.buf<int>((i64)max_value + 2);
        compute_offsets(ctx, d.get(), n, max_value, o.get());
        download(ctx, out, o.get(), (i64)max_value + 2);
    });
}

int rama_stage_symmetric_csr(const int32_t* row, const int32_t* col, const float* val, int64_t m,
                             int32_t n_nodes, int32_t mode, float keep_above, int32_t* out_off,
                             int32_t* out_adj, float* out_w, int64_t cap, int64_t* out_nnz) {
    return guarded([&] {
        if (m < 0 || n_nodes < 0 || mode < 0 || mode > 2) throw Error(RAMA_ERR_INVALID, "bad size or mode");
        Ctx ctx(pick_device(), nullptr, false);
        DBuf<int> r = upload(ctx, row, m), c = upload(ctx, col, m);
        DBuf<float> v = upload(ctx, val, m);
        Csr G = mode == 2 ? symmetric_csr_forest(ctx, r.get(), c.get(), v.get(), m, n_nodes)
                          : symmetric_csr(ctx, r.get(), c.get(), v.get(), m, n_nodes, mode == 1, keep_above);
        *out_nnz = G.nnz;
        if (G.nnz > cap) throw Error(RAMA_ERR_INVALID, "adjacency buffer too small");
        download(ctx, out_off, G.off.get(), (i64)n_nodes + 1);
        download(ctx, out_adj, G.adj.get(), G.nnz);
        download(ctx, out_w, G.w.get(), G.nnz);
    });
}

int rama_stage_compress_labels(const int32_t* data, int64_t n, int32_t max_label, int32_t* out) {
    return guarded([&] {
        Ctx ctx(pick_device(), nullptr, false);
        DBuf<int> d = upload(ctx, data, n);
        DBuf<int> o = ctx.buf<int>(n);
        compress_labels(ctx, d.get(), n, max_label, o.get());
        download(ctx, out, o.get(), n);
    });
}

int rama_stage_matching(const int32_t* row, const int32_t* col, const float* val, int64_t m, int32_t rows,
                        int32_t cols, float mm, int32_t* out_map, int64_t* out_map_size, int32_t* out_matched) {
    return guarded([&] {
        Ctx ctx(pick_device(), nullptr, false);
        DCoo A = upload_dcoo(ctx, row, col, val, m, rows, cols);
        const int* C = A.col.get();
        const int Un = reduce<int>(ctx, [=] __device__(i64 e) { return C[e]; }, A.nnz, -1, OpMax()) + 1;
        Csr U = symmetric_csr(ctx, A.row.get(), A.col.get(), A.val.get(), A.nnz, Un, false, 0.0f);
        DBuf<int> mp;
        *out_matched = matching_vertex_based(ctx, U, mm, mp);
        *out_map_size = Un;
        download(ctx, out_map, mp.get(), Un);
    });
}

int rama_stage_conflicted_cycles(const int32_t* row, const int32_t* col, const float* val, int64_t m, int32_t rows,
                                 int32_t cols, int32_t max_len, float factor, float tol, int32_t* o1, int32_t* o2,
                                 int32_t* o3, int64_t cap, int64_t* out_n, int64_t* out_raw) {
    return guarded([&] {
        Ctx ctx(pick_device(), nullptr, false);
        DCoo A = upload_dcoo(ctx, row, col, val, m, rows, cols);
        Triangles T = conflicted_cycles(ctx, A, max_len, factor, tol);
        *out_n = T.n;
        *out_raw = T.raw;
        if (T.n > cap) throw Error(RAMA_ERR_INVALID, "triangle buffer too small");
        download(ctx, o1, T.t1.get(), T.n);
        download(ctx, o2, T.t2.get(), T.n);
        download(ctx, o3, T.t3.get(), T.n);
    });
}

int rama_stage_message_passing(const int32_t* row, const int32_t* col, const float* val, int64_t m, const int32_t* t1,
                               const int32_t* t2, const int32_t* t3, int64_t T, int32_t pre, int32_t iters,
                               int32_t* out_i, int32_t* out_j, float* out_c, int64_t cap, int64_t* out_m,
                               double* out_lb) {
    return guarded([&] {
        Ctx ctx(pick_device(), nullptr, false);
        DCoo A = upload_dcoo(ctx, row, col, val, m, 0, 0);
        Triangles tri;
        tri.t1 = upload(ctx, t1, T);
        tri.t2 = upload(ctx, t2, T);
        tri.t3 = upload(ctx, t3, T);
        tri.n = T;
        tri.raw = T;
        normalize_triangles(ctx, tri);
        MPState S;
        mp_build(ctx, A, tri, S);
        for (int k = 0; k < pre; ++k) mp_send_to_triplets(ctx, S);
        for (int k = 0; k < iters; ++k) {
            out_lb[k] = mp_lower_bound(ctx, S);
            mp_iteration(ctx, S);
        }
        out_lb[iters] = mp_lower_bound(ctx, S);
        *out_m = S.E;
        if (S.E > cap) throw Error(RAMA_ERR_INVALID, "edge buffer too small");
        download(ctx, out_i, S.ei.get(), S.E);
        download(ctx, out_j, S.ej.get(), S.E);
        download(ctx, out_c, S.ec.get(), S.E);
    });
}

int rama_stage_dual_solver(const int32_t* row, const int32_t* col, const float* val, int64_t m, int32_t rows,
                           int32_t cols, int32_t max_len, int32_t num_iter, float factor, int32_t num_outer, float tol,
                           int32_t* out_row, int32_t* out_col, float* out_val, int64_t cap, int64_t* out_m,
                           int32_t* out_dims, double* out_lb) {
    return guarded([&] {
        Ctx ctx(pick_device(), nullptr, false);
        DCoo A = upload_dcoo(ctx, row, col, val, m, rows, cols);
        *out_lb = dual_solver(ctx, A, max_len, num_iter, factor, num_outer, tol);
        *out_m = A.nnz;
        out_dims[0] = A.rows;
        out_dims[1] = A.cols;
        if (A.nnz > cap) throw Error(RAMA_ERR_INVALID, "edge buffer too small");
        download(ctx, out_row, A.row.get(), A.nnz);
        download(ctx, out_col, A.col.get(), A.nnz);
        download(ctx, out_val, A.val.get(), A.nnz);
    });
}

int rama_stage_connected_components(const int32_t* off, const int32_t* cols, int32_t n, int64_t nnz,
                                    int32_t* out_labels) {
    return guarded([&] {
        Ctx ctx(pick_device(), nullptr, false);
        DBuf<int> o = upload(ctx, off, (i64)n + 1), a = upload(ctx, cols, nnz);
        DBuf<int> l = ctx.buf<int>(n);
        connected_components(ctx, o.get(), a.get(), n, l.get());
        download(ctx, out_labels, l.get(), n);
    });
}

int rama_stage_max_spanning_forest(const int32_t* row, const int32_t* col, const float* val, int64_t m,
                                   int32_t* out_row, int32_t* out_col, float* out_val, int64_t* out_m) {
    return guarded([&] {
        Ctx ctx(pick_device(), nullptr, false);
        int n = 0;
        std::vector<int> pr, pc;
        std::vector<float> pw;
        for (int64_t e = 0; e < m; ++e) {
            n = std::max(n, std::max(row[e], col[e]) + 1);
            if (val[e] > 0.0f) {
                pr.push_back(row[e]);
                pc.push_back(col[e]);
                pw.push_back(val[e]);
            }
        }
        const i64 P = (i64)pr.size();
        DBuf<int> dr = upload(ctx, pr.data(), P), dc = upload(ctx, pc.data(), P);
        DBuf<float> dw = upload(ctx, pw.data(), P);
        DBuf<int> ids;
        const i64 k = max_spanning_forest(ctx, dr.get(), dc.get(), dw.get(), P, n, ids);
        std::vector<int> h(k);
        download(ctx, h.data(), ids.get(), k);
        for (i64 q = 0; q < k; ++q) {
            out_row[q] = pr[h[q]];
            out_col[q] = pc[h[q]];
            out_val[q] = pw[h[q]];
        }
        *out_m = k;
    });
}

int rama_stage_woc_mapping(const int32_t* row, const int32_t* col, const float* val, int64_t m, int32_t rows,
                           int32_t cols, int32_t* out_map, int64_t* out_map_size, int32_t* out_count) {
    return guarded([&] {
        Ctx ctx(pick_device(), nullptr, false);
        DCoo A = upload_dcoo(ctx, row, col, val, m, rows, cols);
        DBuf<int> mp;
        i64 sz = 0;
        *out_count = woc_mapping(ctx, A, mp, sz);
        *out_map_size = sz;
        download(ctx, out_map, mp.get(), sz);
    });
}

}  // extern "C"
