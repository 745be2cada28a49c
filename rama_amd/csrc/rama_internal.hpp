// rama_internal.hpp — device-resident data structures and the stage
// functions that make up the solver.  All buffers live in HBM on the
// context's device; host code only sees sizes and a handful of scalars.
#pragma once

#include <chrono>
#include <memory>

#include "prims.cuh"

namespace rama {

struct Csr;

// Directed multicut graph in COO form, sorted by (row, col) with row < col —
// the reference's dCOO (include/dCOO.h:13-71).  rows/cols follow the
// reference's rows_/cols_ bookkeeping exactly because the driver's decisions
// (matching->MST crossover, label-map sizes) depend on them.
struct DCoo {
    DBuf<int> row, col;
    DBuf<float> val;
    i64 nnz = 0;
    int rows = 0, cols = 0;
    // largest node id that has an edge (-1 if none): export_undirected's
    // node count is max_node + 1, which can be below cols once self-loops
    // have been removed
    int max_node = -1;
    int max_dim() const { return rows > cols ? rows : cols; }
    // Symmetric CSR of exactly these arrays (every edge, max_dim() rows),
    // when a stage built it for a later one to reuse.
    mutable std::shared_ptr<Csr> sym;
};

// Symmetric adjacency in CSR form (both directions, sorted by (row, col)).
struct Csr {
    DBuf<int> off;  // n + 1
    DBuf<int> adj;  // nnz
    DBuf<float> w;  // nnz
    int n = 0;
    i64 nnz = 0;
};

// Solver options as seen by the device code.
struct Options {
    int max_cycle_length_lb = 5;
    int num_dual_itr_lb = 10;
    int max_cycle_length_primal = 3;
    int num_dual_itr_primal = 5;
    int num_outer_itr_dual = 1;
    float mean_multiplier_mm = 0.05f;
    float matching_thresh_crossover_ratio = 0.05f;
    float tri_memory_factor = 2.0f;
    bool only_compute_lb = false;
    int max_time_sec = -1;
    bool dump_timeline = false;
    bool verbose = true;
    bool sanitize_graph = false;
};
Options to_options(const rama_options_c& o);

// ---- dcoo.hip
DCoo dcoo_from_edges(Ctx& ctx, const int* row_in, const int* col_in, const float* val_in, i64 m);
DCoo dcoo_from_sorted(Ctx& ctx, DBuf<int>&& row, DBuf<int>&& col, DBuf<float>&& val, i64 m);
DCoo contract(Ctx& ctx, const DCoo& A, const int* node_map, int max_label, bool* bad, double* diag_sum);
// `known_count`: number of edges passing the filter if the caller knows it (-1: count here).
Csr symmetric_csr(Ctx& ctx, const int* row, const int* col, const float* val, i64 m, int n_nodes,
                  bool filter, float keep_above, i64 known_count = -1);
// symmetric_csr of every edge of a spanning forest (row offsets found per row:
// its rows leave long runs of empty rows).
Csr symmetric_csr_forest(Ctx& ctx, const int* row, const int* col, const float* val, i64 m, int n_nodes);
// The entries of F with weight > keep_above (kept of them), row order kept.
Csr csr_filter(Ctx& ctx, const Csr& F, float keep_above, i64 kept);
Csr symmetric_csr_slots(Ctx& ctx, const int* row, const int* col, const float* val, i64 m, int n_nodes,
                        const int* nslot, const float* keep_above, i64 known_count);
void compute_offsets(Ctx& ctx, const int* sorted_ids, i64 n, int max_value, int* out);
// Returns the number of distinct labels if want_count (one read-back), else -1.
int compress_labels(Ctx& ctx, const int* data, i64 n, int max_label, int* out, bool want_count = true);
void compress_matching(Ctx& ctx, const int* map, i64 n, int* out);
void map_node_labels(Ctx& ctx, const int* cur, i64 cur_size, int* orig, i64 orig_size);
double dcoo_lb(Ctx& ctx, const float* val, i64 n);
float dcoo_min(Ctx& ctx, const float* val, i64 n);
int max_id(Ctx& ctx, const int* a, const int* b, i64 n);
struct Sanitized {
    DBuf<int> ids;  // original id -> compact id (exclusive scan of "used")
    i64 size = 0;
};
Sanitized sanitize_graph(Ctx& ctx, DBuf<int>& i, DBuf<int>& j, DBuf<float>& c, i64& m);
void desanitize_labels(Ctx& ctx, const Sanitized& s, const int* labels, int* out);

// Per-instance statistics of packed instances (see dcoo.hip).
struct SlotStats {
    int max_col;       // largest local col (diagonal included), -1 if none
    int max_row;       // largest local row (= last row), -1 if none
    int max_kept_col;  // largest local col of a kept edge, -1 if none
    int kept;          // kept edges
    int bad;           // a negative self-loop was contracted
};
void slot_stats_clear(Ctx& ctx, SlotStats* st, int Q);
void slot_stats_edges(Ctx& ctx, const int* row, const int* col, i64 n, const int* nbase, int Q, SlotStats* st);
void contract_packed(Ctx& ctx, const DCoo& A, const int* node_map, int max_label, const int* new_nslot,
                     const int* new_nbase, int Q, DCoo& B, SlotStats* st);

// ---- matching.hip
// export_undirected + vertex matching; returns the (uncompressed) mapping of
// size U.n and the number of matched vertices.
int matching_vertex_based(Ctx& ctx, const Csr& U, float mean_multiplier, DBuf<int>& map);
void matching_packed(Ctx& ctx, const Csr& U, const int* nslot, int Q, const i64* hbeg, const i64* dbeg,
                     float mean_multiplier, DBuf<int>& map, int* matched_out);

// ---- cycles.hip
struct Triangles {
    DBuf<int> t1, t2, t3;
    i64 n = 0;    // after normalisation
    i64 raw = 0;  // slots filled by the search (reference triangles_v1.size())
    int max_id = -1;  // known bound on the vertex ids (-1: unknown, scan them)
};
Triangles conflicted_cycles(Ctx& ctx, const DCoo& A, int max_cycle_length, float tri_memory_factor, float tol_ratio);
template <int S>
void run_stage_packed(Ctx& ctx, const Csr& G, const int* rv1, const int* rv2, i64 R, const int* nslot, int Q,
                      const i64* d_base, const i64* d_budget, const i64* d_out, int* t1, int* t2, int* t3,
                      i64* h_total);
void normalize_triangles(Ctx& ctx, Triangles& T);
// On the first `valid` slots: drop triangles with a repeated vertex, sort,
// deduplicate (in place); returns the count.  Ids are below n_nodes.
i64 rearrange_triangles(Ctx& ctx, Triangles& T, i64 valid, int n_nodes);

// The message-passing edge sweep gives each thread two edges (e and
// e + kBlock of its block's kMpEdgesPerBlock), so both edges' loads are in
// flight together; its lower-bound partials are per block of that many
// edges, in the single and the packed launches alike.
constexpr int kMpEdgesPerBlock = 2 * kBlock;
inline unsigned mp_edge_blocks(i64 E) { return grid_for(E, kMpEdgesPerBlock); }

// ---- mp.hip
struct MessagePassing;  // opaque
struct MPState {
    DBuf<int> ei, ej;            // message-passing edges (sorted, unique)
    DBuf<float> ec;              // edge costs
    DBuf<float> share;           // per edge in a triangle (k): its cost / #triangles, what each triangle absorbs
    DBuf<unsigned char> inact;   // per edge: 1 when it is in no triangle (its cost never changes)
    DBuf<u32> pos;               // slot 3t+k -> its place in delta (runs by edge, (triangle, slot) order)
    DBuf<int> c12, c13, c23;     // triangle -> number k of its edge among the edges in a triangle
    DBuf<float> t12, t13, t23;   // triangle costs
    DBuf<float> delta;           // per-slot messages (3T), edge-major
    DBuf<double> lbparts;        // partial lower bounds
    DBuf<int> dims;              // [0] max col of (ei, ej), on device (from A.max_node, A's largest col, and the chords)
    DBuf<int> act;               // the edges in at least one triangle, ascending (Ea of them): act[k] = edge index
    DBuf<int> hq;                // Ea + 1: edge k's messages are delta[hq[k] .. hq[k+1])
    DBuf<float> eca;             // Ea: cost of edge k while the iterations run (stored into ec by mp_store_costs)
    int max_col = -1, last_row = -1;  // host copies once read back (-1: not yet)
    i64 E = 0, T = 0, Ea = 0;
};
void mp_build(Ctx& ctx, const DCoo& A, Triangles& tri, MPState& S);
// Packed instances in one message-passing state (see mp_run_packed).
struct MPPack {
    int Q;
    const i64* tbeg;  // Q+1, device
    const i64* ebeg;  // Q+1, device: edges of each instance
    const i64* abeg;  // Q+1, device: its edges in a triangle, as positions in MPState::act
    const int* tblk;  // Q+1, device: prefix of per-instance triangle blocks
    const int* eblk;  // Q+1, device: prefix of per-instance edge blocks (over its edges in a triangle)
    const i64* pbeg;  // Q, device: start of each instance's partials
};
// abeg[q] = position in S.act of the first edge of instance q with an
// index of at least ebeg[q] (ebeg: Q+1 device values).
void mp_active_ranges(Ctx& ctx, const MPState& S, const i64* ebeg, int Q, i64* abeg);
// Runs `steps` iterations for every instance with d_on[q] != 0; lb_out[q]
// (host) receives each instance's final bound.  seg_*beg_host are the host
// copies of tbeg/ebeg.
void mp_run_packed(Ctx& ctx, MPState& S, const MPPack& pk, int total_tblocks, int total_eblocks, int steps,
                   const i64* seg_tbeg_host, const i64* seg_ebeg_host, const int* d_on,
                   double* lb_out);
// ec[act[k]] = eca[k]: the iterations update the costs of the edges in a
// triangle in eca; ec receives them before it is read.
void mp_store_costs(Ctx& ctx, MPState& S);
double mp_lower_bound(Ctx& ctx, MPState& S);
double mp_fused_lower_bound(Ctx& ctx, MPState& S);
void mp_send_to_triplets(Ctx& ctx, MPState& S);
void mp_iteration(Ctx& ctx, MPState& S);
double mp_run(Ctx& ctx, MPState& S, int steps);
DCoo mp_reparametrized(Ctx& ctx, MPState& S);

// ---- dual.hip
// want_lb = false: the caller ignores the returned bound, so it is not
// computed where nothing else needs it (0 is returned then).
double dual_solver(Ctx& ctx, DCoo& A, int max_cycle_length, int num_iter, float tri_memory_factor, int num_outer,
                   float tol_ratio, bool want_lb = true);

// ---- forest.hip
void connected_components(Ctx& ctx, const int* off, const int* adj, int n, int* labels);
// Forest edge indices (into the given edge list) of the maximum spanning
// forest under (cost desc, index asc); returns the count, ascending order.
i64 max_spanning_forest(Ctx& ctx, const int* r, const int* c, const float* w, i64 m, int n, DBuf<int>& ids);
// edge_contractions_woc(A).find_contraction_mapping(); returns #forest entries.
int woc_mapping(Ctx& ctx, const DCoo& A, DBuf<int>& map, i64& map_size);

// ---- lockstep.hip
// A batch of instances packed back to back (instance q owns input edges
// [edge_off[q], edge_off[q+1]) with its own 0-based ids) and its outputs.
struct PackedBatch {
    int k;
    const int64_t* edge_off;
    const int32_t* i;
    const int32_t* j;
    const float* c;
    Options o;
    int32_t* labels_out;
    const int64_t* label_off;
    int64_t* n_labels;
    double* lbs;
    int32_t* status;
    std::string* first_error;
    std::mutex* emu;
};
void solve_packed(Ctx& ctx, PackedBatch& B, int workers);
// Receives each instance's outcome (from the lockstep thread or a worker).
struct LockstepSink {
    virtual ~LockstepSink() = default;
    virtual void done(int id, int status, const std::string& error, const int32_t* labels, i64 n, double lb) = 0;
};
// The lockstep engine behind an opaque handle: admit groups of instances,
// advance all admitted ones by one contraction round per call.
struct LockstepHandle;
constexpr int kLockstepMaxSlots = 4096;  // instances in one pack at a time
LockstepHandle* lockstep_create(Ctx& ctx, const Options& o, int workers, LockstepSink& sink);
void lockstep_admit(LockstepHandle* h, int k, const int64_t* edge_off, const int32_t* i, const int32_t* j,
                    const float* c, const int* ids);
void lockstep_round(LockstepHandle* h);
int lockstep_slots(const LockstepHandle* h);
void lockstep_destroy(LockstepHandle* h);  // waits for handed-off instances

// ---- solver.hip
struct SolveResult {
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    DBuf<int> labels;
    i64 num_labels = 0;
    double lb = 0.0;
    int iterations = 0;
    std::vector<std::vector<int>> timeline;
};
SolveResult solve(Ctx& ctx, DCoo& A, const Options& o);
void solve_rounds(Ctx& ctx, DCoo& A, SolveResult& out, const Options& o, int start_iter, bool use_mm);

}  // namespace rama
