/*
 * rama_hip.h — C ABI of the MI355X-native primal-dual multicut solver.
 *
 * Every entry point returns 0 on success and a non-zero status on failure;
 * the message of the last failure on the calling thread is returned by
 * rama_last_error().  No C++ exception crosses this boundary.
 *
 * Orientation of the reference interface that each entry replaces is given
 * next to it (paths relative to the RAMA source tree).
 */
#ifndef RAMA_HIP_H
#define RAMA_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
#define RAMA_OK 0
#define RAMA_ERR_INVALID 1     /* bad argument (sizes, ids, capacity)            */
#define RAMA_ERR_HIP 2         /* a HIP runtime call failed                        */
#define RAMA_ERR_BAD_CONTRACTION 3 /* "Found bad contractions" (src/rama_cuda.cu:136-137) */
#define RAMA_ERR_NO_DEVICE 4   /* no usable gfx950 device                          */
#define RAMA_ERR_OOM 5         /* device allocation failed                         */
#define RAMA_ERR_INTERNAL 6

/*
 * Solver options.  Field-for-field the reference's multicut_solver_options
 * (include/multicut_solver_options.h:4-19) minus the two file-name strings,
 * which belong to the command-line front end.
 */
typedef struct rama_options_c {
    int32_t max_cycle_length_lb;            /* default 5    */
    int32_t num_dual_itr_lb;                /* default 10   */
    int32_t max_cycle_length_primal;        /* default 3    */
    int32_t num_dual_itr_primal;            /* default 5    */
    int32_t num_outer_itr_dual;             /* default 1    */
    float mean_multiplier_mm;               /* default 0.05 */
    float matching_thresh_crossover_ratio;  /* default 0.05 */
    float tri_memory_factor;                /* default 2.0  */
    int32_t only_compute_lb;                /* default 0    */
    int32_t max_time_sec;                   /* default -1   */
    int32_t dump_timeline;                  /* default 0    */
    int32_t verbose;                        /* default 1    */
    int32_t sanitize_graph;                 /* default 0    */
} rama_options_c;

/* Per-contraction label snapshots (the reference's timeline,
 * src/rama_cuda.cu:145-150). Owned by the library; free with
 * rama_timeline_free. */
typedef struct rama_timeline rama_timeline;

/* Message of the last failure on this thread ("" if none). */
const char* rama_last_error(void);

/* Library version string. */
const char* rama_version(void);

/* Defaults of multicut_solver_options() (include/multicut_solver_options.h:5-17). */
void rama_options_default(rama_options_c* out);

/* multicut_solver_options(const std::string&) (include/multicut_solver_options.h:22-53):
 * "P", "PD", "PD+", "D".  An unknown name leaves the PD defaults in *out and
 * returns 1, as the reference silently yields PD; 0 otherwise. */
int rama_options_preset(const char* name, rama_options_c* out);

/*
 * Solve from host buffers.  Replaces
 *   rama_cuda(const std::vector<int>& i, const std::vector<int>& j,
 *             const std::vector<float>& costs, const multicut_solver_options&)
 *   (include/rama_cuda.h:8, src/rama_cuda.cu:166-193).
 * labels_out receives num_labels_out[0] labels (at most labels_cap; the call
 * fails with RAMA_ERR_INVALID if more are needed and reports the size needed
 * in num_labels_out).  time_ms_out receives the solve time as the reference
 * measures it (src/rama_cuda.cu:183-186).  timeline_out may be NULL.
 */
int rama_solve_host(const int32_t* i, const int32_t* j, const float* costs, int64_t num_edges,
                    const rama_options_c* opts,
                    int32_t* labels_out, int64_t labels_cap, int64_t* num_labels_out,
                    double* lb_out, int32_t* time_ms_out, rama_timeline** timeline_out);

/*
 * Solve from device buffers on `device`, ordered on `stream` (a hipStream_t;
 * NULL = the library's own stream).  Replaces
 *   rama_cuda(device_vector<int>&& i, device_vector<int>&& j,
 *             device_vector<float>&& costs, const multicut_solver_options&, int device)
 *   (include/rama_cuda.h:9, src/rama_cuda.cu:195-211)
 * and rama_cuda_gpu_pointers (src/rama_py.cu:61-74).  Inputs are not
 * modified.  d_labels_out is a device buffer of labels_cap int32.
 */
int rama_solve_device(const int32_t* d_i, const int32_t* d_j, const float* d_costs, int64_t num_edges,
                      const rama_options_c* opts, int device, void* stream,
                      int32_t* d_labels_out, int64_t labels_cap, int64_t* num_labels_out,
                      double* lb_out, int32_t* time_ms_out, rama_timeline** timeline_out);

/*
 * Allocation-free form of rama_solve_device.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and results as rama_solve_device, without the
 * timeline; stream NULL means the null stream itself (no stream is created).
 * Every device buffer the solve needs is carved out of one static device
 * array of RAMA_NOALLOC_SCRATCH_BYTES per GPU, which the library holds from
 * the first such call on a GPU until the process ends; the call makes no
 * hipMalloc, hipFree, hipHostMalloc, stream or event creation.
 * num_edges above RAMA_NOALLOC_MAX_EDGES fails with RAMA_ERR_INVALID before
 * any work; a graph whose solve needs more than the array holds fails with
 * RAMA_ERR_OOM before the kernel that would use the block is launched.
 * scratch_peak_out (may be NULL) receives the bytes of the array the solve
 * used at most.  All calls on one GPU share its array: calls that overlap
 * on one GPU (from several threads or streams) race and must be serialised
 * by the caller.  The kernel timing switch (rama_profile_enable) creates
 * events when on; leave it off here.
 */
#define RAMA_NOALLOC_SCRATCH_BYTES ((int64_t)256 << 20)
#define RAMA_NOALLOC_MAX_EDGES ((int64_t)1000000) /* a PD+ grid solve peaks near 240 bytes per edge */
int rama_solve_device_noalloc(const int32_t* d_i, const int32_t* d_j, const float* d_costs, int64_t num_edges,
                              const rama_options_c* opts, int device, void* stream,
                              int32_t* d_labels_out, int64_t labels_cap, int64_t* num_labels_out,
                              double* lb_out, int32_t* time_ms_out, int64_t* scratch_peak_out);

int64_t rama_timeline_count(const rama_timeline* t);
int64_t rama_timeline_length(const rama_timeline* t, int64_t k);
const int32_t* rama_timeline_labels(const rama_timeline* t, int64_t k);
void rama_timeline_free(rama_timeline* t);

/* Objective of a labelling: sum of costs of edges whose ends carry different
 * labels (include/rama_utils.h:192-205).  Host buffers, no device work. */
double rama_objective(const int32_t* labels, int64_t num_labels, const int32_t* i, const int32_t* j,
                      const float* costs, int64_t num_edges);

/* ------------------------------------------------------------------------
 * Stage entry points.  Host buffers in and out; each runs one stage of the
 * hot path on the GPU.  They exist so an integrator (and the parity tests)
 * can drive a single stage; the solver itself keeps everything resident.
 * A dCOO is passed as (row, col, val, nnz, rows, cols) with rows/cols the
 * reference's rows_/cols_ (include/dCOO.h:65-66).
 * ------------------------------------------------------------------------ */

/* dCOO::init on an unsorted directed edge list: orient (min,max), stable sort
 * by (row,col), derive rows_/cols_ (src/dCOO.cu:28-55). out_dims = {rows, cols}. */
int rama_stage_dcoo_init(const int32_t* i, const int32_t* j, const float* c, int64_t m,
                         int32_t* out_row, int32_t* out_col, float* out_val, int32_t* out_dims);

/* dCOO::contract_cuda + has_bad_contractions + remove_diagonal
 * (src/dCOO.cu:57-95,106-142, src/rama_cuda.cu:26-30,136-140).
 * Output capacity is m; out_dims = {rows, cols} as the reference computes them
 * before removing the diagonal; out_bad = 1 if a negative self-loop appeared;
 * out_diag_sum = sum of the contracted self-loop costs. */
int rama_stage_contract(const int32_t* row, const int32_t* col, const float* val, int64_t m,
                        const int32_t* node_map, int64_t map_size,
                        int32_t* out_row, int32_t* out_col, float* out_val, int64_t* out_m,
                        int32_t* out_dims, int32_t* out_bad, double* out_diag_sum);

/* dCOO::contract_cuda + remove_diagonal, then to_undirected + dCOO(symmetric)
 * + compute_row_offsets of the result (src/dCOO.cu:57-95,97-142,144-147,
 * include/rama_utils.h:107-131, src/conflicted_cycles_cuda.cu:208-217): the
 * contracted graph's symmetric adjacency as the next cycle search receives
 * it, with out_n = max(rows, cols) rows (out_off holds out_n + 1 offsets, up
 * to off_cap) and each row's neighbours ascending.  Short edge lists get it from the contraction's own
 * sort of both edge directions; it equals the CSR that
 * rama_stage_symmetric_csr (mode 0) builds from rama_stage_contract's output. */
int rama_stage_contract_symmetric(const int32_t* row, const int32_t* col, const float* val, int64_t m,
                                  const int32_t* node_map, int64_t map_size, int32_t* out_n, int32_t* out_off,
                                  int64_t off_cap, int32_t* out_adj, float* out_w, int64_t cap, int64_t* out_nnz);

/* compute_offsets(sorted ids, max_value) (include/rama_utils.h:403-413); out has max_value+2 entries. */
int rama_stage_compute_offsets(const int32_t* ids, int64_t n, int32_t max_value, int32_t* out);

/* to_undirected + dCOO(symmetric) + compute_row_offsets
 * (include/rama_utils.h:107-131, src/dCOO.cu:144-147,164-172,
 * src/conflicted_cycles_cuda.cu:208-217) of a (row, col)-sorted edge list
 * with row < col: the symmetric adjacency with n_nodes rows, each row's
 * neighbours ascending (a repeated pair keeps its input order).  mode 0:
 * every edge; 1: the edges with cost > keep_above (is_positive_edge); 2:
 * every edge, row offsets found one thread per row (the spanning-forest
 * form).  out_off has n_nodes+1 entries, out_adj / out_w capacity cap;
 * out_nnz = 2 x the edges taken. */
int rama_stage_symmetric_csr(const int32_t* row, const int32_t* col, const float* val, int64_t m,
                             int32_t n_nodes, int32_t mode, float keep_above, int32_t* out_off,
                             int32_t* out_adj, float* out_w, int64_t cap, int64_t* out_nnz);

/* compress_label_sequence(data, max_label) (include/rama_utils.h:160-177). */
int rama_stage_compress_labels(const int32_t* data, int64_t n, int32_t max_label, int32_t* out);

/* export_undirected + filter_edges_by_matching_vertex_based
 * (src/dCOO.cu:164-172, src/maximum_matching_vertex_based.cu:76-137).
 * out_map has capacity rows+cols; out_map_size = undirected rows_;
 * out_matched = number of matched vertices. */
int rama_stage_matching(const int32_t* row, const int32_t* col, const float* val, int64_t m,
                        int32_t rows, int32_t cols, float mean_multiplier_mm,
                        int32_t* out_map, int64_t* out_map_size, int32_t* out_matched);

/* conflicted_cycles_cuda (src/conflicted_cycles_cuda.cu:222-343) followed by
 * normalize_triangles (include/rama_utils.h:325-350).  Writes the sorted
 * unique triangles (v1<v2<v3) into out_t1..3 (capacity cap) and their number
 * into out_n; out_raw = number of triangle slots the search filled (the
 * reference's triangles_v1.size()). */
int rama_stage_conflicted_cycles(const int32_t* row, const int32_t* col, const float* val, int64_t m,
                                 int32_t rows, int32_t cols, int32_t max_cycle_length,
                                 float tri_memory_factor, float tol_ratio,
                                 int32_t* out_t1, int32_t* out_t2, int32_t* out_t3, int64_t cap,
                                 int64_t* out_n, int64_t* out_raw);

/* multicut_message_passing (src/multicut_message_passing.cu:65-360): build the
 * edge/triangle problem, then run `pre_triplet_sends` extra
 * send_messages_to_triplets() calls followed by `num_iterations` iteration()
 * calls.  out_lb[k] = lower_bound() before iteration k and out_lb[num_iterations]
 * the final one.  The reparametrized edges (sorted) go to out_i/out_j/out_c
 * (capacity cap, count in out_m). */
int rama_stage_message_passing(const int32_t* row, const int32_t* col, const float* val, int64_t m,
                               const int32_t* t1, const int32_t* t2, const int32_t* t3, int64_t num_tri,
                               int32_t pre_triplet_sends, int32_t num_iterations,
                               int32_t* out_i, int32_t* out_j, float* out_c, int64_t cap, int64_t* out_m,
                               double* out_lb);

/* dual_solver (src/dual_solver.cu:37-63) on a sorted directed dCOO.  The
 * reparametrized dCOO goes to out_row/out_col/out_val (capacity cap) with
 * out_dims = {rows, cols}; the lower bound to out_lb. */
int rama_stage_dual_solver(const int32_t* row, const int32_t* col, const float* val, int64_t m,
                           int32_t rows, int32_t cols, int32_t max_cycle_length, int32_t num_iter,
                           float tri_memory_factor, int32_t num_outer_itr, float tol_ratio,
                           int32_t* out_row, int32_t* out_col, float* out_val, int64_t cap,
                           int64_t* out_m, int32_t* out_dims, double* out_lb);

/* Connected components of a symmetric CSR graph; labels are the minimum node
 * id of each component, as ECL-CC's computeCC_gpu (external/ECL-CC/ECL-CC_11.cu:315-350). */
int rama_stage_connected_components(const int32_t* offsets, const int32_t* cols, int32_t num_nodes,
                                    int64_t nnz, int32_t* out_labels);

/* Maximum spanning forest of the attractive (cost > 0) edges of a directed
 * edge list (replaces MST_boruvka::maximum_spanning_tree,
 * external/cudaMST/gpuMST/gpuMST.cu:364-388).  Forest edges in input order. */
int rama_stage_max_spanning_forest(const int32_t* row, const int32_t* col, const float* val, int64_t m,
                                   int32_t* out_row, int32_t* out_col, float* out_val, int64_t* out_m);

/* edge_contractions_woc(A).find_contraction_mapping()
 * (src/edge_contractions_woc.cu:341-423, src/edge_contractions_woc_thrust.cu:544-599).
 * out_map capacity max(rows, cols); out_map_size = mapping length;
 * out_count = the reference's second return (#forest entries). */
int rama_stage_woc_mapping(const int32_t* row, const int32_t* col, const float* val, int64_t m,
                           int32_t rows, int32_t cols,
                           int32_t* out_map, int64_t* out_map_size, int32_t* out_count);

/* ------------------------------------------------------------------------
 * Many instances on one GPU (no counterpart in the reference, which solves
 * one instance per call).  Every instance is solved by the plain solver, so
 * results are identical to rama_solve_host on the same instance; a pool of
 * worker threads, each with its own stream, overlaps the solves on the
 * device.  For more than four concurrent streams set GPU_MAX_HW_QUEUES
 * (HIP's hardware-queue count) before the first HIP call in the process.
 * ------------------------------------------------------------------------ */

/* Batched solve: instance q owns edges [edge_off[q], edge_off[q+1]) of
 * (i, j, c), with its own 0-based node ids, and receives n_labels[q] labels
 * at labels_out + label_off[q] (capacity label_off[q+1] - label_off[q]) and
 * its lower bound in lbs[q]; status[q] (may be NULL) its status. */
int rama_solve_batch_host(int32_t k, const int64_t* edge_off, const int32_t* i, const int32_t* j, const float* c,
                          const rama_options_c* opts, int32_t workers, int32_t* labels_out, const int64_t* label_off,
                          int64_t* n_labels, double* lbs, int32_t* status);

/* Serving loop with continuous admission: requests are queued in arrival
 * order and admitted by `workers` lanes (threads with a stream each) while
 * the device-memory estimate of the requests in flight stays within
 * hbm_budget_bytes.  A lane solves a lone request with the plain solver and
 * a group of queued requests as one lockstep pack, which admits further
 * requests between its rounds while every lane is busy
 * (RAMA_SERVE_MODE=workers: one request per lane).  Each request's graph
 * and dual state stay resident on the device until its labels are back on
 * the host. */
typedef struct rama_server rama_server;
rama_server* rama_server_create(int device, int workers, int64_t hbm_budget_bytes, const rama_options_c* opts);
/* Copies the request; returns its ticket (>= 0) or -1. */
int64_t rama_server_submit(rama_server* s, const int32_t* i, const int32_t* j, const float* c, int64_t m);
/* 1 if the ticket's result is ready, 0 if not, -1 if unknown. */
int rama_server_poll(rama_server* s, int64_t ticket);
/* Blocks until the ticket is done; returns its status, labels, lower bound,
 * latency (submission to labels on the host) and device solve time. */
int rama_server_wait(rama_server* s, int64_t ticket, int32_t* labels_out, int64_t cap, int64_t* n_labels,
                     double* lb, double* latency_ms, double* solve_ms);
void rama_server_destroy(rama_server* s);

/* ------------------------------------------------------------------------
 * Kernel timing (no counterpart in the reference; its cumulative wall-clock
 * macros, include/time_measure_util.h:11-70, are the closest).  When enabled,
 * instrumented launches are bracketed by HIP events on the solver's stream
 * and aggregated per kernel name with the algorithmic bytes each launch must
 * move.  Entries stay valid until the next reset.
 * ------------------------------------------------------------------------ */
void rama_profile_enable(int on);
void rama_profile_reset(void);
int64_t rama_profile_count(void);
int rama_profile_entry(int64_t k, const char** name, int64_t* launches, double* total_ms, double* total_bytes);
/* Work of entry k that has no byte roofline: the cycle-search counts report
 * their expansions (units) and the neighbour-list entries a merge walk of
 * each expansion's two lists reads (entries). */
int rama_profile_work(int64_t k, double* units, double* entries);
/* Restrict timing to launches of one kernel name (NULL or "": all). */
void rama_profile_only(const char* name);

/* ------------------------------------------------------------------------
 * Launch-shape table (no counterpart in the reference, whose launches use
 * fixed shapes, e.g. 256 threads in src/conflicted_cycles_cuda.cu:240-241
 * and src/maximum_matching_vertex_based.cu:8, and Thrust's own sort
 * policies).  Knobs, each keyed by a size: sort_path_u64 / sort_path_u32
 * (pairs -> 0 three launches per pass, 1 onesweep, 2 top-digit + per-bucket
 * blocks), sort_threads (tiles -> 512 or 1024 threads per 4096-pair tile),
 * scan_chunks (items -> blocks of the integer scan), pick_blocks (CSR
 * entries) and match_blocks (vertices) for the matching grids,
 * mp_fused_max_blocks (message-passing edges -> largest edge grid that sums
 * its own lower bound), cycle_chunk_pct (expansions -> growth of the chunked
 * 4/5-cycle counts).  None of them changes a result.  The table is read on
 * first use from RAMA_TUNING_TABLE, else tuning/gfx950.table beside the
 * library; "none" or no file keeps the built-in defaults.
 * ------------------------------------------------------------------------ */

/* Replace the table by the file at `path` ("knob max_size value" lines;
 * max_size -1 = no limit); NULL or "" = the built-in defaults. */
int rama_tuning_load(const char* path);
/* Add or replace the row (knob, max_size) of the table in use. */
int rama_tuning_set(const char* knob, int64_t max_size, int64_t value);
/* Value of `knob` for `size` (-1 for an unknown knob). */
int64_t rama_tuning_get(const char* knob, int64_t size);
/* Where the table in use came from: a path, "defaults", or "modified ..." */
const char* rama_tuning_source(void);
/* Mean device time in ms of `reps` stable sorts of n random (key, u32)
 * pairs (key_bytes 4 or 8, keys of `bits` bits) by a given path (0, 1, 2 as
 * sort_path_*) and threads per tile (512 or 1024): what the tuner measures. */
int rama_tune_time_sort(int64_t n, int32_t key_bytes, int32_t bits, int32_t path, int32_t threads, int32_t reps,
                        double* ms_out);
/* Device-to-device copy rate in GB/s (bytes read + bytes written) of a
 * streaming copy kernel (16-byte non-temporal loads, grid-stride) over two
 * buffers of `bytes`: the fastest of eight launch shapes (4, 8 or 16
 * blocks per CU, or one pass over the whole array; 4 or 8 loads in flight
 * per thread), each the mean of `reps` timed copies.  The practical HBM rate beside the nominal peak.  No counterpart
 * in the reference. */
int rama_stream_copy_gbps(int64_t bytes, int32_t reps, double* gbps_out);

/* ------------------------------------------------------------------------
 * Device memory cache (no counterpart in the reference, whose thrust
 * vectors allocate and free per call).  Blocks a finished solve used stay
 * cached for the next solve on the same device, up to RAMA_CACHE_BYTES
 * (default 32 GiB) per device.  This returns them to the driver; device < 0
 * releases every device's cache.  Call it with no solve running.
 * ------------------------------------------------------------------------ */
void rama_release_cached(int device);

#ifdef __cplusplus
}
#endif

#endif /* RAMA_HIP_H */
