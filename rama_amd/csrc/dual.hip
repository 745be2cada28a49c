// dual.hip — the dual driver: per cycle length, search conflicted cycles,
// run message passing until the lower bound stalls, and replace the graph by
// its reparametrisation (src/dual_solver.cu:7-63).
#include "rama_internal.hpp"

namespace rama {



// dual_update_cycle_length (src/dual_solver.cu:7-35).  Returns the lower
// bound; `num_tri` receives the raw triangle count of the search.
static double dual_update_cycle_length(Ctx& ctx, DCoo& A, int c, int steps, float factor, float tol, i64* num_tri,
                                       bool want_lb) {
    Triangles tri = conflicted_cycles(ctx, A, c, factor, tol);
    *num_tri = tri.raw;
    if (tri.raw == 0) return want_lb ? dcoo_lb(ctx, A.val.get(), A.nnz) : 0.0;
    MPState S;
    mp_build(ctx, A, tri, S);
    tri = Triangles();
    double lb;
    if (ctx.verbose) {
        // host-driven loop so every iteration's bound can be printed
        lb = mp_lower_bound(ctx, S);
        double prev = 0.0;
        for (int it = 0; it < steps; ++it) {
            std::printf("dual updates cycle length: %d, iteration: %d, lower bound: %.6f\n", c, it, lb);
            if (it > 0 && (lb - prev) < 1e-3) break;
            mp_iteration(ctx, S);
            prev = lb;
            lb = mp_fused_lower_bound(ctx, S);
        }
    } else {
        lb = mp_run(ctx, S, steps);
    }
    if (ctx.verbose) std::printf("dual updates cycle length: %d, final lower bound: %.6f\n", c, lb);
    A = mp_reparametrized(ctx, S);
    return lb;
}

double dual_solver(Ctx& ctx, DCoo& A, int max_len, int num_iter, float factor, int num_outer, float tol,
                   bool want_lb) {
    if (max_len < 3 || num_iter == 0 || num_outer == 0) return want_lb ? dcoo_lb(ctx, A.val.get(), A.nnz) : 0.0;
    // the bound also drives the outer stall test when there are several passes
    const bool need = want_lb || num_outer > 1;
    double fin = 0.0, prev = 0.0;
    i64 num_tri = 0;
    try {
        for (int o = 0; o < num_outer; ++o) {
            for (int c = 3; c <= max_len; ++c)
                fin = dual_update_cycle_length(ctx, A, c, num_iter, factor, tol, &num_tri, need);
            if (num_tri == 0 || (o > 0 && (fin - prev) < 1e-3)) break;
            prev = fin;
        }
    } catch (const Error& e) {
        // out of device memory: trivial bound, as the reference's bad_alloc
        // handler (src/dual_solver.cu:58-61); A keeps its last complete state
        if (e.status != RAMA_ERR_OOM) throw;
        std::fprintf(stderr, "Dual solver out of memory, returning trivial lower bound\n");
        fin = dcoo_lb(ctx, A.val.get(), A.nnz);
    }
    return fin;
}

}  // namespace rama
