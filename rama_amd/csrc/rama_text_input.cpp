// rama_text_input — the command-line front end (src/rama_text_input.cu:8-50
// of RAMA): read a MULTICUT text file, solve, print the objective, the
// lower bound and the solve time, and optionally write the clustering.
//
//   rama_text_input -f instance.txt [-o clusters.txt] [max_cycle_dual ...]
//
// The solution file has one line per cluster id 0 .. max label, each the
// ids of the nodes with that label followed by a tab, as the reference
// writes it (an id without nodes gives an empty line).  Failures print the
// library's message and exit with status 1.
#include <algorithm>
#include <fstream>
#include <iostream>
#include <vector>

#include "rama_cuda.h"

int main(int argc, char** argv) {
    multicut_solver_options opts;
    const int e = opts.from_cl(argc, argv);
    if (e != -1) return e;
    try {
        std::vector<int> i, j;
        std::vector<float> costs;
        std::tie(i, j, costs) = read_file(opts.input_file);
        std::vector<int> labels;
        double lb = 0.0;
        int ms = 0;
        std::vector<std::vector<int>> timeline;
        std::tie(labels, lb, ms, timeline) = rama_cuda(i, j, costs, opts);
        if (!opts.only_compute_lb) {
            const double obj = get_obj(labels, i, j, costs);
            std::cout << "\tcost w.r.t original objective: " << obj << "\n";
        }
        std::cout << "\tfinal lower bound: " << lb << "\n";
        std::cout << "\tGPU compute time: " << ms << "ms\n";
        if (!opts.output_sol_file.empty()) {
            std::ofstream sol(opts.output_sol_file);
            if (!sol) throw std::runtime_error("Could not open solution file " + opts.output_sol_file);
            std::cout << "Writing solution to file: " << opts.output_sol_file << "\n";
            // nodes bucketed by label in increasing node order (one pass
            // instead of one scan of all nodes per cluster)
            const int num_clusters = labels.empty() ? 0 : *std::max_element(labels.begin(), labels.end()) + 1;
            std::vector<int> start((size_t)num_clusters + 1, 0), order(labels.size());
            for (int l : labels)
                if (l >= 0) ++start[(size_t)l + 1];
            for (int c = 0; c < num_clusters; ++c) start[c + 1] += start[c];
            std::vector<int> fill(start.begin(), start.end() - 1);
            for (size_t v = 0; v < labels.size(); ++v)
                if (labels[v] >= 0) order[fill[labels[v]]++] = (int)v;
            for (int c = 0; c < num_clusters; ++c) {
                for (int p = start[c]; p < start[c + 1]; ++p) sol << order[p] << "\t";
                sol << "\n";
            }
        }
    } catch (const std::exception& ex) {
        std::cerr << "rama_text_input: " << ex.what() << "\n";
        return 1;
    }
    return 0;
}
