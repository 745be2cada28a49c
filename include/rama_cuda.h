// rama_cuda.h — C++ entry points with the reference's signatures
// (include/rama_cuda.h:8-9, src/multicut_text_parser.cpp:6-31 of RAMA),
// header-only over the C ABI of librama_hip.so.  Failures are thrown as
// std::runtime_error carrying rama_last_error(), as the reference throws.
#pragma once

#include <fstream>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "multicut_solver_options.h"
#include "rama_hip.h"

namespace rama_detail {
inline void check(int rc) {
    if (rc != RAMA_OK) throw std::runtime_error(rama_last_error());
}
inline std::vector<std::vector<int>> take_timeline(rama_timeline* t) {
    std::vector<std::vector<int>> out;
    if (!t) return out;
    for (int64_t k = 0; k < rama_timeline_count(t); ++k) {
        const int32_t* p = rama_timeline_labels(t, k);
        out.emplace_back(p, p + rama_timeline_length(t, k));
    }
    rama_timeline_free(t);
    return out;
}
}  // namespace rama_detail

// rama_cuda(i, j, costs, opts) -> (node labels, lower bound, solve ms, timeline)
inline std::tuple<std::vector<int>, double, int, std::vector<std::vector<int>>> rama_cuda(
    const std::vector<int>& i, const std::vector<int>& j, const std::vector<float>& costs,
    const multicut_solver_options& opts) {
    if (i.size() != j.size() || i.size() != costs.size()) throw std::runtime_error("Input shapes must match");
    int64_t cap = 1;
    for (size_t e = 0; e < i.size(); ++e) cap = std::max<int64_t>(cap, std::max(i[e], j[e]) + 1);
    std::vector<int> labels((size_t)cap);
    int64_t n = 0;
    double lb = 0.0;
    int32_t ms = 0;
    rama_timeline* tl = nullptr;
    const rama_options_c o = opts.to_c();
    rama_detail::check(rama_solve_host(i.data(), j.data(), costs.data(), (int64_t)costs.size(), &o, labels.data(), cap,
                                       &n, &lb, &ms, &tl));
    labels.resize((size_t)n);
    return {labels, lb, (int)ms, rama_detail::take_timeline(tl)};
}

// Device-pointer form (rama_cuda_gpu_pointers, src/rama_py.cu:61-74): writes
// the labels into node_labels (device memory the caller sized) and returns
// the timeline.
inline std::vector<std::vector<int>> rama_cuda_gpu_pointers(const int* i, const int* j, const float* edge_costs,
                                                            int* node_labels, int num_nodes, int num_edges,
                                                            int device, const multicut_solver_options& opts,
                                                            void* stream = nullptr) {
    int64_t n = 0;
    double lb = 0.0;
    int32_t ms = 0;
    rama_timeline* tl = nullptr;
    const rama_options_c o = opts.to_c();
    // the reference writes node_mapping.size() = largest node id + 1 labels
    // into a buffer the caller sized for num_nodes; more labels than that
    // fail with the size error instead of writing past the buffer
    rama_detail::check(rama_solve_device(i, j, edge_costs, num_edges, &o, device, stream, node_labels,
                                         (int64_t)num_nodes, &n, &lb, &ms, &tl));
    return rama_detail::take_timeline(tl);
}

// read_file (src/multicut_text_parser.cpp:6-31): "MULTICUT" then "i j cost" lines.
inline std::tuple<std::vector<int>, std::vector<int>, std::vector<float>> read_file(const std::string& filename) {
    std::ifstream f(filename);
    if (!f.is_open()) throw std::runtime_error("Could not open multicut input file " + filename);
    std::string init_line;
    std::getline(f, init_line);
    if (init_line != "MULTICUT") throw std::runtime_error("first line must be 'MULTICUT'");
    std::vector<int> iv, jv;
    std::vector<float> cv;
    int a, b;
    float c;
    while (f >> a >> b >> c) {
        iv.push_back(a);
        jv.push_back(b);
        cv.push_back(c);
    }
    return {iv, jv, cv};
}

// get_obj (include/rama_utils.h:192-205)
inline double get_obj(const std::vector<int>& labels, const std::vector<int>& i, const std::vector<int>& j,
                      const std::vector<float>& costs) {
    return rama_objective(labels.data(), (int64_t)labels.size(), i.data(), j.data(), costs.data(),
                          (int64_t)costs.size());
}
