// rama_py.cpp — the pybind11 module `rama_py` with RAMA's Python surface
// (src/rama_py.cu:21-125), on top of librama_hip.so's C ABI.
//
//   multicut_solver_options   (), (str), (11 positional args); 13 fields; __repr__
//   rama_cuda(i, j, costs, opts) -> (labels, lb, time_ms, timeline)
//   rama_cuda_gpu_pointers(i_ptr, j_ptr, costs_ptr, labels_ptr, num_nodes, num_edges, device, opts) -> timeline
//   read_multicut_file(path) -> (i, j, costs)
//   rama_torch(i, j, costs, opts) -> [labels, lb]   (device tensors; torch is
//                                                     reached through Python, so
//                                                     the module links no torch)
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include "rama_cuda.h"

namespace py = pybind11;

static std::vector<py::object> rama_torch(py::object ti, py::object tj, py::object tc,
                                          const multicut_solver_options& opts) {
    py::module_ torch = py::module_::import("torch");
    auto check = [&](py::object t, const char* name) {
        if (!t.attr("is_cuda").cast<bool>()) throw std::runtime_error(std::string(name) + " must be a CUDA tensor");
        if (!t.attr("is_contiguous")().cast<bool>()) throw std::runtime_error(std::string(name) + " must be contiguous");
        if (t.attr("dim")().cast<int>() != 1) throw std::runtime_error(std::string(name) + " should be one-dimensional");
    };
    check(ti, "_i");
    check(tj, "_j");
    check(tc, "_costs");
    const int64_t m = ti.attr("size")(0).cast<int64_t>();
    if (tj.attr("size")(0).cast<int64_t>() != m || tc.attr("size")(0).cast<int64_t>() != m)
        throw std::runtime_error("Input shapes must match");
    if (!ti.attr("dtype").equal(tj.attr("dtype")))
        throw std::runtime_error("Node indices i, j should be of same type");
    if (!ti.attr("dtype").equal(torch.attr("int32")) || !tc.attr("dtype").equal(torch.attr("float32")))
        throw std::runtime_error("i, j must be int32 and costs float32");
    const int device = tc.attr("device").attr("index").cast<int>();
    if (device < 0) throw std::runtime_error("Invalid device ID");
    const int64_t n_cap = std::max(ti.attr("max")().attr("item")().cast<int64_t>(),
                                   tj.attr("max")().attr("item")().cast<int64_t>()) + 2;
    py::object labels = torch.attr("empty")(n_cap, py::arg("dtype") = torch.attr("int32"),
                                            py::arg("device") = ti.attr("device"));
    const int64_t stream = torch.attr("cuda").attr("current_stream")(device).attr("cuda_stream").cast<int64_t>();
    const rama_options_c o = opts.to_c();
    int64_t n = 0;
    double lb = 0.0;
    int32_t ms = 0;
    const auto ptr = [](py::object t) { return reinterpret_cast<void*>(t.attr("data_ptr")().cast<intptr_t>()); };
    rama_detail::check(rama_solve_device(static_cast<const int32_t*>(ptr(ti)), static_cast<const int32_t*>(ptr(tj)),
                                         static_cast<const float*>(ptr(tc)), m, &o, device,
                                         reinterpret_cast<void*>(stream), static_cast<int32_t*>(ptr(labels)), n_cap,
                                         &n, &lb, &ms, nullptr));
    py::object node_mapping = labels.attr("narrow")(0, 0, n).attr("clone")();
    py::object lb_t = torch.attr("full")(py::make_tuple(1), lb, py::arg("dtype") = torch.attr("float64"),
                                         py::arg("device") = ti.attr("device"));
    return {node_mapping, lb_t};
}

PYBIND11_MODULE(rama_py, m) {
    m.doc() = "Bindings for RAMA: Rapid algorithm for multicut (MI355X HIP build). "
              "For running purely primal algorithm initialize multicut_solver_options with \"P\". "
              "For algorithm with best quality call with \"PD+\" where \"PD\" is default algorithm. "
              "For only computing the lower bound call with \"D\". ";
    py::class_<multicut_solver_options>(m, "multicut_solver_options")
        .def(py::init<>())
        .def(py::init<const std::string&>())
        .def(py::init<const int&, const int&, const int&, const int&, const int&, const float&, const float&,
                      const float&, const bool&, const int&, const bool&>())
        .def_readwrite("max_cycle_length_lb", &multicut_solver_options::max_cycle_length_lb)
        .def_readwrite("num_dual_itr_lb", &multicut_solver_options::num_dual_itr_lb)
        .def_readwrite("max_cycle_length_primal", &multicut_solver_options::max_cycle_length_primal)
        .def_readwrite("num_dual_itr_primal", &multicut_solver_options::num_dual_itr_primal)
        .def_readwrite("num_outer_itr_dual", &multicut_solver_options::num_outer_itr_dual)
        .def_readwrite("mean_multiplier_mm", &multicut_solver_options::mean_multiplier_mm)
        .def_readwrite("matching_thresh_crossover_ratio", &multicut_solver_options::matching_thresh_crossover_ratio)
        .def_readwrite("tri_memory_factor", &multicut_solver_options::tri_memory_factor)
        .def_readwrite("only_compute_lb", &multicut_solver_options::only_compute_lb)
        .def_readwrite("max_time_sec", &multicut_solver_options::max_time_sec)
        .def_readwrite("verbose", &multicut_solver_options::verbose)
        .def_readwrite("dump_timeline", &multicut_solver_options::dump_timeline)
        .def_readwrite("sanitize_graph", &multicut_solver_options::sanitize_graph)
        .def("__repr__", [](const multicut_solver_options& a) { return a.get_string(); });

    m.def("rama_cuda", [](const std::vector<int>& i, const std::vector<int>& j, const std::vector<float>& edge_costs,
                          const multicut_solver_options& opts) {
        py::gil_scoped_release nogil;
        return rama_cuda(i, j, edge_costs, opts);
    });

    m.def("rama_cuda_gpu_pointers", [](const long i_ptr, const long j_ptr, const long edge_costs_ptr,
                                       const long node_labels_out_ptr, const int num_nodes, const int num_edges,
                                       const int gpuDeviceID, const multicut_solver_options& opts) {
        py::gil_scoped_release nogil;
        return rama_cuda_gpu_pointers(reinterpret_cast<const int*>(i_ptr), reinterpret_cast<const int*>(j_ptr),
                                      reinterpret_cast<const float*>(edge_costs_ptr),
                                      reinterpret_cast<int*>(node_labels_out_ptr), num_nodes, num_edges, gpuDeviceID,
                                      opts);
    });

    m.def("read_multicut_file", [](const std::string& filename) { return read_file(filename); });

    m.def("rama_torch", &rama_torch, "RAMA solver with torch interface (ROCm tensors).");
}
