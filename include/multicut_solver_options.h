// multicut_solver_options — the solver's option set with the reference's
// fields, defaults, presets and string form
// (include/multicut_solver_options.h:4-127 of RAMA).  from_cl() parses the
// reference's command line (its :82-105, which uses CLI11) with a small
// parser of its own; to_c() converts to the C ABI struct of rama_hip.h.
#pragma once

#include <sys/stat.h>

#include <cstdlib>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "rama_hip.h"

struct multicut_solver_options {
    std::string input_file;
    std::string output_sol_file = "";
    int max_cycle_length_lb = 5;
    int num_dual_itr_lb = 10;
    int max_cycle_length_primal = 3;
    int num_dual_itr_primal = 5;
    int num_outer_itr_dual = 1;
    float mean_multiplier_mm = 0.05;
    float matching_thresh_crossover_ratio = 0.05;
    float tri_memory_factor = 2.0;
    bool only_compute_lb = false;
    int max_time_sec = -1;
    bool dump_timeline = false;
    bool verbose = true;
    bool sanitize_graph = false;

    multicut_solver_options() {}

    // Presets "P", "PD", "PD+", "D"; each prints its one-line description.
    // An unknown name keeps the PD defaults (the reference constructs but
    // never throws its error).
    explicit multicut_solver_options(const std::string& solver_type) {
        if (solver_type == "PD") {
            std::cout << "Running solver type PD which offers best compute time versus quality tradeoff." << std::endl;
        } else if (solver_type == "P") {
            std::cout << "Running purely primal solver (better runtime, worse quality)" << std::endl;
            max_cycle_length_lb = 0;
            num_dual_itr_lb = 0;
            max_cycle_length_primal = 0;
            num_dual_itr_primal = 0;
        } else if (solver_type == "PD+") {
            std::cout << "Running PD+ solver (worse runtime, better quality)" << std::endl;
            max_cycle_length_lb = 5;
            num_dual_itr_lb = 10;
            max_cycle_length_primal = 5;
            num_dual_itr_primal = 10;
        } else if (solver_type == "D") {
            std::cout << "Running dual solver to compute only the lower bound." << std::endl;
            max_cycle_length_lb = 5;
            num_dual_itr_lb = 10;
            num_outer_itr_dual = 5;
            only_compute_lb = true;
        }
    }

    multicut_solver_options(const int _max_cycle_length_lb, const int _num_dual_itr_lb,
                            const int _max_cycle_length_primal, const int _num_dual_itr_primal,
                            const int _num_outer_itr_dual, const float _mean_multiplier_mm,
                            const float _matching_thresh_crossover_ratio, const float _tri_memory_factor,
                            const bool _only_compute_lb, const int _max_time_sec, const bool _dump_timeline = false,
                            const bool _sanitize_graph = false)
        : max_cycle_length_lb(_max_cycle_length_lb),
          num_dual_itr_lb(_num_dual_itr_lb),
          max_cycle_length_primal(_max_cycle_length_primal),
          num_dual_itr_primal(_num_dual_itr_primal),
          num_outer_itr_dual(_num_outer_itr_dual),
          mean_multiplier_mm(_mean_multiplier_mm),
          matching_thresh_crossover_ratio(_matching_thresh_crossover_ratio),
          tri_memory_factor(_tri_memory_factor),
          only_compute_lb(_only_compute_lb),
          max_time_sec(_max_time_sec),
          dump_timeline(_dump_timeline),
          sanitize_graph(_sanitize_graph) {}

    std::string get_string() const {
        return std::string("<multicut_solver_options>:") + "max_cycle_length_lb: " + std::to_string(max_cycle_length_lb) +
               ", num_dual_itr_lb: " + std::to_string(num_dual_itr_lb) +
               ", num_dual_itr_primal: " + std::to_string(num_dual_itr_primal) +
               ", max_cycle_length_primal: " + std::to_string(max_cycle_length_primal) +
               ", num_outer_itr_dual: " + std::to_string(num_outer_itr_dual) +
               ", mean_multiplier_mm: " + std::to_string(mean_multiplier_mm) +
               ", matching_thresh_crossover_ratio: " + std::to_string(matching_thresh_crossover_ratio) +
               ", tri_memory_factor: " + std::to_string(tri_memory_factor) +
               ", only_compute_lb: " + std::to_string(only_compute_lb) +
               ", max_time_sec: " + std::to_string(max_time_sec) +
               ", sanitize_graph: " + std::to_string(sanitize_graph) + ", verbose: " + std::to_string(verbose) + "\n";
    }

    void print() const { std::cout << this->get_string(); }

    // Command line of the reference's front end (include/multicut_solver_options.h:82-105):
    //   -f,--file,file_pos PATH (required, must exist)   -o,--out_sol_file PATH
    //   positionals after the file: max_cycle_dual [0,5], dual_itr >= 0,
    //   max_cycle_primal [0,5], dual_itr_primal >= 0, dual_itr_outer >= 0,
    //   mean_multiplier_mm >= 0, matching_thresh_crossover_ratio >= 0,
    //   tri_memory_factor > 0
    //   flags --only_lb, --dump_timeline, --sanitize_graph; -h,--help
    // Returns -1 when the solver should run, else the process exit code
    // (0 after --help; CLI11's codes for errors: 104 conversion, 105
    // validation, 106 missing required option, 109 unexpected arguments,
    // 114 an option without its value).
    int from_cl(int argc, char** argv) {
        enum { kConversion = 104, kValidation = 105, kRequired = 106, kExtras = 109, kMismatch = 114 };
        auto fail = [](int code, const std::string& msg) {
            std::cerr << msg << "\nRun with --help for more information.\n";
            return code;
        };
        struct Pos {
            const char* name;
            int* i;
            float* f;
            double lo, hi;
            bool lo_open;
        };
        const double inf = 1e300;
        std::vector<Pos> pos = {
            {"max_cycle_dual", &max_cycle_length_lb, nullptr, 0, 5, false},
            {"dual_itr", &num_dual_itr_lb, nullptr, 0, inf, false},
            {"max_cycle_primal", &max_cycle_length_primal, nullptr, 0, 5, false},
            {"dual_itr_primal", &num_dual_itr_primal, nullptr, 0, inf, false},
            {"dual_itr_outer", &num_outer_itr_dual, nullptr, 0, inf, false},
            {"mean_multiplier_mm", nullptr, &mean_multiplier_mm, 0, inf, false},
            {"matching_thresh_crossover_ratio", nullptr, &matching_thresh_crossover_ratio, 0, inf, false},
            {"tri_memory_factor", nullptr, &tri_memory_factor, 0, inf, true},
        };
        bool have_file = false;
        size_t next = 0;
        const std::string prog = argc > 0 ? argv[0] : "rama_text_input";
        for (int k = 1; k < argc; ++k) {
            std::string a = argv[k], val;
            const auto eq = a.find('=');
            const bool is_long = a.rfind("--", 0) == 0;
            if (is_long && eq != std::string::npos) {
                val = a.substr(eq + 1);
                a = a.substr(0, eq);
            }
            auto take = [&](std::string& dst) -> int {
                if (!val.empty() || (is_long && eq != std::string::npos)) {
                    dst = val;
                    return 0;
                }
                if (k + 1 >= argc) return fail(kMismatch, a + " requires a value");
                dst = argv[++k];
                return 0;
            };
            if (a == "-h" || a == "--help") {
                std::cout << "Solver for multicut problem. \nUsage: " << prog
                          << " [OPTIONS] file_pos [max_cycle_dual] [dual_itr] [max_cycle_primal] [dual_itr_primal]"
                             " [dual_itr_outer] [mean_multiplier_mm] [matching_thresh_crossover_ratio]"
                             " [tri_memory_factor]\n\nOptions:\n"
                             "  -h,--help                   Print this help message and exit\n"
                             "  -f,--file,file_pos PATH     Path to multicut instance (.txt) (required)\n"
                             "  -o,--out_sol_file PATH      Path to save node labeling (.txt)\n"
                             "  --only_lb                   Only compute the lower bound\n"
                             "  --dump_timeline             Return the output of each contraction step\n"
                             "  --sanitize_graph            Input has nodes without edges (their labels are -1)\n";
                return 0;
            } else if (a == "-f" || a == "--file") {
                if (int e = take(input_file)) return e;
                have_file = true;
            } else if (a == "-o" || a == "--out_sol_file") {
                if (int e = take(output_sol_file)) return e;
            } else if (a == "--only_lb") {
                only_compute_lb = true;
            } else if (a == "--dump_timeline") {
                dump_timeline = true;
            } else if (a == "--sanitize_graph") {
                sanitize_graph = true;
            } else if (a.size() > 1 && a[0] == '-' && !(a[1] >= '0' && a[1] <= '9') && a[1] != '.') {
                return fail(kExtras, "The following argument was not expected: " + a);
            } else if (!have_file) {
                input_file = a;
                have_file = true;
            } else if (next < pos.size()) {
                const Pos& p = pos[next++];
                char* end = nullptr;
                const double v = p.i ? (double)std::strtol(a.c_str(), &end, 10) : std::strtod(a.c_str(), &end);
                if (a.empty() || end == nullptr || *end != '\0')
                    return fail(kConversion, "Could not convert: " + std::string(p.name) + " = " + a);
                if (v < p.lo || (p.lo_open && v <= p.lo) || v > p.hi)
                    return fail(kValidation, std::string(p.name) + ": Value " + a + " not in range [" +
                                                 std::to_string(p.lo) + " - " +
                                                 (p.hi >= inf ? std::string("inf") : std::to_string(p.hi)) +
                                                 (p.lo_open ? "], exclusive of the lower end" : "]"));
                if (p.i) *p.i = (int)v; else *p.f = (float)v;
            } else {
                return fail(kExtras, "The following argument was not expected: " + a);
            }
        }
        if (!have_file) return fail(kRequired, "--file is required");
        struct stat st;
        if (stat(input_file.c_str(), &st) != 0)
            return fail(kValidation, "--file: Path does not exist: " + input_file);
        return -1;
    }

    rama_options_c to_c() const {
        rama_options_c o;
        o.max_cycle_length_lb = max_cycle_length_lb;
        o.num_dual_itr_lb = num_dual_itr_lb;
        o.max_cycle_length_primal = max_cycle_length_primal;
        o.num_dual_itr_primal = num_dual_itr_primal;
        o.num_outer_itr_dual = num_outer_itr_dual;
        o.mean_multiplier_mm = mean_multiplier_mm;
        o.matching_thresh_crossover_ratio = matching_thresh_crossover_ratio;
        o.tri_memory_factor = tri_memory_factor;
        o.only_compute_lb = only_compute_lb ? 1 : 0;
        o.max_time_sec = max_time_sec;
        o.dump_timeline = dump_timeline ? 1 : 0;
        o.verbose = verbose ? 1 : 0;
        o.sanitize_graph = sanitize_graph ? 1 : 0;
        return o;
    }
};
