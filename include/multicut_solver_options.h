// multicut_solver_options — the solver's option set with the reference's
// fields, defaults, presets and string form
// (include/multicut_solver_options.h:4-127 of RAMA).  The command-line
// parser (from_cl, CLI11) belongs to the CLI front end and is not part of
// this header; to_c() converts to the C ABI struct of rama_hip.h.
#pragma once

#include <iostream>
#include <stdexcept>
#include <string>

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
