"""Writes tests/golden/known_answers.json.

The vectors are the inputs and expected outputs that the reference's own
tests assert (file:line given per case), plus the README triangle worked by
hand from the reference code.  They are data only; nothing of the
reference is executed.
"""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))

CASES = {
    # test/compute_offsets_test.cu:8-33
    "compute_offsets": [
        {"ids": [0, 1, 2, 2, 3, 4, 4], "max_value": 4, "offsets": [0, 1, 2, 4, 5, 7]},
        {"ids": [2, 2, 4, 4, 4, 5, 7, 7], "max_value": 7, "offsets": [0, 0, 0, 2, 2, 5, 6, 6, 8]},
    ],
    # test/compute_unique_test.cu:8-23 (unique values 0..4 of a sorted vector)
    "compress_labels": [
        {"data": [0, 1, 2, 2, 3, 4, 4], "max_label": 4, "labels": [0, 1, 2, 2, 3, 4, 4], "num_unique": 5},
    ],
    # test/cc_gpu_test.cu:53-69: CSR graphs and their component counts
    "connected_components": [
        {"offsets": [0, 1, 2, 4, 5, 6], "cols": [1, 0, 3, 4, 2, 2], "num_components": 2},
        {"offsets": [0, 0, 1, 2], "cols": [2, 1], "num_components": 2},
    ],
    # test/conflicted_cycles_test.cu:28-110: dCOO(i, j, costs) with budget
    # factor 10; raw triangle counts for max cycle length 3, 4, 5
    "conflicted_cycles": [
        {"name": "3-cycle", "i": [0, 1, 2], "j": [1, 2, 0], "costs": [-1, 2, 1], "factor": 10.0, "counts": [1, 1, 1]},
        {"name": "4-cycle", "i": [0, 1, 2, 3], "j": [1, 2, 3, 0], "costs": [-1, 2, 1, 3], "factor": 10.0,
         "counts": [0, 2, 2]},
        {"name": "4-cycle non-conflicted", "i": [0, 1, 2, 3], "j": [1, 2, 3, 0], "costs": [-1, -2, 1, 3],
         "factor": 10.0, "counts": [0, 0, 0]},
        {"name": "5-cycle", "i": [0, 1, 2, 3, 4], "j": [1, 2, 3, 4, 0], "costs": [-1, 2, 1, 3, 2], "factor": 10.0,
         "counts": [0, 0, 3]},
    ],
    # test/multicut_message_passing_test.cu:4-30
    "message_passing": [
        {"i": [0, 0, 0, 1, 2], "j": [1, 2, 3, 3, 3], "costs": [1.0, 1.0, -2.0, 1.0, 1.0],
         "t1": [0, 0], "t2": [1, 2], "t3": [3, 3], "pre_triplet_sends": 1, "iterations": 3,
         "initial_lb": -2.0, "final_lb": 0.0, "tol": 1e-6},
    ],
    # test/maximum_matching_vertex_based_test.cu:8-32 (mean_multiplier_mm = 0)
    "matching": [
        {"i": [0, 1, 0, 3, 3], "j": [1, 2, 2, 0, 1], "costs": [2.0, 3.0, -1.0, 1.0, 2.0], "mean_multiplier_mm": 0.0,
         "matched_vertices": 4},
    ],
    # test/edge_contractions_woc_test.cu:7-80: pairs that must end in different clusters
    "woc_mapping": [
        {"i": [0, 1, 1, 0, 0], "j": [1, 2, 3, 3, 2], "costs": [2.0, 3.0, 1.0, -3.0, -2.0], "separated": [[0, 3], [0, 2]]},
        {"i": [0, 1, 1, 0, 0, 0, 3, 3, 3], "j": [1, 2, 3, 3, 2, 4, 4, 5, 2],
         "costs": [2.0, 3.0, 1.0, -3.0, -2.0, 0.5, 0.75, 0.9, 0.8], "separated": [[0, 3], [0, 2]]},
        {"i": [0, 1, 1, 0, 3, 3, 3, 1, 0, 0], "j": [1, 2, 3, 4, 4, 5, 2, 5, 2, 3],
         "costs": [2.0, 3.0, 1.0, 0.5, 0.75, 0.9, 0.8, -2.0, -3.0, -1.0], "separated": [[1, 5], [0, 2], [0, 3]]},
    ],
    # test/rama_cuda_test.cu:4-12 (must run without error, PD defaults) and the
    # README triangle of test/test_pytorch.py:4-6 in P mode; labels and
    # objective worked by hand from src/maximum_matching_vertex_based.cu:76-137
    # and src/rama_cuda.cu:88-158.
    "solve": [
        {"name": "rama_cuda_test", "i": [0, 0, 1], "j": [1, 2, 2], "costs": [1.5, -3.0, 1.0], "preset": "PD"},
        {"name": "readme_triangle_P", "i": [0, 1, 2], "j": [1, 2, 0], "costs": [1.1, -2.0, 3.0], "preset": "P",
         "labels": [0, 1, 0], "objective": -0.9},
    ],
}


def main():
    path = os.path.join(HERE, "known_answers.json")
    with open(path, "w") as f:
        json.dump(CASES, f, indent=1)
    print(path)


if __name__ == "__main__":
    main()
