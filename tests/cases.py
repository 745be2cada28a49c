"""Known-answer checks shared by the oracle (CPU) and GPU test modules.

Each function takes a backend module exposing the stage functions of
include/rama_hip.h (rama_amd._capi on the GPU, oracle.oracle on the CPU).
"""
import numpy as np


def check_compute_offsets(B, golden):
    for case in golden["compute_offsets"]:
        out = B.compute_offsets(case["ids"], case["max_value"])
        assert out.tolist() == case["offsets"]


def check_compress_labels(B, golden):
    for case in golden["compress_labels"]:
        out = B.compress_labels(case["data"], case["max_label"])
        assert out.tolist() == case["labels"]
        assert len(set(out.tolist())) == case["num_unique"]


def check_connected_components(B, golden):
    for case in golden["connected_components"]:
        off, cols = case["offsets"], case["cols"]
        lab = B.connected_components(off, cols)
        assert len(set(lab.tolist())) == case["num_components"]
        for v in range(len(off) - 1):
            assert lab[v] >= 0
            for k in range(off[v], off[v + 1]):
                assert lab[cols[k]] == lab[v]


def check_conflicted_cycles(B, golden):
    for case in golden["conflicted_cycles"]:
        # dCOO(i.begin, i.end, j.begin, j.end, ...) takes i as the column ids
        r, c, v, (rows, cols) = B.dcoo_init(case["i"], case["j"], case["costs"])
        for L, want in zip((3, 4, 5), case["counts"]):
            _, _, _, raw = B.conflicted_cycles(r, c, v, rows, cols, L, case["factor"])
            assert raw == want, (case["name"], L, raw, want)


def check_message_passing(B, golden):
    for case in golden["message_passing"]:
        r, c, v, _ = B.dcoo_init(case["i"], case["j"], case["costs"])
        _, _, _, lbs0 = B.message_passing(r, c, v, case["t1"], case["t2"], case["t3"], 0, 0)
        assert abs(lbs0[0] - case["initial_lb"]) <= case["tol"]
        _, _, _, lbs = B.message_passing(r, c, v, case["t1"], case["t2"], case["t3"], case["pre_triplet_sends"],
                                         case["iterations"])
        assert abs(lbs[-1] - case["final_lb"]) <= case["tol"]


def check_matching(B, golden):
    for case in golden["matching"]:
        r, c, v, (rows, cols) = B.dcoo_init(case["i"], case["j"], case["costs"])
        _, matched = B.matching(r, c, v, rows, cols, case["mean_multiplier_mm"])
        assert matched == case["matched_vertices"]


def check_woc(B, golden):
    for case in golden["woc_mapping"]:
        r, c, v, (rows, cols) = B.dcoo_init(case["i"], case["j"], case["costs"])
        m, _ = B.woc_mapping(r, c, v, rows, cols)
        for a, b in case["separated"]:
            assert m[a] != m[b]


def check_solve(B, golden):
    for case in golden["solve"]:
        opts = B.options(case["preset"], verbose=0)
        res = B.solve(case["i"], case["j"], case["costs"], opts)
        labels = np.asarray(res[0])
        if "labels" in case:
            assert labels.tolist() == case["labels"]
            assert abs(B.objective(labels, case["i"], case["j"], case["costs"]) - case["objective"]) < 1e-6


ALL = [check_compute_offsets, check_compress_labels, check_connected_components, check_conflicted_cycles,
       check_message_passing, check_matching, check_woc, check_solve]
