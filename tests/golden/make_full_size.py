"""Writes full_size_oracle.json: the oracle's result on the two full-size
single-instance workloads (the 1024x2048 grid with long-range offsets in PD,
the 192^3 grid in PD+).  Each entry holds the objective, the lower bound,
the cluster count and a hash of the partition (labels renumbered by first
occurrence), so a GPU test can check the whole partition without the
oracle's minutes of CPU time.  The D-mode entry (lower bound only) holds
the bound and no partition.  Run from the repository root:

    python3 tests/golden/make_full_size.py [workload ...]

Named workloads are (re)computed and merged into the existing file; no name
recomputes all of them.
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from rama_amd import instances  # noqa: E402


def partition_hash(labels):
    """sha256 of the labels renumbered 0, 1, ... in order of first occurrence."""
    lab = np.asarray(labels, np.int64)
    _, first, inv = np.unique(lab, return_index=True, return_inverse=True)
    rank = np.empty(len(first), np.int64)
    rank[np.argsort(first, kind="stable")] = np.arange(len(first))
    return hashlib.sha256(rank[inv].astype(np.int32).tobytes()).hexdigest()


def objective(labels, i, j, c):
    lab = np.asarray(labels)
    cut = lab[i] != lab[j]
    return float(np.asarray(c, np.float64)[cut].sum())


WORKLOADS = {
    "grid2d_voronoi_1024x2048_PD": (lambda: instances.grid2d_voronoi(H=1024, W=2048, seed=0), "PD"),
    "grid3d_voronoi_192_PDplus": (lambda: instances.grid3d_voronoi(D=192, seed=0), "PD+"),
    "grid2d_voronoi_1024x2048_D": (lambda: instances.grid2d_voronoi(H=1024, W=2048, seed=0), "D"),
}


def main():
    from oracle import oracle as O

    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "full_size_oracle.json")
    names = sys.argv[1:] or list(WORKLOADS)
    out = {}
    if os.path.exists(path):
        with open(path) as f:
            out = json.load(f)
    for name in names:
        gen, preset = WORKLOADS[name]
        i, j, c, _ = gen()
        t = time.time()
        lab, lb, _ = O.solve(i, j, c, O.options(preset))
        if preset == "D":
            out[name] = {"preset": preset, "edges": int(len(c)), "lower_bound": lb,
                         "oracle_seconds": round(time.time() - t, 1)}
        else:
            out[name] = {"preset": preset, "edges": int(len(c)), "objective": objective(lab, i, j, c),
                         "lower_bound": lb, "clusters": int(len(np.unique(lab))),
                         "partition_sha256": partition_hash(lab), "oracle_seconds": round(time.time() - t, 1)}
        print(name, out[name], flush=True)
    with open(path, "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
