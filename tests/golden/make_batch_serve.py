"""Writes batch_serve_oracle.json: the oracle's PD result on every instance
of the two many-instance workloads, so GPU tests can check each batched or
served result without minutes of CPU time.

  batch  the 512 instances of the batched workload (n log-uniform in
         [1e3, 2e5], even k Voronoi grids, odd k Erdos-Renyi with N(0.3, 1)
         costs, seed k; instances.batch_sizes(512, 1e3, 2e5, seed=12345))
  serve  the 48-request pool of the serving workload (lognormal edge
         counts, median 5e4, sigma 1, cap 2e7; instances.serve_pool())

Each entry holds the edge count, the objective, the lower bound, the cluster
count and the sha256 of the partition (labels renumbered by first
occurrence, as make_full_size.py does).  The oracle runs in a process pool.
Run from the repository root:

    python3 tests/golden/make_batch_serve.py [--procs 8]
"""
import argparse
import json
import multiprocessing as mp
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from rama_amd import instances  # noqa: E402
from tests.golden.make_full_size import objective, partition_hash  # noqa: E402

BATCH = dict(count=512, lo=1e3, hi=2e5, seed=12345)
SERVE = dict(count=48, median=5e4, sigma=1.0, cap=2e7, seed=2024)


def _batch_job(k):
    sizes = instances.batch_sizes(BATCH["count"], BATCH["lo"], BATCH["hi"], seed=BATCH["seed"])
    i, j, c, _ = instances.batch_instance(k, int(sizes[k]))
    return ("batch", k) + _solve(i, j, c)


def _serve_job(k):
    # the pool is cheap to rebuild (about a second); every worker builds it once
    global _POOL
    if "_POOL" not in globals():
        _POOL = instances.serve_pool(**SERVE)
    i, j, c = _POOL[k]
    return ("serve", k) + _solve(i, j, c)


def _solve(i, j, c):
    from oracle import oracle as O

    t = time.time()
    lab, lb, _ = O.solve(i, j, c, O.options("PD"))
    return ({"edges": int(len(c)), "objective": objective(lab, i, j, c), "lower_bound": lb,
             "clusters": int(len(np.unique(lab))), "labels": int(len(lab)), "partition_sha256": partition_hash(lab),
             "oracle_seconds": round(time.time() - t, 2)},)


def _run(job):
    kind, k = job
    return _batch_job(k) if kind == "batch" else _serve_job(k)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--procs", type=int, default=min(8, os.cpu_count() or 1))
    a = ap.parse_args()
    from oracle import build as oracle_build

    oracle_build.build(force=False)
    # largest instances first so the pool drains evenly
    sizes = instances.batch_sizes(BATCH["count"], BATCH["lo"], BATCH["hi"], seed=BATCH["seed"])
    jobs = [("batch", int(k)) for k in np.argsort(-sizes, kind="stable")] + [("serve", k) for k in range(SERVE["count"])]
    out = {"batch": {"generator": BATCH, "preset": "PD", "instances": {}},
           "serve": {"generator": SERVE, "preset": "PD", "instances": {}}}
    t0 = time.time()
    with mp.get_context("spawn").Pool(a.procs) as pool:
        for n, (kind, k, res) in enumerate(pool.imap_unordered(_run, jobs)):
            out[kind]["instances"][str(k)] = res
            if n % 64 == 0:
                print("%d/%d done, %.0f s" % (n + 1, len(jobs), time.time() - t0), flush=True)
    for kind in out:
        inst = out[kind]["instances"]
        out[kind]["instances"] = {k: inst[k] for k in sorted(inst, key=int)}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "batch_serve_oracle.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=0, separators=(",", ":"))
    print("wrote", path, "in %.0f s" % (time.time() - t0))


if __name__ == "__main__":
    main()
