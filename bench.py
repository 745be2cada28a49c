"""Benchmark: plain PD solve of the 12.47M-edge 2-D grid instance on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]

Workload: 1024x2048 grid, 4-neighbourhood plus long-range offsets (9,0),
(0,9), (27,0), (0,27) (N = 2,097,152 nodes, E = 12,469,248 edges), 2000-seed
Voronoi planted partition, logit-Beta costs with 10% noisy cross edges,
seed 0, int32 ids, fp32 costs.  One step = one full PD solve (default
options) through the device C ABI, inputs already resident in HBM.
value = E / seconds per solve (PD edges/s).

Also reported on the same JSON line:
  roofline      the dominant kernel's algorithmic HBM bytes per launch over
                its average launch time, timed with HIP events on the
                solver's stream over the timed steps;
  cpu_baseline  the CPU restatement (oracle, 1 thread) timed on a bounded
                256x1024 crop of the same generator.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--H", type=int, default=1024)
    ap.add_argument("--W", type=int, default=2048)
    ap.add_argument("--preset", default="PD")
    return ap.parse_args()


def cpu_baseline():
    """Oracle (serial C++ restatement) on a bounded crop of the same workload."""
    from oracle import oracle as O
    from rama_amd import instances

    i, j, c, _ = instances.grid2d_voronoi(H=256, W=1024, n_seeds=250, seed=0)
    t = time.perf_counter()
    O.solve(i, j, c, O.options("PD"))
    dt = time.perf_counter() - t
    return {"value": len(c) / dt, "unit": "edges/s", "cores": 1, "kind": "port",
            "sample": "PD solve of a 256x1024 crop (%d edges, 250 Voronoi seeds) of the same generator, %.1f s"
                      % (len(c), dt)}


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py measures one GPU: this build runs the solver on a single MI355X; "
              "multi-GPU runs are not part of it (see DESIGN.md for how the path would shard).", file=sys.stderr)
        sys.exit(2)
    import torch

    from rama_amd import _capi, instances, profiler

    i, j, c, _ = instances.grid2d_voronoi(H=a.H, W=a.W, seed=0)
    E = len(c)
    dev = torch.device("cuda", 0)
    di = torch.from_numpy(i).to(dev)
    dj = torch.from_numpy(j).to(dev)
    dc = torch.from_numpy(c).to(dev)
    n_nodes = int(max(i.max(), j.max())) + 1
    labels = torch.empty(n_nodes + 1, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream().cuda_stream
    opts = _capi.options(a.preset, verbose=0)
    torch.cuda.synchronize()

    def step():
        return _capi.solve_device(di.data_ptr(), dj.data_ptr(), dc.data_ptr(), E, labels.data_ptr(), labels.numel(),
                                  opts, device=0, stream=stream)

    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()
    profiler.enable(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        n_lab, lb, _, _ = step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / a.steps
    stats = profiler.read()
    profiler.enable(False)

    lab = labels[:n_lab].cpu().numpy()
    obj = _capi.objective(lab, i, j, c)
    roof = profiler.roofline(stats, HBM_PEAK_GBS)
    out = {
        "metric": "PD edges/s",
        "value": E / dt,
        "unit": "edges/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "grid2d_voronoi 1024x2048 + offsets (9,0),(0,9),(27,0),(0,27); 2000 seeds; PD",
                   "edges": E, "nodes": n_nodes, "mode": a.preset, "objective": obj, "lower_bound": lb,
                   "clusters": int(len(np.unique(lab)))},
        "roofline": roof,
        "kernels": profiler.top(stats, 8),
    }
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
