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
                solver's stream over the timed steps (the warm-up steps time
                every kernel to pick the dominant one and fill `kernels`;
                the timed steps record events around that kernel only);
  batched       the 512-instance batched workload (n log-uniform in [1e3,
                2e5], Voronoi grids and Erdos-Renyi graphs, PD): solves/s of
                one packed batch vs the same instances solved one at a time
                by the plain host entry, and whether every result is identical;
  serving       Poisson arrivals of PD requests (lognormal edge counts,
                median 5e4, sigma 1, cap 2e7) at 0.5, 0.9 and 1.5 x the plain
                loop's measured capacity: solves/s and p50/p99 latency of the
                serving loop (16 lanes, 200 GB HBM budget) against the plain
                one-request-at-a-time loop, and whether results are identical;
  configs       the same grid in D mode (lower bound only) and the 192^3
                3D grid in PD+ (21,123,072 edges), each through the same
                device entry: ms per solve, edges/s, objective, bound;
  cpu_baseline  the CPU restatement (oracle, 1 thread) timed on a bounded
                512x1024 crop of the same generator, plus the 64x64 grid
                workload in full.
"""
import argparse
import ctypes
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
    ap.add_argument("--workload", choices=["grid2d", "grid3d"], default="grid2d",
                    help="grid2d: H x W grid with long-range offsets, 2000 seeds; "
                         "grid3d: D^3 grid, 6-neighbourhood, 20000 seeds (run with --preset PD+)")
    ap.add_argument("--D", type=int, default=192)
    ap.add_argument("--mode", choices=["plain", "batch", "serve", "copy", "tune"], default="plain",
                    help="plain: the default line; batch: 512 mixed instances batched vs one at a time; "
                         "serve: Poisson arrivals against the serving loop and the plain loop; "
                         "copy: one 1 GiB device copy (calibrates HBM byte counters); "
                         "tune: measure the launch-shape knobs and write the tuning table")
    ap.add_argument("--tune-out", default=os.path.join(ROOT, "rama_amd", "tuning", "gfx950.table"),
                    help="tune mode: where the table is written")
    ap.add_argument("--count", type=int, default=512, help="instances in batch mode")
    ap.add_argument("--batched-only", action="store_true",
                    help="batch mode: run the batched solve alone (for profiling), no one-at-a-time loop")
    ap.add_argument("--workers", type=int, default=16)
    ap.add_argument("--requests", type=int, default=160, help="requests per offered load in serve mode")
    ap.add_argument("--kernels", type=int, default=8, help="profiled kernels listed on the line")
    ap.add_argument("--no-batch", action="store_true", help="plain mode: leave out the batched block")
    ap.add_argument("--no-serve", action="store_true", help="plain mode: leave out the serving block")
    ap.add_argument("--no-configs", action="store_true",
                    help="plain mode: leave out the 3D PD+ grid and the D-mode blocks")
    return ap.parse_args()


def config_block(name, gen, preset, steps=2):
    """One more single-instance configuration, solved through the device
    entry like the main line (inputs resident, one warm-up solve, `steps`
    timed solves): ms per solve, edges/s, objective, bound, clusters."""
    import torch

    from rama_amd import _capi

    i, j, c, _ = gen()
    E = len(c)
    dev = torch.device("cuda", 0)
    di, dj, dc = (torch.from_numpy(x).to(dev) for x in (i, j, c))
    labels = torch.empty(int(max(i.max(), j.max())) + 2, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream().cuda_stream
    opts = _capi.options(preset, verbose=0)

    def solve():
        return _capi.solve_device(di.data_ptr(), dj.data_ptr(), dc.data_ptr(), E, labels.data_ptr(), labels.numel(),
                                  opts, device=0, stream=stream)

    solve()
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(steps):
        n_lab, lb, _, _ = solve()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t) / steps
    out = {"workload": name, "mode": preset, "edges": E, "steps": steps, "warmup": 1, "ms_per_solve": dt * 1e3,
           "edges_per_s": E / dt, "lower_bound": lb}
    if n_lab > 0:
        lab = labels[:n_lab].cpu().numpy()
        out["objective"] = _capi.objective(lab, i, j, c)
        out["clusters"] = int(len(np.unique(lab)))
    del di, dj, dc, labels
    torch.cuda.empty_cache()
    _capi.lib().rama_release_cached(0)
    return out


def cpu_baseline():
    """Oracle (serial C++ restatement, one thread) on a bounded crop of the
    same workload (a quarter of the grid at the same cluster density), and
    the 64x64 grid workload in full."""
    from oracle import oracle as O
    from rama_amd import instances

    i, j, c, _ = instances.grid2d_voronoi(H=512, W=1024, n_seeds=500, seed=0)
    t = time.perf_counter()
    O.solve(i, j, c, O.options("PD"))
    dt = time.perf_counter() - t
    i0, j0, c0, _ = instances.grid2d_rectangles(seed=0)
    reps = 20
    t = time.perf_counter()
    for _ in range(reps):
        O.solve(i0, j0, c0, O.options("PD"))
    dt0 = (time.perf_counter() - t) / reps
    out = {"value": len(c) / dt, "unit": "edges/s", "cores": 1, "kind": "port",
           "sample": "PD solve of a 512x1024 crop (%d edges, 500 Voronoi seeds) of the same generator, %.1f s"
                     % (len(c), dt),
           "host_cpus": os.cpu_count(),
           "grid64": {"edges": len(c0), "ms_per_solve": dt0 * 1e3, "edges_per_s": len(c0) / dt0}}
    rec = full_size_recorded()
    if rec:
        out["full_size_recorded"] = rec
    return out


def full_size_recorded():
    """The oracle's full-size solves, timed once when their fixture was
    written (too long to repeat in a bench run; not measured in this run)."""
    path = os.path.join(ROOT, "tests", "golden", "full_size_oracle.json")
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        rec = json.load(f)
    return {k: {"edges": v["edges"], "seconds": v["oracle_seconds"], "edges_per_s": v["edges"] / v["oracle_seconds"]}
            for k, v in rec.items() if "oracle_seconds" in v}


def copy_peak(torch, nbytes=2 << 30, reps=5):
    """Measured device-to-device copy rate in GB/s (read + write bytes)."""
    src = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    dst.copy_(src)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    gbs = 2.0 * nbytes * reps / (e0.elapsed_time(e1) * 1e-3) / 1e9
    del src, dst
    torch.cuda.empty_cache()
    return gbs


def copy_mode():
    """One copy of a 1 GiB buffer: under rocprofv3 --pmc its FETCH_SIZE and
    WRITE_SIZE (KiB) against the known 1 GiB each give the counters' scale."""
    import torch

    src = torch.empty(1 << 28, dtype=torch.float32, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    return {"mode": "copy", "bytes_read": src.numel() * 4, "bytes_written": dst.numel() * 4,
            "kib_each": src.numel() * 4 / 1024}


def _batch_instances(count):
    from rama_amd import instances

    sizes = instances.batch_sizes(count, 1e3, 2e5, seed=12345)
    return [instances.batch_instance(k, int(n))[:3] for k, n in enumerate(sizes)]


def batched_block(count=512, workers=16, batched_only=False):
    """512 instances (n log-uniform in [1e3, 2e5]; even: Voronoi grids, odd:
    Erdos-Renyi), batched on one GPU (lockstep packed solve; `workers`
    threads continue the instances that switch to the spanning-forest
    contraction) vs the same instances solved one at a time by the plain
    host entry.  Both include the host-to-device copies."""
    from rama_amd import _capi

    insts = _batch_instances(count)
    E = int(sum(len(x[2]) for x in insts))
    opts = _capi.options("PD", verbose=0)
    _capi.solve_batch(insts[:4], opts, workers=1)  # warm up the runtime
    out = {"workload": "%d instances, n log-uniform [1e3, 2e5], Voronoi grids + Erdos-Renyi, PD" % count,
           "instances": count, "edges": E, "workers": workers}
    if not batched_only:
        for x in insts[:4]:
            _capi.solve(*x, opts)
        # the plain form, one instance after the other
        t = time.perf_counter()
        one = [_capi.solve(*x, opts)[:2] for x in insts]
        t_one = time.perf_counter() - t
    t = time.perf_counter()
    bat = _capi.solve_batch(insts, opts, workers=workers)
    t_bat = time.perf_counter() - t
    out["batched_solves_per_s"] = count / t_bat
    out["batched_s"] = t_bat
    out["total_objective"] = float(sum(_capi.objective(l, *x) for (l, _), x in zip(bat, insts)))
    if not batched_only:
        out["one_at_a_time_solves_per_s"] = count / t_one
        out["one_at_a_time_s"] = t_one
        out["speedup"] = t_one / t_bat
        out["identical_results"] = bool(all(np.array_equal(x[0], y[0]) and x[1] == y[1] for x, y in zip(one, bat)))
        out["one_at_a_time_total_objective"] = float(sum(_capi.objective(l, *x) for (l, _), x in zip(one, insts)))
    return out


def serving_block(loads=(0.5, 0.9, 1.5), requests=160, workers=16):
    """Poisson arrivals (lognormal edge counts, median 5e4, sigma 1, cap 2e7)
    at the given multiples of the plain loop's capacity, against the serving
    loop (lanes with continuous admission into lockstep packs, 200 GB HBM
    budget) and the plain one-request-at-a-time loop.  Latency runs from
    submission to labels on the host."""
    from rama_amd import _capi, instances

    pool = instances.serve_pool()
    opts = _capi.options("PD", verbose=0)
    # capacity of the plain loop: solves/s over the pool, one at a time
    with _capi.Server(workers=1, opts=opts, mode="workers") as srv:
        srv.wait(srv.submit(*pool[0]))
        t = time.perf_counter()
        for x in pool:
            srv.wait(srv.submit(*x))
        cap = len(pool) / (time.perf_counter() - t)

    def run(n_lanes, stream, mode):
        with _capi.Server(workers=n_lanes, hbm_budget_bytes=200 << 30, opts=opts, mode=mode) as srv:
            t0 = time.perf_counter()
            tickets = []
            for q, at in stream:
                d = t0 + at - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                tickets.append((q, srv.submit(*pool[q])))
            lat = []
            res = {}
            for q, tk in tickets:
                lab, lb, l_ms, _ = srv.wait(tk)
                lat.append(l_ms)
                res.setdefault(q, (lab, lb))
            wall = time.perf_counter() - t0
        lat = np.asarray(lat)
        return {"solves_per_s": len(stream) / wall, "p50_ms": float(np.percentile(lat, 50)),
                "p99_ms": float(np.percentile(lat, 99))}, res

    out = {"workload": "Poisson arrivals, lognormal edges (median 5e4, sigma 1, cap 2e7), "
                       "Voronoi grids + Erdos-Renyi, PD; %d requests per load from a %d-request pool"
                       % (requests, len(pool)),
           "lanes": workers, "hbm_budget_GB": 200, "plain_capacity_per_s": cap, "loads": {}}
    for k, f in enumerate(loads):
        stream = instances.poisson_stream(f * cap, requests, len(pool), seed=100 + k)
        served, r1 = run(workers, stream, "lockstep")
        plain, r2 = run(1, stream, "workers")
        same = all(np.array_equal(r1[q][0], r2[q][0]) and r1[q][1] == r2[q][1] for q in r1)
        obj1 = float(sum(_capi.objective(r1[q][0], *pool[q]) for q, _ in stream))
        obj2 = float(sum(_capi.objective(r2[q][0], *pool[q]) for q, _ in stream))
        out["loads"][str(f)] = {"offered_per_s": f * cap, "served": served, "plain_loop": plain,
                                "identical_results": bool(same), "objective_served": obj1, "objective_plain": obj2}
    return out


# Launch-shape candidates the tuner measures (tuning knobs of include/rama_hip.h).
TUNE_SORT_SIZES = [1 << k for k in range(14, 25)] + [24 << 20]
TUNE_SORT_KEYS = ((8, 42, "sort_path_u64"), (4, 21, "sort_path_u32"))
TUNE_SOLVE_KNOBS = {
    "pick_blocks": (1024, 2048, 4096, 8192),
    "match_blocks": (512, 1024, 2048, 4096),
    "mp_fused_max_blocks": (0, 64, 256, 1024),
    "scan_chunks": (1024, 2048, 4096, 6144, 8192),
    "cycle_chunk_pct": (100, 105, 110, 125, 150),
}


def _buckets(sizes, values):
    """Rows (max_size, value) from per-size choices: size s_k covers up to the
    geometric midpoint with the next sampled size, the last one everything
    larger; neighbouring rows with one value merge."""
    rows = []
    for k, (s, v) in enumerate(zip(sizes, values)):
        hi = int((s * sizes[k + 1]) ** 0.5) if k + 1 < len(sizes) else -1
        if rows and rows[-1][1] == v:
            rows[-1] = (hi, v)
        else:
            rows.append((hi, v))
    return rows


def tune_mode(a):
    """Measure the launch-shape knobs on this GPU and write the tuning table.

    Sorts: every path (three launches per pass, onesweep, top-digit +
    per-bucket blocks) at 512 and 1024 threads per tile, timed by
    rama_tune_time_sort on random keys of the widths the solver sorts most
    (42-bit u64 (row, col) keys of a 2M-node graph, 21-bit u32 node keys) at
    sizes 16K .. 24M pairs; each size bucket gets its fastest path, and
    threads per tile the shape fastest summed over both key types.  Grid
    bounds and chunking: coordinate descent on the PD solve time (median of
    three) of the 1024x2048 grid workload.  None of the knobs changes a
    result; the solve's partition is checked to stay the same."""
    import ctypes

    import torch

    from rama_amd import _capi, instances

    L = _capi.lib()
    L.rama_tuning_load(None)  # measure from the built-in defaults
    reps = 5

    def tsort(n, kb, bits, path, thr):
        ms = ctypes.c_double()
        _capi._check(L.rama_tune_time_sort(n, kb, bits, path, thr, reps, ctypes.byref(ms)))
        return ms.value

    report = {"sorts": {}, "solve": {}}
    rows = {}
    thr_time = {512: [0.0] * len(TUNE_SORT_SIZES), 1024: [0.0] * len(TUNE_SORT_SIZES)}
    for kb, bits, knob in TUNE_SORT_KEYS:
        choice = []
        for si, n in enumerate(TUNE_SORT_SIZES):
            times = {}
            for path in (0, 1, 2):
                if path == 2 and n > 1024 * 4096:
                    continue
                for thr in (512, 1024):
                    times[(path, thr)] = tsort(n, kb, bits, path, thr)
            best_path = min((0, 1, 2), key=lambda p: min(times.get((p, t), 1e30) for t in (512, 1024)))
            choice.append(best_path)
            for t in (512, 1024):
                thr_time[t][si] += times[(best_path, t)]
            report["sorts"]["%s n=%d" % (knob, n)] = {"%d/%d" % k: round(v, 4) for k, v in times.items()}
            progress("tune %s n=%d: path %d (%s)" % (knob, n, best_path,
                                                    ", ".join("%d/%d %.3f" % (k[0], k[1], v) for k, v in times.items())))
        rows[knob] = _buckets(TUNE_SORT_SIZES, choice)
    tiles = [(n + 4095) // 4096 for n in TUNE_SORT_SIZES]
    rows["sort_threads"] = _buckets(tiles, [512 if thr_time[512][k] < thr_time[1024][k] else 1024
                                            for k in range(len(tiles))])
    for knob, rr in rows.items():
        for ms, v in rr:
            _capi._check(L.rama_tuning_set(knob.encode(), ms, v))

    i, j, c, _ = instances.grid2d_voronoi(H=1024, W=2048, seed=0)
    dev = torch.device("cuda", 0)
    di, dj, dc = (torch.from_numpy(x).to(dev) for x in (i, j, c))
    labels = torch.empty(int(max(i.max(), j.max())) + 2, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream().cuda_stream
    opts = _capi.options("PD", verbose=0)

    def solve_ms():
        ts = []
        for _ in range(3):
            torch.cuda.synchronize()
            t = time.perf_counter()
            n_lab, _, _, _ = _capi.solve_device(di.data_ptr(), dj.data_ptr(), dc.data_ptr(), len(c), labels.data_ptr(),
                                                labels.numel(), opts, device=0, stream=stream)
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t)
        return float(np.median(ts)) * 1e3, labels[:n_lab].cpu().numpy().copy()

    base_ms, ref = solve_ms()
    base_ms, ref = solve_ms()
    for knob, cands in TUNE_SOLVE_KNOBS.items():
        res = {}
        for v in cands:
            _capi._check(L.rama_tuning_set(knob.encode(), -1, v))
            ms, lab = solve_ms()
            if not np.array_equal(lab, ref):
                raise RuntimeError("tuning knob %s=%d changed the partition" % (knob, v))
            res[v] = ms
        best = min(res, key=res.get)
        _capi._check(L.rama_tuning_set(knob.encode(), -1, best))
        rows[knob] = [(-1, best)]
        report["solve"][knob] = {str(k): round(v, 3) for k, v in res.items()}
        progress("tune %s: %d (%s)" % (knob, best, ", ".join("%d %.2f ms" % kv for kv in res.items())))
    final_ms, _ = solve_ms()
    props = torch.cuda.get_device_properties(0)
    lines = ["# rama_amd launch-shape table for %s (%s), written by `python bench.py --mode tune`."
             % (getattr(props, "gcnArchName", "gfx950").split(":")[0], props.name),
             "# Each knob: rows 'knob max_size value' in increasing max_size; -1 = no upper limit.",
             "# Sorts: fastest of (0 three-launch, 1 onesweep, 2 top-digit + buckets) x (512, 1024 threads per",
             "# tile) per size; the rest: coordinate descent on the 1024x2048 grid PD solve",
             "# (%.2f ms with the defaults, %.2f ms with this table)." % (base_ms, final_ms)]
    for k in ("sort_path_u64", "sort_path_u32", "sort_threads") + tuple(TUNE_SOLVE_KNOBS):
        for ms, v in rows[k]:
            lines.append("%s %d %d" % (k, ms, v))
    os.makedirs(os.path.dirname(os.path.abspath(a.tune_out)), exist_ok=True)
    with open(a.tune_out, "w") as f:
        f.write("\n".join(lines) + "\n")
    report.update({"table": a.tune_out, "default_ms": base_ms, "tuned_ms": final_ms, "rows": rows})
    return report


def _tuning_source():
    """The launch-shape table the library runs with (a path relative to the
    repository, or "defaults")."""
    from rama_amd import _capi

    src = _capi.lib().rama_tuning_source().decode()
    return os.path.relpath(src, ROOT) if os.path.isabs(src) else src


def progress(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py measures one GPU: the solver runs on a single MI355X; "
              "multi-GPU runs are not supported.", file=sys.stderr)
        sys.exit(2)
    if a.mode == "copy":
        print(json.dumps(copy_mode()))
        return
    if a.mode == "tune":
        print(json.dumps({"mode": "tune", "tuning": tune_mode(a)}))
        return
    # more hardware queues so the batch workers' and serving lanes' streams
    # run side by side (set before the first HIP call of the process)
    os.environ.setdefault("GPU_MAX_HW_QUEUES", "16")
    if a.mode != "plain":
        import torch  # noqa: F401  (one HIP runtime for the process)

        if a.mode == "batch":
            b = batched_block(a.count, a.workers, a.batched_only)
            print(json.dumps({"metric": "batched solves/s", "value": b["batched_solves_per_s"], "unit": "solves/s",
                              "n_gpus": 1, "higher_is_better": True, "dtype": "fp32", "data": "synthetic",
                              "vs_baseline": None, "batched": b}))
        else:
            loads = (0.5, 0.9, 1.5, 6.0)
            sv = serving_block(loads, a.requests, a.workers)
            print(json.dumps({"metric": "served solves/s + p99 latency", "unit": "solves/s", "n_gpus": 1,
                              "value": sv["loads"]["6.0"]["served"]["solves_per_s"], "higher_is_better": True,
                              "dtype": "fp32", "data": "synthetic", "vs_baseline": None, "serving": sv}))
        return
    import torch

    from rama_amd import _capi, instances, profiler

    if a.workload == "grid3d":
        i, j, c, _ = instances.grid3d_voronoi(D=a.D, seed=0)
        workload = "grid3d_voronoi %d^3, 6-neighbourhood; 20000 seeds; %s" % (a.D, a.preset)
    else:
        i, j, c, _ = instances.grid2d_voronoi(H=a.H, W=a.W, seed=0)
        workload = "grid2d_voronoi %dx%d + offsets (9,0),(0,9),(27,0),(0,27); 2000 seeds; %s" % (a.H, a.W, a.preset)
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

    # warm-up steps with every kernel timed: the per-kernel table and the
    # choice of the dominant kernel; the timed steps then record events
    # around that kernel's launches only
    profiler.only(None)
    profiler.enable(True)
    for _ in range(max(a.warmup, 1)):
        step()
    torch.cuda.synchronize()
    all_stats = profiler.read()
    profiler.enable(False)
    dom = profiler.dominant(all_stats)
    profiler.only(dom)
    profiler.enable(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        n_lab, lb, _, _ = step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / a.steps
    stats = profiler.read()
    profiler.enable(False)
    profiler.only(None)

    lab = labels[:n_lab].cpu().numpy()
    obj = _capi.objective(lab, i, j, c)
    roof = profiler.roofline(stats, HBM_PEAK_GBS, name=dom)
    if roof is not None:
        # HBM bytes per launch of the same kernel from the rocprofv3
        # FETCH_SIZE / WRITE_SIZE runs summarised under profiles/
        table = profiler.traffic_table(os.path.join(ROOT, "profiles", "traffic_pd_grid2d_1024x2048.md"))
        if a.workload == "grid2d" and a.H == 1024 and a.W == 2048 and a.preset == "PD" and roof["kernel"] in table:
            roof["traffic"] = table[roof["kernel"]]
        # practical peak: the library's streaming copy kernel (2 GiB buffers);
        # torch's device copy beside it
        gbps = ctypes.c_double()
        _capi._check(_capi.lib().rama_stream_copy_gbps(2 << 30, 5, ctypes.byref(gbps)))
        roof["measured_copy_peak"] = gbps.value
        roof["frac_of_copy_peak"] = round(roof["achieved"] / gbps.value, 4)
        roof["torch_copy_peak"] = copy_peak(torch)
    out = {
        # the headline metric: PD edges/s is `value`; the served solves/s and
        # p99 latency against the sequential loop are in the serving block
        "metric": ("PD edges/s; served solves/s + p99 latency vs sequential at equal objective"
                   if a.preset == "PD" and not a.no_serve else "%s edges/s" % a.preset),
        "value": E / dt,
        "unit": "edges/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": max(a.warmup, 1),
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": workload,
                   "edges": E, "nodes": n_nodes, "mode": a.preset, "objective": obj, "lower_bound": lb,
                   "clusters": int(len(np.unique(lab)))},
        "roofline": roof,
        "kernels": profiler.top(all_stats, a.kernels),
        "tuning": _tuning_source(),
    }
    progress("plain PD: %.2f ms per solve" % (dt * 1e3))
    del di, dj, dc, labels
    torch.cuda.empty_cache()
    _capi.lib().rama_release_cached(0)
    if not a.no_configs:
        out["configs"] = [
            config_block("grid2d_voronoi 1024x2048 + offsets (9,0),(0,9),(27,0),(0,27); 2000 seeds; D (lower bound only)",
                         lambda: instances.grid2d_voronoi(H=1024, W=2048, seed=0), "D", steps=3),
            config_block("grid3d_voronoi 192^3, 6-neighbourhood; 20000 seeds; PD+",
                         lambda: instances.grid3d_voronoi(D=192, seed=0), "PD+", steps=2),
        ]
        progress("configs: D %.2f ms, 3D PD+ %.2f ms" % (out["configs"][0]["ms_per_solve"],
                                                         out["configs"][1]["ms_per_solve"]))
    if not a.no_batch:
        out["batched"] = batched_block(512, a.workers)
        progress("batched: %.1f solves/s vs %.1f one at a time" % (out["batched"]["batched_solves_per_s"],
                                                                   out["batched"]["one_at_a_time_solves_per_s"]))
    if not a.no_serve:
        out["serving"] = serving_block((0.5, 0.9, 1.5), a.requests, a.workers)
        progress("serving done")
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
