"""The many-instance workloads against the oracle.

batch: all 512 instances of the batched workload (n log-uniform in [1e3,
2e5], Voronoi grids and Erdos-Renyi graphs) solved in one packed batch;
every instance's partition, objective, cluster count and lower bound must
equal what the oracle produced for it alone (tests/golden/batch_serve_oracle.json,
written by tests/golden/make_batch_serve.py).

serve: a Poisson stream of requests drawn from the serving workload's
48-request pool (lognormal edge counts, median 5e4) submitted to the
serving loop at its arrival times, with lanes and a 200 GB budget; every
served result must equal the oracle's for that request.

Partitions are compared exactly (sha256 of the labels renumbered by first
occurrence); lower bounds within 1e-9 relative (the GPU sums its partials in
another order than the serial oracle).
"""
import json
import os
import threading
import time

import numpy as np
import pytest

from rama_amd import instances as I
from tests.golden.make_full_size import objective, partition_hash

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def fixture():
    with open(os.path.join(ROOT, "tests", "golden", "batch_serve_oracle.json")) as f:
        return json.load(f)


def _check(inst, ref, lab, lb, what):
    i, j, c = inst
    assert len(lab) == ref["labels"], (what, len(lab), ref["labels"])
    assert partition_hash(lab) == ref["partition_sha256"], (what, objective(lab, i, j, c), ref["objective"])
    assert objective(lab, i, j, c) == ref["objective"], what
    assert len(np.unique(lab)) == ref["clusters"], what
    assert abs(lb - ref["lower_bound"]) <= 1e-9 * max(1.0, abs(ref["lower_bound"])), (what, lb, ref["lower_bound"])
    assert lb <= ref["objective"] + 1e-6 * abs(ref["objective"]), what


def test_batch_512_equals_oracle(gpu, fixture):
    g = fixture["batch"]["generator"]
    sizes = I.batch_sizes(g["count"], g["lo"], g["hi"], seed=g["seed"])
    insts = [I.batch_instance(k, int(n))[:3] for k, n in enumerate(sizes)]
    print("batch: %d instances, %d edges, largest %d nodes" % (len(insts), sum(len(x[2]) for x in insts), sizes.max()),
          flush=True)
    t = time.perf_counter()
    res = gpu.solve_batch(insts, gpu.options("PD", verbose=0), workers=16)
    print("batched solve %.2f s" % (time.perf_counter() - t), flush=True)
    ref = fixture["batch"]["instances"]
    assert len(res) == len(ref) == 512
    for k, (inst, (lab, lb)) in enumerate(zip(insts, res)):
        _check(inst, ref[str(k)], lab, lb, "batch instance %d (n=%d)" % (k, sizes[k]))


def test_serve_poisson_stream_equals_oracle(gpu, fixture):
    g = fixture["serve"]["generator"]
    pool = I.serve_pool(**g)
    ref = fixture["serve"]["instances"]
    stream = I.poisson_stream(rate=150.0, count=120, pool_size=len(pool), seed=7)
    opts = gpu.options("PD", verbose=0)
    with gpu.Server(workers=16, hbm_budget_bytes=200 << 30, opts=opts) as srv:
        t0 = time.perf_counter()
        tickets = []
        for q, at in stream:
            d = t0 + at - time.perf_counter()
            if d > 0:
                time.sleep(d)
            tickets.append((q, srv.submit(*pool[q])))
        lat = []
        for n, (q, tk) in enumerate(tickets):
            lab, lb, l_ms, s_ms = srv.wait(tk)
            assert l_ms >= s_ms > 0
            lat.append(l_ms)
            _check(pool[q], ref[str(q)], lab, lb, "request %d (pool %d)" % (n, q))
        wall = time.perf_counter() - t0
    print("served %d requests in %.2f s, p50 %.1f ms, p99 %.1f ms"
          % (len(tickets), wall, np.percentile(lat, 50), np.percentile(lat, 99)), flush=True)


def test_server_submit_while_waiting(gpu, fixture):
    # submitters and waiters on separate threads (ctypes releases the GIL):
    # the ticket map changes under a blocked wait; a second wait on a
    # collected ticket reports an error instead of touching freed state
    pool = I.serve_pool(**fixture["serve"]["generator"])
    small = [q for q in range(len(pool)) if len(pool[q][2]) < 40000][:6]
    ref = fixture["serve"]["instances"]
    opts = gpu.options("PD", verbose=0)
    errors = []
    with gpu.Server(workers=4, opts=opts) as srv:
        tickets = []
        lock = threading.Lock()

        def submitter(qs):
            for q in qs:
                t = srv.submit(*pool[q])
                with lock:
                    tickets.append((q, t))
                time.sleep(0.002)

        first = srv.submit(*pool[small[0]])
        waiter_out = {}

        def waiter():
            try:
                waiter_out["r"] = srv.wait(first)
            except Exception as e:  # noqa: BLE001
                errors.append(e)

        th = [threading.Thread(target=waiter)] + [threading.Thread(target=submitter, args=(small * 3,)) for _ in range(3)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        assert not errors, errors
        lab, lb, _, _ = waiter_out["r"]
        _check(pool[small[0]], ref[str(small[0])], lab, lb, "first request")
        with pytest.raises(gpu.RamaError):
            srv._caps[first] = len(lab) + 1
            srv.wait(first)
        # two waiters blocked on one ticket: one takes the result, the other
        # gets an error
        big = max(range(len(pool)), key=lambda q: len(pool[q][2]))
        t2 = srv.submit(*pool[big])
        cap = srv._caps.pop(t2)
        outs = []

        def raw_wait():
            lab2 = np.zeros(cap, np.int32)
            n2, lb2, a, b = (gpu.ctypes.c_int64(), gpu.ctypes.c_double(), gpu.ctypes.c_double(),
                             gpu.ctypes.c_double())
            rc = gpu.lib().rama_server_wait(srv._h, t2, gpu._p(lab2, gpu.I32P), cap, gpu.ctypes.byref(n2),
                                            gpu.ctypes.byref(lb2), gpu.ctypes.byref(a), gpu.ctypes.byref(b))
            with lock:
                outs.append((rc, lab2[: n2.value].copy(), lb2.value))

        ws = [threading.Thread(target=raw_wait) for _ in range(2)]
        for t in ws:
            t.start()
        for t in ws:
            t.join()
        rcs = sorted(o[0] for o in outs)
        assert rcs[0] == 0 and rcs[1] != 0, rcs
        ok = [o for o in outs if o[0] == 0][0]
        _check(pool[big], ref[str(big)], ok[1], ok[2], "double-waited request")
        for q, t in tickets:
            lab, lb, _, _ = srv.wait(t)
            _check(pool[q], ref[str(q)], lab, lb, "threaded request (pool %d)" % q)
