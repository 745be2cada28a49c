"""Batched solving and the serving loop: every instance's result in the mix
is identical (labels and lower bound, bit for bit) to solving it alone,
because each instance runs the unchanged plain solver on its own stream."""
import numpy as np
import pytest

from rama_amd import instances as I

pytestmark = pytest.mark.gpu


def mixed_instances(count, lo=300, hi=6000, seed=3):
    sizes = I.batch_sizes(count, lo, hi, seed=seed)
    return [I.batch_instance(k, int(n))[:3] for k, n in enumerate(sizes)]


def test_batch_equals_alone(gpu):
    insts = mixed_instances(24)
    opts = gpu.options("PD", verbose=0)
    alone = [gpu.solve(i, j, c, opts)[:2] for (i, j, c) in insts]
    for workers in (1, 8):
        mixed = gpu.solve_batch(insts, opts, workers=workers)
        for (la, lba), (lm, lbm) in zip(alone, mixed):
            assert lba == lbm
            assert np.array_equal(la, lm)


def test_server_mix_equals_alone(gpu):
    insts = mixed_instances(20, seed=5)
    opts = gpu.options("PD", verbose=0)
    alone = [gpu.solve(i, j, c, opts)[:2] for (i, j, c) in insts]
    with gpu.Server(workers=6, hbm_budget_bytes=200 << 30, opts=opts) as srv:
        order = np.random.default_rng(0).permutation(len(insts))
        tickets = {int(q): srv.submit(*insts[q]) for q in order}
        for q, t in tickets.items():
            lab, lb, lat, sms = srv.wait(t)
            assert lb == alone[q][1]
            assert np.array_equal(lab, alone[q][0])
            assert lat >= sms > 0


def test_server_budget_admits_one_at_a_time(gpu):
    # a 1-byte budget still runs every request (one in flight at a time)
    insts = mixed_instances(4, seed=9)
    opts = gpu.options("P", verbose=0)
    with gpu.Server(workers=3, hbm_budget_bytes=1, opts=opts) as srv:
        ts = [srv.submit(*x) for x in insts]
        res = [srv.wait(t) for t in ts]
    for (i, j, c), (lab, lb, _, _) in zip(insts, res):
        ref = gpu.solve(i, j, c, opts)
        assert np.array_equal(lab, ref[0]) and lb == ref[1]


def test_server_reports_errors(gpu):
    with gpu.Server(workers=2, opts=gpu.options("PD", verbose=0)) as srv:
        t = srv.submit([0, -5], [1, 2], [1.0, -1.0])
        with pytest.raises(gpu.RamaError):
            srv.wait(t)
