"""Batched solving and the serving loop: every instance's result in the mix
is identical (labels and lower bound, bit for bit) to solving it alone.  The
lockstep path packs many instances into one graph but keeps each one's
thresholds, stop rules and reduction order; the worker path runs the plain
solver per instance on its own stream."""
import numpy as np
import pytest

from rama_amd import instances as I
from tests.instances_small import named

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


def test_server_staggered_arrivals_equal_alone(gpu):
    # requests arrive in waves while earlier ones are mid-solve; with one
    # lane the first runs alone, the next queued ones form a pack and later
    # waves join that pack between its rounds
    import time

    insts = mixed_instances(24, lo=500, hi=20000, seed=11)
    opts = gpu.options("PD", verbose=0)
    alone = _alone(gpu, insts, opts)
    with gpu.Server(workers=1, opts=opts) as srv:
        tickets = []
        for w in range(4):
            tickets += [srv.submit(*insts[q]) for q in range(6 * w, 6 * w + 6)]
            time.sleep(0.003)
        res = [srv.wait(t) for t in tickets]
    _assert_same(alone, [r[:2] for r in res])


@pytest.mark.parametrize("preset", ["P", "D"])
def test_server_modes_equal_alone(gpu, preset):
    insts = mixed_instances(10, seed=13)
    opts = gpu.options(preset, verbose=0)
    alone = _alone(gpu, insts, opts)
    with gpu.Server(workers=2, opts=opts) as srv:
        res = [srv.wait(t) for t in [srv.submit(*x) for x in insts]]
    _assert_same(alone, [r[:2] for r in res])


def test_server_worker_mode_equals_lockstep(gpu, monkeypatch):
    insts = mixed_instances(8, seed=17)
    opts = gpu.options("PD", verbose=0)
    with gpu.Server(workers=3, opts=opts) as srv:
        a = [srv.wait(t)[:2] for t in [srv.submit(*x) for x in insts]]
    monkeypatch.setenv("RAMA_SERVE_MODE", "workers")
    with gpu.Server(workers=3, opts=opts) as srv:
        b = [srv.wait(t)[:2] for t in [srv.submit(*x) for x in insts]]
    _assert_same(a, b)


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
    # a bad request fails alone; the ones around it are served
    i, j, c = named("grid64")
    opts = gpu.options("PD", verbose=0)
    ref = gpu.solve(i, j, c, opts)
    with gpu.Server(workers=2, opts=opts) as srv:
        t0 = srv.submit(i, j, c)
        t = srv.submit([0, -5], [1, 2], [1.0, -1.0])
        t1 = srv.submit(i, j, c)
        with pytest.raises(gpu.RamaError):
            srv.wait(t)
        for tk in (t0, t1):
            lab, lb, _, _ = srv.wait(tk)
            assert lb == ref[1] and np.array_equal(lab, ref[0])


# ---- the lockstep (packed) batch path: every instance of the batch shares one
# graph and each round's launches; results must still equal solving alone.

def _alone(gpu, insts, opts):
    return [gpu.solve(i, j, c, opts)[:2] for (i, j, c) in insts]


def _assert_same(alone, mixed):
    for q, ((la, lba), (lm, lbm)) in enumerate(zip(alone, mixed)):
        assert lba == lbm, (q, lba, lbm)
        assert np.array_equal(la, lm), q


@pytest.mark.parametrize("preset", ["P", "PD", "PD+", "D"])
def test_packed_modes_equal_alone(gpu, preset):
    from tests.instances_small import named
    insts = mixed_instances(10, seed=21) + [named("grid64"), named("er800_dense")]
    opts = gpu.options(preset, verbose=0)
    _assert_same(_alone(gpu, insts, opts), gpu.solve_batch(insts, opts, workers=4))


@pytest.mark.parametrize("preset", ["PD", "PD+"])
def test_packed_hubs_equal_alone(gpu, preset):
    # supernode-like hubs among ordinary instances: the heavy-pair paths of
    # the packed cycle search (per-lane probing, whole-wave intersections)
    # write into each instance's own triangle region
    from tests.instances_small import hub_graph
    insts = [hub_graph(1200 + 100 * k, 8 + k, 150 + 40 * k, seed=30 + k) for k in range(4)]
    insts += mixed_instances(6, seed=23)
    opts = gpu.options(preset, verbose=0)
    _assert_same(_alone(gpu, insts, opts), gpu.solve_batch(insts, opts, workers=4))


def test_packed_duplicates_gaps_and_tiny(gpu):
    from tests.instances_small import named, random_graph
    rng = np.random.default_rng(5)
    gap_i, gap_j, gap_c = random_graph(400, 1500, seed=13)
    # unused ids below the largest one, and a single-edge instance
    insts = [named("random_dup"), (gap_i * 3, gap_j * 3, gap_c), (np.array([0]), np.array([5]), np.array([-1.0], np.float32)),
             (np.array([0, 1, 0]), np.array([1, 2, 2]), np.array([1.5, 2.0, -3.0], np.float32))]
    insts += [random_graph(int(n), int(3 * n), seed=int(s)) for n, s in zip(rng.integers(50, 300, 6), range(6))]
    opts = gpu.options("PD", verbose=0)
    _assert_same(_alone(gpu, insts, opts), gpu.solve_batch(insts, opts, workers=3))


def test_packed_equals_per_instance_path(gpu, monkeypatch):
    insts = mixed_instances(16, seed=31)
    opts = gpu.options("PD", verbose=0)
    packed = gpu.solve_batch(insts, opts, workers=4)
    monkeypatch.setenv("RAMA_BATCH_MODE", "workers")
    per_instance = gpu.solve_batch(insts, opts, workers=4)
    _assert_same(per_instance, packed)


def test_packed_forest_from_start(gpu):
    # crossover ratio above 1: every instance uses the spanning forest from
    # round 0, handed to the workers right after the packed initial dual
    insts = mixed_instances(8, seed=41)
    opts = gpu.options("PD", verbose=0, matching_thresh_crossover_ratio=1.5)
    _assert_same(_alone(gpu, insts, opts), gpu.solve_batch(insts, opts, workers=2))


def test_packed_reports_bad_instance(gpu):
    insts = mixed_instances(3, seed=51) + [([0, -5], [1, 2], [1.0, -1.0])]
    with pytest.raises(gpu.RamaError):
        gpu.solve_batch(insts, gpu.options("PD", verbose=0), workers=2)
