"""Message-passing loops replayed from one captured HIP graph against the
same loops launched kernel by kernel (RAMA_MP_GRAPH=0): labels and lower
bounds identical bit for bit, for single solves and the lockstep batch.
Graph solves are repeated with direct solves in between, so every replay
of a fresh capture is checked."""
import numpy as np
import pytest

from rama_amd import instances as I
from tests.instances_small import named

pytestmark = pytest.mark.gpu

CASES = ["grid96x128_lr", "grid3d_24", "er3000", "er800_dense", "random_dup", "hubs"]


@pytest.mark.parametrize("preset", ["PD", "PD+", "D"])
def test_graph_replay_equals_direct(gpu, monkeypatch, preset):
    opts = gpu.options(preset, verbose=0)
    for name in CASES:
        i, j, c = named(name)
        monkeypatch.setenv("RAMA_MP_GRAPH", "0")
        ref = gpu.solve(i, j, c, opts)[:2]
        for rep in range(3):
            monkeypatch.setenv("RAMA_MP_GRAPH", "1")
            lab, lb = gpu.solve(i, j, c, opts)[:2]
            assert lb == ref[1], (name, rep)
            assert np.array_equal(lab, ref[0]), (name, rep)
            monkeypatch.setenv("RAMA_MP_GRAPH", "0")
            lab0, lb0 = gpu.solve(i, j, c, opts)[:2]
            assert lb0 == ref[1] and np.array_equal(lab0, ref[0]), (name, rep)


def test_graph_replay_equals_direct_batched(gpu, monkeypatch):
    sizes = I.batch_sizes(24, 300, 8000, seed=21)
    insts = [I.batch_instance(k, int(n))[:3] for k, n in enumerate(sizes)]
    opts = gpu.options("PD", verbose=0)
    monkeypatch.setenv("RAMA_MP_GRAPH", "0")
    ref = gpu.solve_batch(insts, opts, workers=1)
    for _ in range(2):
        monkeypatch.setenv("RAMA_MP_GRAPH", "1")
        got = gpu.solve_batch(insts, opts, workers=1)
        for (la, lba), (lg, lbg) in zip(ref, got):
            assert lba == lbg
            assert np.array_equal(la, lg)
