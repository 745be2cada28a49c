"""Every way the cycle search can count and emit gives the oracle's
triangles: the 2-hop tables (5-cycles) and wave-wide sets (4-cycles) against
per-thread counting (RAMA_CYCLE_TABLES=0), and the per-expansion emit
against the queued one (RAMA_EMIT_FORM), on graphs with triangles, hubs,
repeated edges, and budgets that cut the enumeration short."""
import numpy as np
import pytest

from rama_amd import instances as I
from tests.instances_small import named

pytestmark = pytest.mark.gpu

GRAPHS = ["grid96x128_lr", "grid3d_24", "er3000", "er800_dense", "random_dup", "hubs"]
PATHS = [("1", "dense"), ("1", "queue"), ("0", "dense"), ("0", "queue")]


def _contracted(oracle, name, groups):
    i, j, c = named(name)
    r, co, v, dims = oracle.dcoo_init(i, j, c)
    n = max(dims)
    rng = np.random.Generator(np.random.PCG64(7))
    mp = rng.integers(0, max(2, n // groups), size=n).astype(np.int32)
    r2, c2, v2, dims2, _, _ = oracle.contract(r, co, v, mp)
    return r2, c2, v2, dims2


@pytest.mark.parametrize("name", GRAPHS)
@pytest.mark.parametrize("groups", [1, 3])
def test_cycle_paths_equal_oracle(gpu, oracle, monkeypatch, name, groups):
    if groups == 1:
        i, j, c = named(name)
        r, co, v, (rows, cols) = oracle.dcoo_init(i, j, c)
    else:
        r, co, v, (rows, cols) = _contracted(oracle, name, groups)
    for max_len in (4, 5):
        for factor in (2.0, 0.3):
            o = oracle.conflicted_cycles(r, co, v, rows, cols, max_len, factor)
            for tables, form in PATHS:
                monkeypatch.setenv("RAMA_CYCLE_TABLES", tables)
                monkeypatch.setenv("RAMA_EMIT_FORM", form)
                g = gpu.conflicted_cycles(r, co, v, rows, cols, max_len, factor)
                case = (name, groups, max_len, factor, tables, form)
                assert g[3] == o[3], case
                for a, b in zip(g[:3], o[:3]):
                    assert np.array_equal(a, b), case


def test_cycle_paths_batched_equal(gpu, monkeypatch):
    sizes = I.batch_sizes(16, 300, 6000, seed=31)
    insts = [I.batch_instance(k, int(n))[:3] for k, n in enumerate(sizes)]
    opts = gpu.options("PD+", verbose=0)
    ref = gpu.solve_batch(insts, opts, workers=1)
    for tables, form in PATHS:
        monkeypatch.setenv("RAMA_CYCLE_TABLES", tables)
        monkeypatch.setenv("RAMA_EMIT_FORM", form)
        got = gpu.solve_batch(insts, opts, workers=1)
        for (la, lba), (lg, lbg) in zip(ref, got):
            assert lba == lbg, (tables, form)
            assert np.array_equal(la, lg), (tables, form)
