"""Seeded solves of mid-size graphs (20k-100k nodes, up to about 420k
edges) against the oracle: sizes where the sorts take the three-launch
passes and the scans several chunks, with repeated pairs and hubs, across
P/PD/PD+/D.  Each solve must reproduce the oracle's partition and bound, or
the oracle's error."""
import numpy as np
import pytest

from tests.instances_small import hub_graph, random_graph

pytestmark = pytest.mark.gpu


def _instance(seed):
    rng = np.random.Generator(np.random.PCG64(9000 + seed))
    n = int(rng.integers(20000, 100000))
    if seed % 3 == 2:
        return hub_graph(n, int(rng.integers(8, 40)), int(rng.integers(200, 2000)), seed=seed)
    return random_graph(n, int(n * rng.uniform(2.0, 4.0)), seed=seed, neg_frac=float(rng.uniform(0.2, 0.6)),
                        dup=(seed % 3 == 1))


@pytest.mark.parametrize("seed", [0, 1, 2, 3, 5, 7, 11, 17])
def test_mid_size_solve_equals_oracle(gpu, oracle, seed):
    i, j, c = _instance(seed)
    preset = ["P", "PD", "PD+", "D"][seed % 4]
    try:
        o, olb, _ = oracle.solve(i, j, c, oracle.options(preset))
    except Exception as e:
        with pytest.raises(Exception) as ge:
            gpu.solve(i, j, c, gpu.options(preset, verbose=0))
        assert str(e) in str(ge.value)
        return
    g, glb, _, _ = gpu.solve(i, j, c, gpu.options(preset, verbose=0))
    assert len(g) == len(o)
    _, ig = np.unique(g, return_inverse=True)
    _, io = np.unique(o, return_inverse=True)
    pairs = set(zip(ig.tolist(), io.tolist()))
    assert len(pairs) == len(set(ig.tolist())) == len(set(io.tolist())), "partition differs"
    assert abs(glb - olb) <= 1e-9 * max(1.0, abs(olb))
