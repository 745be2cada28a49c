"""Small seeded instances for parity tests (sizes the oracle finishes in seconds)."""
import numpy as np

from rama_amd import instances as I


def random_graph(n, m, seed, neg_frac=0.4, dup=False):
    rng = np.random.Generator(np.random.PCG64(seed))
    a = rng.integers(0, n, size=3 * m)
    b = rng.integers(0, n, size=3 * m)
    ok = a != b
    lo, hi = np.minimum(a[ok], b[ok]), np.maximum(a[ok], b[ok])
    keys = lo.astype(np.int64) * n + hi
    if not dup:
        keys = np.unique(keys)
        keys = keys[rng.permutation(len(keys))[:m]]
    else:
        keys = keys[:m]
    i, j = (keys // n).astype(np.int32), (keys % n).astype(np.int32)
    sw = rng.uniform(size=len(i)) < 0.5
    i, j = np.where(sw, j, i), np.where(sw, i, j)
    c = rng.normal(0.5, 1.0, size=len(i)).astype(np.float32)
    c[rng.uniform(size=len(i)) < neg_frac] *= -1
    return i, j, c


def named(name):
    if name == "grid64":
        return I.grid2d_rectangles()[:3]
    if name == "grid96x128_lr":
        return I.grid2d_voronoi(H=96, W=128, n_seeds=12, seed=3)[:3]
    if name == "grid3d_24":
        return I.grid3d_voronoi(D=24, n_seeds=40, seed=5)[:3]
    if name == "er3000":
        return I.erdos_renyi(3000, avg_degree=8, seed=7)[:3]
    if name == "er800_dense":
        return I.erdos_renyi(800, avg_degree=24, seed=9)[:3]
    if name == "random_dup":
        return random_graph(500, 3000, seed=11, dup=True)
    if name == "hubs":
        return hub_graph(1500, 12, 300, seed=13)
    raise KeyError(name)


def hub_graph(n, hubs, fan, seed):
    """Supernode-like hubs: nodes 0..hubs-1 each joined to `fan` random
    nodes, hub pairs and random pairs besides, mixed signs.  Hub pairs
    intersect neighbour lists of hundreds of entries (the wave-cooperative
    paths of the cycle search)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    a, b = [], []
    for h in range(hubs):
        nb = rng.choice(np.arange(hubs, n), size=fan, replace=False)
        a.append(np.full(fan, h)), b.append(nb)
    hh = np.array([(x, y) for x in range(hubs) for y in range(x + 1, hubs)])
    a.append(hh[:, 0]), b.append(hh[:, 1])
    a.append(rng.integers(hubs, n, size=4 * n)), b.append(rng.integers(hubs, n, size=4 * n))
    a, b = np.concatenate(a), np.concatenate(b)
    ok = a != b
    lo, hi = np.minimum(a[ok], b[ok]), np.maximum(a[ok], b[ok])
    keys = np.unique(lo.astype(np.int64) * n + hi)
    i, j = (keys // n).astype(np.int32), (keys % n).astype(np.int32)
    c = rng.normal(0.3, 1.0, size=len(i)).astype(np.float32)
    return i, j, c


NAMES = ["grid64", "grid96x128_lr", "grid3d_24", "er3000", "er800_dense"]
