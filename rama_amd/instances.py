"""Seeded synthetic multicut instances (positive cost = attractive).

Edge costs come from a planted partition: each edge gets a probability p of
being cut (Beta-distributed, depending on whether its ends share a cluster),
p is clipped to [1e-3, 1 - 1e-3] and the cost is log((1 - p) / p) computed in
fp64 and stored as fp32.  Node ids are row-major grid indices; every edge is
emitted once with i < j.

  grid2d_rectangles  64x64 4-grid, 16 planted rectangles
  grid2d_voronoi     2-D grid + long-range offsets, Voronoi clusters
  grid3d_voronoi     3-D 6-grid, Voronoi clusters, mixture costs
  erdos_renyi        G(n, m) without self-loops or duplicates, N(0.3, 1) costs
  batch_sizes / batch_instance   512 mixed instances (grids and ER), log-uniform sizes
  serve_pool / poisson_stream    serving requests: lognormal edge counts, Poisson arrivals
"""
import math

import numpy as np


def _costs(p):
    p = np.clip(p, 1e-3, 1.0 - 1e-3)
    return np.log((1.0 - p) / p).astype(np.float32)


def _grid_edges(shape, offsets):
    """Edges (i, j) of an N-d grid for the given non-negative offsets, i < j."""
    shape = tuple(shape)
    idx = np.arange(int(np.prod(shape)), dtype=np.int64).reshape(shape)
    I, J = [], []
    for off in offsets:
        src = tuple(slice(0, s - o) for s, o in zip(shape, off))
        dst = tuple(slice(o, s) for s, o in zip(shape, off))
        I.append(idx[src].ravel())
        J.append(idx[dst].ravel())
    return np.concatenate(I), np.concatenate(J)


def _voronoi_labels(shape, n_seeds, rng):
    from scipy.spatial import cKDTree

    seeds = rng.uniform(0.0, 1.0, size=(n_seeds, len(shape))) * np.asarray(shape, dtype=np.float64)
    tree = cKDTree(seeds)
    grids = np.meshgrid(*[np.arange(s, dtype=np.float64) for s in shape], indexing="ij")
    pts = np.stack([g.ravel() for g in grids], axis=1)
    labels = np.empty(len(pts), dtype=np.int32)
    step = 1 << 20
    for b in range(0, len(pts), step):
        _, lab = tree.query(pts[b:b + step], k=1)
        labels[b:b + step] = lab
    return labels


def grid2d_rectangles(H=64, W=64, cuts=3, seed=0):
    """64x64 4-connected grid, 4x4 planted rectangles, Beta(2,8)/Beta(8,2) costs."""
    rng = np.random.Generator(np.random.PCG64(seed))
    rc = np.sort(rng.choice(np.arange(1, H), size=cuts, replace=False))
    cc = np.sort(rng.choice(np.arange(1, W), size=cuts, replace=False))
    rb = np.searchsorted(rc, np.arange(H), side="right")
    cb = np.searchsorted(cc, np.arange(W), side="right")
    lab = (rb[:, None] * (cuts + 1) + cb[None, :]).ravel()
    i, j = _grid_edges((H, W), [(0, 1), (1, 0)])
    same = lab[i] == lab[j]
    p = np.where(same, rng.beta(2, 8, size=len(i)), rng.beta(8, 2, size=len(i)))
    return i.astype(np.int32), j.astype(np.int32), _costs(p), lab


def grid2d_voronoi(H=1024, W=2048, offsets=((0, 1), (1, 0), (9, 0), (0, 9), (27, 0), (0, 27)), n_seeds=2000,
                   noise=0.10, seed=0):
    """2-D grid with long-range edges, Voronoi clusters, 10% noisy cross edges."""
    rng = np.random.Generator(np.random.PCG64(seed))
    lab = _voronoi_labels((H, W), n_seeds, rng)
    i, j = _grid_edges((H, W), offsets)
    same = lab[i] == lab[j]
    p_in = rng.beta(2, 8, size=len(i))
    p_x = rng.beta(8, 2, size=len(i))
    flip = rng.uniform(size=len(i)) < noise
    p = np.where(same, p_in, np.where(flip, p_in, p_x))
    return i.astype(np.int32), j.astype(np.int32), _costs(p), lab


def grid3d_voronoi(D=192, n_seeds=20000, seed=0):
    """3-D 6-connected grid, Voronoi clusters, mixture intra costs."""
    rng = np.random.Generator(np.random.PCG64(seed))
    shape = (D, D, D)
    lab = _voronoi_labels(shape, n_seeds, rng)
    i, j = _grid_edges(shape, [(0, 0, 1), (0, 1, 0), (1, 0, 0)])
    same = lab[i] == lab[j]
    m = len(i)
    mix = rng.uniform(size=m) < 0.7
    p_in = np.where(mix, rng.beta(2, 8, size=m), rng.beta(5, 5, size=m))
    p = np.where(same, p_in, rng.beta(8, 2, size=m))
    return i.astype(np.int32), j.astype(np.int32), _costs(p), lab


def erdos_renyi(n, avg_degree=8, mu=0.3, sigma=1.0, seed=0):
    """G(n, m = n*avg_degree/2) without self-loops or duplicates, costs ~ N(mu, sigma)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    m = n * avg_degree // 2
    keys = np.empty(0, dtype=np.int64)
    while len(keys) < m:
        need = int((m - len(keys)) * 1.1) + 16
        a = rng.integers(0, n, size=need)
        b = rng.integers(0, n, size=need)
        ok = a != b
        lo, hi = np.minimum(a[ok], b[ok]), np.maximum(a[ok], b[ok])
        keys = np.unique(np.concatenate([keys, lo * n + hi]))
    keys = keys[rng.permutation(len(keys))[:m]]
    keys.sort()
    i, j = keys // n, keys % n
    c = rng.normal(mu, sigma, size=m).astype(np.float32)
    return i.astype(np.int32), j.astype(np.int32), c, None


def small_grid_voronoi(n, seed):
    """Near-square 4-grid with about n nodes and grid2d_voronoi's cluster density."""
    H = max(2, int(math.floor(math.sqrt(n))))
    W = max(2, int(math.ceil(n / H)))
    seeds = max(2, int(round(H * W / 1049)))
    return grid2d_voronoi(H, W, offsets=((0, 1), (1, 0)), n_seeds=seeds, seed=seed)


def batch_sizes(count=512, lo=1e3, hi=2e5, seed=12345):
    rng = np.random.Generator(np.random.PCG64(seed))
    return np.round(np.exp(rng.uniform(math.log(lo), math.log(hi), size=count))).astype(np.int64)


def batch_instance(k, n):
    """Instance k of the batched workload: even k grid, odd k Erdős–Rényi; seed = k."""
    if k % 2 == 0:
        return small_grid_voronoi(int(n), seed=k)
    return erdos_renyi(int(n), avg_degree=8, seed=k)


def serve_pool(count=48, median=5e4, sigma=1.0, cap=2e7, seed=2024):
    """Request pool of the serving workload: request k has
    m = min(cap, round(lognormal(ln median, sigma))) edges; even k is a
    Voronoi grid (about m / 2 nodes), odd k Erdős–Rényi (m / 4 nodes), both
    with seed k.  Returns a list of (i, j, c)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    pool = []
    for k in range(count):
        m = int(min(cap, round(math.exp(rng.normal(math.log(median), sigma)))))
        n = max(64, m // 2 if k % 2 == 0 else m // 4)
        pool.append(batch_instance(k, n)[:3])
    return pool


def poisson_stream(rate, count, pool_size, seed):
    """Arrival schedule of the serving workload: (pool index, arrival time in
    seconds from the start) for `count` requests with exponential gaps of
    mean 1 / rate, each request drawn uniformly from the pool."""
    rng = np.random.Generator(np.random.PCG64(seed))
    seq = rng.integers(0, pool_size, size=count)
    at = np.cumsum(rng.exponential(1.0 / rate, size=count))
    return [(int(q), float(t)) for q, t in zip(seq, at)]


def config(name, **kw):
    """Instances of the five workloads by short name."""
    if name == "grid64":
        return grid2d_rectangles(**kw)
    if name == "grid2d":
        return grid2d_voronoi(**kw)
    if name == "grid3d":
        return grid3d_voronoi(**kw)
    raise KeyError(name)
