"""The oracle pinned by what the mathematics fixes, independent of the
reference's code: brute-force optima, scipy's components and spanning
forests, closed-form characterisations and conservation laws.  A plausible
slip in the restatement (a dropped term, a wrong index, a swapped endpoint)
fails one of these."""
import itertools

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.csgraph import connected_components, minimum_spanning_tree

from tests.instances_small import named, random_graph


def partitions(n):
    """All set partitions of range(n) as label vectors (restricted growth strings)."""
    def rec(k, labels, m):
        if k == n:
            yield list(labels)
            return
        for l in range(m + 1):
            labels.append(l)
            yield from rec(k + 1, labels, max(m, l + 1))
            labels.pop()
    yield from rec(0, [], 0)


def brute_optimum(n, i, j, c):
    best = np.inf
    i, j, c = np.asarray(i), np.asarray(j), np.asarray(c, dtype=np.float64)
    for lab in partitions(n):
        lab = np.asarray(lab)
        best = min(best, float(c[lab[i] != lab[j]].sum()))
    return best


def tiny_graph(seed, n=7, p=0.6):
    rng = np.random.Generator(np.random.PCG64(seed))
    I, J = [], []
    for a, b in itertools.combinations(range(n), 2):
        if rng.uniform() < p:
            I.append(a)
            J.append(b)
    c = rng.normal(0.2, 1.5, size=len(I)).astype(np.float32)
    return np.array(I, np.int32), np.array(J, np.int32), c


@pytest.mark.parametrize("seed", range(12))
@pytest.mark.parametrize("preset", ["PD", "PD+", "P", "D"])
def test_bounds_sandwich_brute_force_optimum(oracle, seed, preset):
    i, j, c = tiny_graph(seed)
    n = int(max(i.max(), j.max())) + 1
    opt = brute_optimum(n, i, j, c)
    lab, lb, _ = oracle.solve(i, j, c, oracle.options(preset))
    assert lb <= opt + 1e-5
    if preset != "D":
        obj = oracle.objective(lab, i, j, c)
        assert obj >= opt - 1e-5
        assert len(lab) == n


def test_lower_bound_improves_on_trivial_bound(oracle):
    # conflicted cycles exist: the message-passing bound must beat sum(min(0, c))
    i, j, c = named("grid96x128_lr")
    r, co, v, (rows, cols) = oracle.dcoo_init(i, j, c)
    trivial = float(np.minimum(v, 0).astype(np.float64).sum())
    *_, lb = oracle.dual_solver(r, co, v, rows, cols, 5, 10)
    assert lb > trivial + 1.0


@pytest.mark.parametrize("seed", range(4))
def test_contraction_sums_and_conservation(oracle, seed):
    i, j, c = random_graph(300, 2000, seed=seed, dup=True)
    r, co, v, dims = oracle.dcoo_init(i, j, c)
    rng = np.random.Generator(np.random.PCG64(seed))
    mp = rng.integers(0, 60, size=max(dims)).astype(np.int32)
    r2, c2, v2, dims2, bad, diag = oracle.contract(r, co, v, mp)
    # brute-force per-key sums (double, then float)
    want = {}
    for a, b, x in zip(mp[r], mp[co], v):
        k = (min(a, b), max(a, b))
        want[k] = want.get(k, 0.0) + float(x)
    off = {k: s for k, s in want.items() if k[0] != k[1]}
    got = {(int(a), int(b)): float(x) for a, b, x in zip(r2, c2, v2)}
    assert set(got) == set(off)
    for k, s in off.items():
        assert got[k] == np.float32(s)
    # total cost is conserved once the self-loops are counted back
    assert abs(sum(got.values()) + diag - float(np.float64(v).sum())) < 1e-3
    assert bad == any(s < 0 for k, s in want.items() if k[0] == k[1])
    assert list(zip(r2, c2)) == sorted(zip(r2, c2)) and np.all(r2 < c2)


@pytest.mark.parametrize("seed", range(3))
def test_connected_components_match_scipy(oracle, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    n, m = 3000, 2500
    a, b = rng.integers(0, n, m), rng.integers(0, n, m)
    M = sp.coo_matrix((np.ones(2 * m), (np.r_[a, b], np.r_[b, a])), shape=(n, n)).tocsr()
    M.sum_duplicates()
    lab = oracle.connected_components(M.indptr.astype(np.int32), M.indices.astype(np.int32))
    k, ref = connected_components(M, directed=False)
    mins = np.full(k, n)
    np.minimum.at(mins, ref, np.arange(n))
    assert np.array_equal(lab, mins[ref])


@pytest.mark.parametrize("name", ["grid64", "er3000", "grid3d_24"])
def test_spanning_forest_weight_matches_scipy(oracle, name):
    i, j, c = named(name)
    r, co, v, dims = oracle.dcoo_init(i, j, c)
    fr, fc, fw = oracle.max_spanning_forest(r, co, v)
    pos = v > 0
    n = max(dims)
    M = sp.coo_matrix((-v[pos].astype(np.float64), (r[pos], co[pos])), shape=(n, n)).tocsr()
    T = minimum_spanning_tree(M)
    assert abs(-T.sum() - float(fw.astype(np.float64).sum())) < 1e-3
    # a forest: |edges| = nodes - components over the attractive graph
    k, _ = connected_components(M, directed=False)
    assert len(fw) == n - k
    kf, _ = connected_components(sp.coo_matrix((np.ones(len(fr)), (fr, fc)), shape=(n, n)), directed=False)
    assert kf == k


def _classes(r, c, v, tol=1e-4):
    thr = float(v.min()) * tol
    att = {(int(a), int(b)) for a, b, x in zip(r, c, v) if x > -thr}
    rep = [(int(a), int(b)) for a, b, x in zip(r, c, v) if x < thr]
    return att, rep


@pytest.mark.parametrize("name", ["grid64", "er800_dense"])
def test_three_cycles_closed_form(oracle, name):
    """Max length 3: exactly the triangles with one repulsive edge whose other
    two edges are attractive (characterised over all node triples)."""
    i, j, c = named(name)
    r, co, v, (rows, cols) = oracle.dcoo_init(i, j, c)
    r2, c2, v2, dims, _, _ = oracle.contract(r, co, v, (np.arange(max(rows, cols)) // 3).astype(np.int32))
    att, rep = _classes(r2, c2, v2)
    adj = {}
    for a, b in att:
        adj.setdefault(a, set()).add(b)
        adj.setdefault(b, set()).add(a)
    want = set()
    for a, b in rep:
        for m in adj.get(a, set()) & adj.get(b, set()):
            want.add(tuple(sorted((a, b, m))))
    t1, t2, t3, _ = oracle.conflicted_cycles(r2, c2, v2, dims[0], dims[1], 3, 1000.0)
    assert set(zip(t1.tolist(), t2.tolist(), t3.tolist())) == want


def test_longer_cycles_are_fan_triangulations(oracle):
    """Max length 5: every triangle shares an edge pattern of a conflicted
    cycle of length <= 5 through one repulsive edge: check on a ring of 5 and
    a ring of 6 (which must yield nothing)."""
    for n, expect in ((5, {(0, 1, 4), (1, 3, 4), (1, 2, 3)}), (6, set())):
        i = list(range(n))
        j = [(k + 1) % n for k in range(n)]
        c = [-1.0] + [1.0] * (n - 1)
        r, co, v, (rows, cols) = oracle.dcoo_init(i, j, c)
        t1, t2, t3, _ = oracle.conflicted_cycles(r, co, v, rows, cols, 5, 10.0)
        assert set(zip(t1.tolist(), t2.tolist(), t3.tolist())) == expect


@pytest.mark.parametrize("name", ["grid64", "er3000"])
def test_matching_is_a_mutual_best_matching(oracle, name):
    i, j, c = named(name)
    r, co, v, (rows, cols) = oracle.dcoo_init(i, j, c)
    m, matched = oracle.matching(r, co, v, rows, cols, 0.05)
    moved = np.nonzero(m != np.arange(len(m)))[0]
    # every moved node points to a smaller node that itself stays put
    assert np.all(m[moved] < moved) and np.all(m[m[moved]] == m[moved])
    # pairs are disjoint and each pair is joined by an attractive edge
    assert len(set(m[moved].tolist())) == len(moved) == matched // 2
    edges = {(int(a), int(b)): float(x) for a, b, x in zip(r, co, v)}
    pos = v[v >= 0].astype(np.float64)
    thr = np.float32(0.05) * np.float32(pos.sum()) / np.float32(len(pos))
    for b in moved:
        assert edges[(int(m[b]), int(b))] > thr


def test_message_passing_bound_is_monotone_and_valid(oracle):
    for seed in range(8):
        i, j, c = tiny_graph(seed, n=8, p=0.7)
        n = int(max(i.max(), j.max())) + 1
        r, co, v, (rows, cols) = oracle.dcoo_init(i, j, c)
        t1, t2, t3, _ = oracle.conflicted_cycles(r, co, v, rows, cols, 5, 10.0)
        if len(t1) == 0:
            continue
        _, _, _, lbs = oracle.message_passing(r, co, v, t1, t2, t3, 0, 8)
        assert np.all(np.diff(lbs) >= -1e-6)
        assert lbs[-1] <= brute_optimum(n, i, j, c) + 1e-5


@pytest.mark.parametrize("rep_cost", [5.0, -1.5])
def test_repeated_pair_stays_outside_the_triangles(oracle, rep_cost):
    # thrust's set_union_by_key keeps all m copies of a key repeated in the
    # first range and set_intersection_by_key matches the first of them, so a
    # repeated (row, col) pair survives message passing as two edges: the
    # first one carries the pair's triangles, the repeat keeps its own cost
    # and only adds min(0, cost) to the bound.
    base = (np.array([0, 0, 1], np.int32), np.array([1, 2, 2], np.int32), np.array([2.0, 3.0, -4.0], np.float32))
    rep = (np.array([0, 0, 0, 1], np.int32), np.array([1, 1, 2, 2], np.int32),
           np.array([2.0, rep_cost, 3.0, -4.0], np.float32))
    t = [np.array([0], np.int32), np.array([1], np.int32), np.array([2], np.int32)]
    bi, bj, bc, blbs = oracle.message_passing(*base, *t, 0, 3)
    ri, rj, rc, rlbs = oracle.message_passing(*rep, *t, 0, 3)
    assert list(ri) == [0, 0, 0, 1] and list(rj) == [1, 1, 2, 2]
    assert rc[1] == np.float32(rep_cost)
    np.testing.assert_array_equal(rc[[0, 2, 3]], bc)
    assert bc[0] != np.float32(2.0)  # the first entry did take part
    np.testing.assert_allclose(rlbs, blbs + min(0.0, rep_cost), rtol=0, atol=1e-9)


def test_partition_hash_is_relabel_invariant():
    import importlib.util
    import os

    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "make_full_size.py")
    spec = importlib.util.spec_from_file_location("make_full_size", path)
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)
    a = np.array([5, 5, 2, 7, 2, 9], np.int32)
    perm = np.array([3, 8, 1, 0, 4, 6, 2, 5, 7, 9])
    assert mk.partition_hash(a) == mk.partition_hash(perm[a])
    assert mk.partition_hash(a) != mk.partition_hash(np.array([5, 5, 2, 7, 7, 9]))
    assert mk.partition_hash(a) != mk.partition_hash(np.array([5, 5, 2, 2, 2, 9]))
