"""Symmetric CSR (to_undirected + dCOO(symmetric) + compute_row_offsets,
include/rama_utils.h:107-131, src/dCOO.cu:144-147,164-172) through the C ABI
against numpy: both directions of every edge taken, each row's neighbours
ascending and a repeated pair in input order (the stable sort of the doubled
list), offsets = row starts.  Covers the three row-offset forms (one thread
per entry; one per row when rows far outnumber entries; the spanning-forest
form) and the cost filter (is_positive_edge)."""
import numpy as np
import pytest

from rama_amd import _capi
from tests.instances_small import named

pytestmark = pytest.mark.gpu

GRAPHS = ["grid64", "grid96x128_lr", "grid3d_24", "er3000", "er800_dense", "random_dup", "hubs"]


def _sorted_edges(i, j, c):
    i, j, c = np.asarray(i), np.asarray(j), np.asarray(c, np.float32)
    a, b = np.minimum(i, j), np.maximum(i, j)
    keep = a != b
    a, b, c = a[keep], b[keep], c[keep]
    o = np.lexsort((b, a))  # stable: by row, then col, repeats in input order
    return a[o].astype(np.int32), b[o].astype(np.int32), c[o]


def _reference(r, c, w, n, keep_above=None):
    if keep_above is not None:
        m = w > np.float32(keep_above)
        r, c, w = r[m], c[m], w[m]
    rows = np.concatenate([r, c])
    cols = np.concatenate([c, r])
    ws = np.concatenate([w, w])
    o = np.lexsort((cols, rows))
    rows, cols, ws = rows[o], cols[o], ws[o]
    off = np.searchsorted(rows, np.arange(n + 1), side="left").astype(np.int32)
    return off, cols.astype(np.int32), ws


def _check(r, c, w, n, mode, keep_above=0.0):
    got = _capi.symmetric_csr(r, c, w, n, mode=mode, keep_above=keep_above)
    ref = _reference(r, c, w, n, keep_above if mode == 1 else None)
    assert np.array_equal(got[0], ref[0])
    assert np.array_equal(got[1], ref[1])
    assert np.array_equal(got[2].view(np.uint32), ref[2].view(np.uint32))


@pytest.mark.parametrize("name", GRAPHS)
def test_symmetric_csr_equals_doubled_stable_sort(gpu, name):
    r, c, w = _sorted_edges(*named(name))
    n = int(max(r.max(), c.max())) + 1
    _check(r, c, w, n, 0)  # rows with entries: one thread per entry
    _check(r, c, w, 4 * len(w) + n + 7, 0)  # rows far outnumber entries: one thread per row
    _check(r, c, w, n, 2)  # the spanning-forest form
    _check(r, c, w, n, 1, 0.0)  # attractive edges only
    _check(r, c, w, n, 1, float(np.median(w)))


def test_symmetric_csr_sparse_forest_rows(gpu):
    # a random forest whose smaller ends gather at low ids over many nodes
    rng = np.random.default_rng(5)
    n = 300_000
    child = rng.choice(np.arange(1, n), 20_000, replace=False)
    parent = rng.integers(0, 50, size=child.size)
    r, c, w = _sorted_edges(parent, child, rng.normal(size=child.size))
    for mode in (0, 2):
        _check(r, c, w, n, mode)


def test_symmetric_csr_empty(gpu):
    e = np.zeros(0, np.int32)
    off, adj, w = _capi.symmetric_csr(e, e, np.zeros(0, np.float32), 5)
    assert np.array_equal(off, np.zeros(6, np.int32)) and adj.size == 0 and w.size == 0
