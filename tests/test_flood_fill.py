"""Flood-fill MIS and aggregation to the nearest root: the fine aggregation of
aggregatorType 4 (then merge-split conditioned) and 5 (as it is), AggMIS's
FloodFillMIS (src/core/aggmis/cuda/AggMIS_MIS_CPU.cu:25-140) and
AggregateToNearest (src/core/aggmis/cuda/AggMIS_Aggregation_CPU.cu:4-92).

The oracle's restatement (orc_flood_fill_mis, orc_aggregate_to_nearest) is
pinned by path graphs worked out by hand from the reference's steps and by
what the definitions fix on mesh and random graphs: no two roots within
distance k, every node within distance k of a root, aggregate ids numbered in
root order, every node in an aggregate of a nearest root, aggregates
connected.  The library's host implementation (fem_flood_fill_aggregate)
must reproduce the oracle bit for bit.  CPU only."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.csgraph as csg

from test_oracle_amg import _mesh_graph, _random_graph


def _path(n):
    rows = np.concatenate([np.arange(n - 1), np.arange(1, n)])
    cols = np.concatenate([np.arange(1, n), np.arange(n - 1)])
    G = sp.csr_matrix((np.ones(rows.size, dtype=np.int8), (rows, cols)), shape=(n, n))
    G.sort_indices()
    return G.indptr.astype(np.int32), G.indices.astype(np.int32), G


def _library(fem, ap, aj, k, seed):
    ap = np.ascontiguousarray(ap, dtype=np.int32)
    aj = np.ascontiguousarray(aj, dtype=np.int32)
    n = ap.size - 1
    roots = np.empty(n, dtype=np.int32)
    agg = np.empty(n, dtype=np.int32)
    start = C.c_int64(-1)
    na = C.c_int64(0)
    st = fem.lib().fem_flood_fill_aggregate(n, ap.ctypes.data, aj.ctypes.data, k, seed, C.byref(start),
                                            roots.ctypes.data, agg.ctypes.data, C.byref(na))
    assert st == 0, fem.lib().fem_last_error(None)
    return roots, agg, int(na.value), int(start.value)


@pytest.mark.parametrize("k,starter,roots_expected,agg_expected", [
    # k = 2 from node 0: root 0 marks 1, 2; node 3 (distance 3) is the only fringe node
    (2, 0, [0, 3, 6], [0, 0, 1, 1, 1, 2, 2]),
    # k = 2 from node 3: 2, 4 then 1, 5 are marked; 0 and 6 enter the fringe with one visit
    # each and 0, pushed first, is taken first
    (2, 3, [0, 3, 6], [0, 0, 1, 1, 1, 2, 2]),
    # k = 1: every other node
    (1, 0, [0, 2, 4, 6], [0, 0, 1, 1, 2, 2, 3]),
    # k = 0: root 5 (marks 4, 6), then the lowest undecided nodes 0 (marks 1) and 2
    # (marks 3); node 1 joins 0, the first of its equally near neighbours, and 4
    # joins 5 (distance 0) rather than 3 (distance 1)
    (0, 5, [0, 2, 5], [0, 0, 1, 1, 2, 2, 2]),
])
def test_path_graph_by_hand(oracle, k, starter, roots_expected, agg_expected):
    ap, aj, _ = _path(7)
    roots, agg, na = oracle.flood_fill(ap, aj, k, starter)
    assert np.flatnonzero(roots == 1).tolist() == roots_expected
    assert set(roots.tolist()) <= {0, 1}
    assert agg.tolist() == agg_expected and na == len(roots_expected)


def test_fringe_prefers_nodes_reached_more_often(oracle):
    """Edges 0-1, 1-2, 0-3, 3-4, 3-6, 4-6, 2-5, 5-6, 5-7; k = 1 from node 0.
    Root 0 marks 1 and 3; the fringe gets 2, then 4 and 6 (one visit each).
    Root 2 (pushed first) marks 5, and 5 reaches 6 a second time and 7 once:
    6, with two visits, is taken before 4, and its mark excludes 4; 7 is the
    last root.  Push order alone would take 4 and exclude 6: {0, 2, 4, 7}.
    Aggregation: 1 and 3 join 0, 5 joins 2 (first of its equally near
    neighbours), 4 joins 6 (distance 0) rather than 3 (distance 1)."""
    edges = [(0, 1), (1, 2), (0, 3), (3, 4), (3, 6), (4, 6), (2, 5), (5, 6), (5, 7)]
    n = 8
    r = [a for a, b in edges] + [b for a, b in edges]
    c = [b for a, b in edges] + [a for a, b in edges]
    G = sp.csr_matrix((np.ones(len(r), dtype=np.int8), (r, c)), shape=(n, n))
    G.sort_indices()
    roots, agg, na = oracle.flood_fill(G.indptr.astype(np.int32), G.indices.astype(np.int32), 1, 0)
    assert np.flatnonzero(roots == 1).tolist() == [0, 2, 6, 7]
    assert agg.tolist() == [0, 0, 1, 0, 2, 1, 2, 3] and na == 4


def _check_invariants(ap, aj, G, k, roots, agg, na):
    n = ap.size - 1
    assert set(np.unique(roots).tolist()) <= {0, 1}
    rs = np.flatnonzero(roots == 1)
    assert rs.size == na >= 1
    D = csg.shortest_path(G, unweighted=True, indices=rs)  # (roots, n)
    # independent at distance k, maximal
    Drr = D[:, rs]
    np.fill_diagonal(Drr, np.inf)
    assert Drr.min() > k
    assert D.min(axis=0).max() <= max(k, 1) or k == 0 and D.min(axis=0).max() <= 1
    # aggregates: ids numbered in root order, nearest root, connected
    assert np.array_equal(agg[rs], np.arange(rs.size))
    assert agg.min() >= 0 and agg.max() == rs.size - 1
    nearest = D.min(axis=0)
    own = D[agg, np.arange(n)]
    assert np.array_equal(own, nearest)
    rows = np.repeat(np.arange(n), np.diff(ap))
    keep = agg[rows] == agg[aj]  # edges inside an aggregate
    inner = sp.csr_matrix((np.ones(int(keep.sum())), (rows[keep], aj[keep])), shape=(n, n))
    ncomp, lab = csg.connected_components(inner, directed=False)
    assert ncomp == rs.size


@pytest.mark.parametrize("k", [0, 1, 2, 3])
def test_invariants_on_mesh_and_random_graphs(oracle, meshgen, k):
    graphs = [_mesh_graph(meshgen, oracle, 9), _mesh_graph(meshgen, oracle, 13, dim=2),
              _random_graph(5, 600, 3)]
    for ap, aj, G in graphs:
        n = ap.size - 1
        for starter in (0, n // 2, n - 1):
            roots, agg, na = oracle.flood_fill(ap, aj, k, starter)
            _check_invariants(ap, aj, G, k, roots, agg, na)


@pytest.mark.parametrize("k", [0, 1, 2, 3])
def test_library_matches_oracle(fem, oracle, meshgen, k):
    graphs = [_mesh_graph(meshgen, oracle, 9), _mesh_graph(meshgen, oracle, 15),
              _mesh_graph(meshgen, oracle, 21, dim=2), _random_graph(7, 2000, 4)]
    for ap, aj, _ in graphs:
        for seed in (0, 1, 12345):
            roots, agg, na, start = _library(fem, ap, aj, k, seed)
            assert 0 <= start < ap.size - 1
            ro, ao, nao = oracle.flood_fill(ap, aj, k, start)
            assert np.array_equal(roots, ro) and np.array_equal(agg, ao) and na == nao


def test_library_rejects_bad_graphs(fem):
    ap = np.array([0, 1, 2], dtype=np.int32)
    aj = np.array([1, 5], dtype=np.int32)  # index out of range
    n = 2
    roots = np.empty(n, dtype=np.int32)
    agg = np.empty(n, dtype=np.int32)
    na = C.c_int64(0)
    st = fem.lib().fem_flood_fill_aggregate(n, ap.ctypes.data, aj.ctypes.data, 2, 0, None, roots.ctypes.data,
                                            agg.ctypes.data, C.byref(na))
    assert st != 0
