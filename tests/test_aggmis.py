"""Merge-split conditioning of aggregatorType 3 (AggMIS's MergeSplitConditioner,
src/core/aggmis/cuda/AggMIS_MergeSplitConditioner.cu, as NewMIS calls it,
src/core/cuda/ComputePermutationMethods.cu:403-407).

The oracle's restatement (oracle/fem_oracle_aggmis.c) is pinned by two path
graphs worked out by hand from the reference's steps, and by invariants on
mesh and random graphs (dense ids, every node assigned, every aggregate
connected, sizes inside the bounds for all but a few); the library's host
conditioner (fem_condition_aggregates) must reproduce the oracle bit for bit.
CPU only."""
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


def _library(fem, ap, aj, agg, nagg, lo, hi):
    ap = np.ascontiguousarray(ap, dtype=np.int32)
    aj = np.ascontiguousarray(aj, dtype=np.int32)
    out = np.array(agg, dtype=np.int32, copy=True)
    na = C.c_int64(0)
    st = fem.lib().fem_condition_aggregates(ap.size - 1, ap.ctypes.data, aj.ctypes.data, out.ctypes.data, nagg, lo,
                                            hi, C.byref(na))
    assert st == 0, fem.lib().fem_last_error(None)
    return out, int(na.value)


def test_split_of_one_long_aggregate(oracle, fem):
    """Path 0..9 as one aggregate, sizes [3, 5]: the unforced split (10 > 5
    and 10 > 2*3) takes roots 9 (farthest from node 0) and 0 (farthest from
    9), grows both one layer per round (node 4 reaches side 0-3 first, node 5
    side 6-9), giving halves of 5; the new aggregate (id 1) is the half of
    root 0.  Nothing is outsized afterwards, so no further step changes it."""
    ap, aj, _ = _path(10)
    agg = np.zeros(10, dtype=np.int32)
    want = np.array([1, 1, 1, 1, 1, 0, 0, 0, 0, 0], dtype=np.int32)
    out, na = oracle.condition(ap, aj, agg, 1, 3, 5)
    assert na == 2 and np.array_equal(out, want), out
    out2, na2 = _library(fem, ap, aj, agg, 1, 3, 5)
    assert na2 == na and np.array_equal(out2, out)


def test_merges_of_singletons_then_forced_merge_and_split(oracle, fem):
    """Path 0..9 as singletons, sizes [3, 5] (desired 4).
    Unforced merges: every pair is equally desirable (2 + 2), each aggregate
    asks its first neighbour, so the mutual pairs form from the left: (0,1),
    (2,3), ..., (8,9); next round sizes 2 (desirability 1 + 1): (0,1), (2,3)
    pair up, the last (8,9) is left -> sizes 4, 4, 2.  No unforced merge or
    merge-split is legal (2 + 4 = 6 is not > 2 * 3), so the outsized ratio is
    1/3 > 0.1: a forced split finds nothing above 5, the forced merge joins
    the 2 to its neighbour (desirability 1 / (6 - 5)) -> sizes 4, 6; the next
    forced split halves the 6 (nodes 4-6 get the new id 2, 7-9 keep 1)."""
    ap, aj, _ = _path(10)
    agg = np.arange(10, dtype=np.int32)
    want = np.array([0, 0, 0, 0, 2, 2, 2, 1, 1, 1], dtype=np.int32)
    out, na = oracle.condition(ap, aj, agg, 10, 3, 5)
    assert na == 3 and np.array_equal(out, want), out
    out2, na2 = _library(fem, ap, aj, agg, 10, 3, 5)
    assert na2 == na and np.array_equal(out2, out)


def _check_invariants(G, agg, na, lo, hi, max_outsized):
    n = agg.size
    assert agg.min() == 0 and agg.max() == na - 1 and np.unique(agg).size == na
    sizes = np.bincount(agg, minlength=na)
    outsized = int(np.sum((sizes < lo) | (sizes > hi)))
    for a in range(na):
        members = np.flatnonzero(agg == a)
        sub = G[members][:, members]
        nc, _ = csg.connected_components(sub, directed=False)
        assert nc == 1, (a, members.size)
    assert outsized <= max_outsized * na, (outsized, na, np.sort(sizes)[:10], np.sort(sizes)[-10:])
    return sizes


@pytest.mark.parametrize("name,lo,hi", [("cube13", 20, 30), ("square33", 8, 12), ("random", 6, 10)])
def test_condition_invariants_and_library_equals_oracle(oracle, meshgen, fem, name, lo, hi):
    """Initial aggregation as aggregatorType 3 builds it (distance-2 MIS,
    first-aggregated-neighbour allocation, no runt removal), then the
    conditioner: dense ids, connected aggregates, at most 10 % outside
    [lo, hi] (the conditioner's own stopping ratio), a size spread far
    narrower than the input's; the library's conditioner equals the
    oracle's."""
    if name == "random":
        ap, aj, G = _random_graph(11, 4000, 3)
    else:
        N = int(name[4:]) if name.startswith("cube") else int(name[6:])
        ap, aj, G = _mesh_graph(meshgen, oracle, N, 3 if name.startswith("cube") else 2)
    agg0, na0 = oracle.aggregate(ap, aj, depth=2, min_size=0, rule=0, seed=77)
    out, na = oracle.condition(ap, aj, agg0, na0, lo, hi)
    s0 = np.bincount(agg0, minlength=na0)
    sizes = _check_invariants(G, out, na, lo, hi, 0.1)
    print(f"{name}: {na0} aggregates of {s0.min()}..{s0.max()} -> {na} of {sizes.min()}..{sizes.max()}")
    out2, na2 = _library(fem, ap, aj, agg0, na0, lo, hi)
    assert na2 == na and np.array_equal(out2, out)


def test_condition_rejects_bad_input(fem):
    ap, aj, _ = _path(4)
    na = C.c_int64(0)
    agg = np.array([0, 0, 2, 2], dtype=np.int32)  # id 1 unused
    L = fem.lib()
    assert L.fem_condition_aggregates(4, ap.ctypes.data, aj.ctypes.data, agg.ctypes.data, 3, 1, 2, C.byref(na)) == 1
    agg = np.array([0, 0, 1, 1], dtype=np.int32)
    assert L.fem_condition_aggregates(4, ap.ctypes.data, aj.ctypes.data, agg.ctypes.data, 2, 3, 2, C.byref(na)) == 1
