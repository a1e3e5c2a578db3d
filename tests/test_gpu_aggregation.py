"""Device aggregation (randomized MIS-2 + node allocation + runt removal)
against the oracle's restatement of the reference algorithm: bit-exact
aggregate ids for the same graph and seed, plus the MIS invariants."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import report




def _graph(oracle, meshgen, N, dim=3):
    xyz, conn = meshgen.cube_mesh(N, seed=6) if dim == 3 else meshgen.square_mesh(N, seed=6)
    rp, ci = oracle.pattern(conn, xyz.shape[0])
    keep = ci != np.repeat(np.arange(xyz.shape[0]), np.diff(rp))
    ap = np.concatenate([[0], np.cumsum(np.add.reduceat(keep, rp[:-1]))]).astype(np.int32)
    return ap, ci[keep].astype(np.int32)


@pytest.mark.gpu
@pytest.mark.parametrize("N,dim,rule,depth", [(11, 3, 0, 2), (11, 3, 1, 2), (17, 3, 0, 2), (33, 2, 1, 2),
                                              (9, 3, 0, 1)])
def test_device_aggregation_bitexact(fem, gpu, oracle, meshgen, N, dim, rule, depth):
    report(f"[gpu] aggregation N={N} dim={dim} rule={rule} depth={depth}")
    ap, aj = _graph(oracle, meshgen, N, dim)
    for seed in (1, 0x5EED):
        g_agg, g_n = fem.device_aggregate(ap, aj, depth=depth, min_size=9, rule=rule, seed=seed)
        o_agg, o_n = oracle.aggregate(ap, aj, depth=depth, min_size=9, rule=rule, seed=seed)
        assert g_n == o_n
        assert np.array_equal(g_agg, o_agg)
        # every node in exactly one aggregate, ids dense
        assert g_agg.min() == 0 and g_agg.max() == g_n - 1
        assert np.all(np.bincount(g_agg) >= 1)


def test_mis_invariants(oracle, meshgen):
    """CPU: the restated MIS is independent and maximal at distance k."""
    import scipy.sparse as sp
    ap, aj = _graph(oracle, meshgen, 9)
    n = ap.size - 1
    G = sp.csr_matrix((np.ones(aj.size), aj, ap), shape=(n, n))
    for k in (1, 2):
        mis = oracle.mis(ap, aj, k, 7)
        assert set(np.unique(mis)) <= {0, 1}
        roots = np.flatnonzero(mis == 1)
        Gk = G.copy()
        for _ in range(k - 1):
            Gk = Gk + Gk @ G
        Gk = (Gk > 0).astype(np.int8)
        # independence: no two roots within distance k
        assert Gk[roots][:, roots].nnz - np.count_nonzero(Gk[roots][:, roots].diagonal()) == 0
        # maximality: every non-root within distance k of a root
        covered = np.asarray(Gk[:, roots].sum(axis=1)).ravel() > 0
        assert np.all(covered[mis == 0])
