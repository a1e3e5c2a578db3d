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


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
def test_aggmis_hierarchy(fem, gpu, oracle, meshgen, precision):
    """aggregatorType 3 (randMisParameters 20027036: MIS depth 2, sizes
    [27, 36]): the level-0 aggregates of the hierarchy equal the oracle's
    pipeline on the same graph and seed (device MIS + first-aggregated-
    neighbour allocation without runt removal, then the merge-split
    conditioner), at most 10 % of them lie outside [27, 36], and PCG on this
    hierarchy reaches the direct solution."""
    import scipy.sparse.linalg as spla
    report(f"[gpu] aggregatorType 3 hierarchy, precision={precision}")
    xyz, tets = meshgen.cube_mesh(21, seed=3)
    kappa = meshgen.lognormal_kappa(3, tets.shape[0], 1.0)
    s = fem.FEMSolver.from_arrays(3, xyz, tets, kappa=kappa, mass_coeff=1e-2)
    s.set_params(solverType=1, tolerance=1e-10, maxIters=300, precision=precision, smootherType=1,
                 preInnerIters=3, postInnerIters=3, topSize=64, aggregatorType=3, randMisParameters=20027036,
                 strengthThreshold=0.0, seed=0x5EED)
    s.setup()
    rp, ci, v, agg = s.level_matrix(0)
    n = rp.size - 1
    rows = np.repeat(np.arange(n), np.diff(rp))
    keep = ci != rows
    ap = np.concatenate([[0], np.cumsum(np.add.reduceat(keep, rp[:-1]))]).astype(np.int32)
    aj = ci[keep].astype(np.int32)
    a0, n0 = oracle.aggregate(ap, aj, depth=2, min_size=0, rule=0, seed=0x5EED)
    want, nw = oracle.condition(ap, aj, a0, n0, 27, 36)
    sizes = np.bincount(agg)
    report(f"      level 0: {n} rows -> {sizes.size} aggregates of {sizes.min()}..{sizes.max()}, "
           f"{s.num_levels()} levels")
    assert sizes.size == nw and np.array_equal(agg, want)
    assert np.sum((sizes < 27) | (sizes > 36)) <= 0.1 * nw
    b = meshgen.gaussian_rhs(3, n)
    x, info = s.solve(b)
    A_user = s.get_matrix_scipy()
    xd = spla.spsolve(A_user.tocsc(), b)
    e = np.linalg.norm(x - xd) / np.linalg.norm(xd)
    report(f"      PCG: {info['iterations']} iterations, rel error {e:.1e}")
    assert info["status"] == 0 and e <= 1e-8
    s.close()


@pytest.mark.gpu
@pytest.mark.parametrize("agg_type", [4, 5])
def test_flood_fill_hierarchy(fem, gpu, oracle, meshgen, agg_type):
    """aggregatorType 4 and 5 (randMisParameters 20027036: flood-fill MIS of
    depth 2, sizes [27, 36]): the level-0 aggregates of the hierarchy equal the
    oracle's flood-fill MIS and nearest-root aggregation from the start node the
    library derives from the seed, merge-split conditioned for type 4 only, and
    PCG on this hierarchy reaches the direct solution."""
    import ctypes as C
    import scipy.sparse.linalg as spla
    report(f"[gpu] aggregatorType {agg_type} hierarchy")
    xyz, tets = meshgen.cube_mesh(21, seed=3)
    kappa = meshgen.lognormal_kappa(3, tets.shape[0], 1.0)
    s = fem.FEMSolver.from_arrays(3, xyz, tets, kappa=kappa, mass_coeff=1e-2)
    s.set_params(solverType=1, tolerance=1e-10, maxIters=300, precision=1, smootherType=1,
                 preInnerIters=3, postInnerIters=3, topSize=64, aggregatorType=agg_type,
                 randMisParameters=20027036, strengthThreshold=0.0, seed=0x5EED)
    s.setup()
    rp, ci, v, agg = s.level_matrix(0)
    n = rp.size - 1
    rows = np.repeat(np.arange(n), np.diff(rp))
    keep = ci != rows
    ap = np.concatenate([[0], np.cumsum(np.add.reduceat(keep, rp[:-1]))]).astype(np.int32)
    aj = ci[keep].astype(np.int32)
    roots = np.empty(n, dtype=np.int32)
    lagg = np.empty(n, dtype=np.int32)
    start, na = C.c_int64(-1), C.c_int64(0)
    assert fem.lib().fem_flood_fill_aggregate(n, ap.ctypes.data, aj.ctypes.data, 2, 0x5EED, C.byref(start),
                                              roots.ctypes.data, lagg.ctypes.data, C.byref(na)) == 0
    ro, want, nw = oracle.flood_fill(ap, aj, 2, start.value)
    if agg_type == 4:
        want, nw = oracle.condition(ap, aj, want, nw, 27, 36)
    sizes = np.bincount(agg)
    report(f"      level 0: {n} rows -> {sizes.size} aggregates of {sizes.min()}..{sizes.max()}, "
           f"{s.num_levels()} levels")
    assert sizes.size == nw and np.array_equal(agg, want)
    b = meshgen.gaussian_rhs(3, n)
    x, info = s.solve(b)
    xd = spla.spsolve(s.get_matrix_scipy().tocsc(), b)
    e = np.linalg.norm(x - xd) / np.linalg.norm(xd)
    report(f"      PCG: {info['iterations']} iterations, rel error {e:.1e}")
    assert info["status"] == 0 and e <= 1e-8
    s.close()
