"""GPU AMG hierarchy against its definition, level by level, evaluated on the
host with scipy from the level matrices and aggregates the library exports:
aggregates partition the rows, P = (I - omega D^-1 A) T (T(i, agg(i)) = 1,
smoothedMG_amg_level.cu:321-387), A_c = P^T A P (:390-400), A_c symmetric.

At 41^3 (68,921 rows) the level-0 products take the thread-per-row sparse
product kernels and the coarse ones the wavefront-per-row kernels."""
from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse as sp

from conftest import report

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("theta", [0.0, 0.005])
@pytest.mark.parametrize("precision", [0, 1])
def test_hierarchy_matches_definition(fem, gpu, meshgen, theta, precision):
    report(f"[gpu] hierarchy vs definition 41^3 theta={theta} precision={precision}")
    xyz, tets = meshgen.cube_mesh(41, seed=2)
    kappa = meshgen.lognormal_kappa(2, tets.shape[0], 1.5)
    s = fem.FEMSolver.from_arrays(3, xyz, tets, kappa=kappa, mass_coeff=1e-2)
    omega = 0.67
    s.set_params(solverType=1, precision=precision, topSize=256, aggregationRule=1, strengthThreshold=theta,
                 proOmega=omega)
    s.setup()
    nl = s.num_levels()
    assert nl >= 3
    for l in range(nl - 1):
        rp, ci, v, agg = s.level_matrix(l)
        n = rp.size - 1
        A = sp.csr_matrix((v, ci, rp), shape=(n, n))
        nc = int(agg.max()) + 1
        assert agg.min() == 0 and np.unique(agg).size == nc  # every aggregate is used
        prp, pci, pv = s.level_prolongator(l)
        P = sp.csr_matrix((pv, pci, prp), shape=(n, nc))
        T = sp.csr_matrix((np.ones(n), (np.arange(n), agg)), shape=(n, nc))
        Pdef = (T - omega * sp.diags(1.0 / A.diagonal()) @ (A @ T)).tocsr()
        dP = abs(P - Pdef).max()
        assert dP <= 1e-14 * abs(Pdef).max(), (l, dP)
        crp, cci, cv, _ = s.level_matrix(l + 1)
        Ac = sp.csr_matrix((cv, cci, crp), shape=(nc, nc))
        Adef = (P.T @ A @ P).tocsr()
        dA = abs(Ac - Adef).max()
        report(f"      level {l}: n={n} nc={nc} |P - def|={dP:.1e} |A_c - P^T A P|={dA:.1e}")
        assert dA <= 1e-12 * abs(Adef).max(), (l, dA)
        assert abs(Ac - Ac.T).max() <= 1e-13 * abs(Ac).max()
        # every product entry is present in the stored pattern (no entry dropped)
        missing = abs(Adef) - abs(Adef).multiply(Ac != 0)
        assert missing.max() <= 1e-12 * abs(Adef).max()
    s.close()


def test_hierarchy_wide_rows_take_global_tables(fem, gpu):
    """Sparse products with rows too long for the per-thread LDS hash tables:
    a 2D 800 x 100 grid (80,000 rows, so the level-0 products take the
    thread-per-row LDS kernels) plus one hub vertex joined to all 800 bottom
    vertices by a fan of triangles.  The hub's row of A P has far more than
    192 distinct columns (the larger table's limit), so its symbolic product
    runs in the global-memory table (sym_global_kernel); the level operators
    must still equal P^T A P."""
    report("[gpu] hierarchy with a hub row (global-memory product tables)")
    nx, ny = 800, 100
    X, Y = np.meshgrid(np.linspace(0.0, 1.0, nx), np.linspace(0.0, 1.0, ny))
    xyz = np.stack([X.ravel(), Y.ravel(), np.zeros(nx * ny)], axis=1)
    idx = np.arange(nx * ny).reshape(ny, nx)
    a, b, c, d = idx[:-1, :-1].ravel(), idx[:-1, 1:].ravel(), idx[1:, 1:].ravel(), idx[1:, :-1].ravel()
    tris = np.concatenate([np.stack([a, b, c], 1), np.stack([a, c, d], 1)])
    hub = nx * ny
    xyz = np.vstack([xyz, [0.5, -0.3, 0.0]])
    bottom = idx[0]
    fan = np.stack([np.full(nx - 1, hub), bottom[1:], bottom[:-1]], 1)
    tris = np.ascontiguousarray(np.concatenate([tris, fan]).astype(np.int32))
    s = fem.FEMSolver.from_arrays(2, xyz, tris, mass_coeff=1e-2)
    omega = 0.67
    s.set_params(solverType=1, precision=0, topSize=256, aggregationRule=1, strengthThreshold=0.0, proOmega=omega)
    s.setup()
    rp, ci, v, agg = s.level_matrix(0)
    n = rp.size - 1
    A = sp.csr_matrix((v, ci, rp), shape=(n, n))
    nc = int(agg.max()) + 1
    prp, pci, pv = s.level_prolongator(0)
    P = sp.csr_matrix((pv, pci, prp), shape=(n, nc))
    widest = int(np.diff((A @ P).tocsr().indptr).max())
    report(f"      n={n} nc={nc} widest row of A P: {widest} columns")
    assert n > 65536 and widest > 192
    for l in range(s.num_levels() - 1):
        rp, ci, v, agg = s.level_matrix(l)
        n = rp.size - 1
        A = sp.csr_matrix((v, ci, rp), shape=(n, n))
        nc = int(agg.max()) + 1
        prp, pci, pv = s.level_prolongator(l)
        P = sp.csr_matrix((pv, pci, prp), shape=(n, nc))
        T = sp.csr_matrix((np.ones(n), (np.arange(n), agg)), shape=(n, nc))
        Pdef = (T - omega * sp.diags(1.0 / A.diagonal()) @ (A @ T)).tocsr()
        assert abs(P - Pdef).max() <= 1e-14 * abs(Pdef).max()
        crp, cci, cv, _ = s.level_matrix(l + 1)
        Ac = sp.csr_matrix((cv, cci, crp), shape=(nc, nc))
        Adef = (P.T @ A @ P).tocsr()
        dA = abs(Ac - Adef).max()
        report(f"      level {l}: n={n} nc={nc} |A_c - P^T A P|={dA:.1e}")
        assert dA <= 1e-12 * abs(Adef).max()
        missing = abs(Adef) - abs(Adef).multiply(Ac != 0)
        assert missing.max() <= 1e-12 * abs(Adef).max()
    s.close()
