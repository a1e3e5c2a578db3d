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
