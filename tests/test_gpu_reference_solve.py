"""GPU solutions against the restated reference solve (oracle/fem_oracle_amg.c:
the reference's AMG-PCG at its own defaults -- randomized MIS-2 aggregation,
omega 0.67 smoothed prolongator, 5/5 partition-local Jacobi sweeps of weight 1,
FP32 hierarchy, dense coarsest solve -- on the host) and an FP64 direct
solve.  Both iterative solves stop on a relative true residual tau, so
|x_gpu - x_ref| <= (|r_gpu| + |r_ref|) / lambda_min(K); the test asserts that
bound (lambda_min by shift-invert Lanczos) and reports the observed gap."""
from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from conftest import report

pytestmark = pytest.mark.gpu

TAU = 1e-10


def _lambda_min(A):
    return float(spla.eigsh(A.tocsc(), k=1, sigma=0.0, which="LM", return_eigenvectors=False)[0])


def _compare(fem, oracle, dim, xyz, conn, kappa, mass, b, name):
    A = oracle.scipy_matrix(xyz, conn, kappa, mass)
    rp, ci, v = oracle.assemble(xyz, conn, kappa, mass)
    ref = oracle.amg_pcg(None, None, None, None, b, tol=TAU, max_iters=500, A=(rp, ci, v))
    assert ref["status"] == 0 and ref["converged"]
    x_ref = ref["x"]
    lmin = _lambda_min(A)
    xd = spla.spsolve(A.tocsc(), b)
    s = fem.FEMSolver.from_arrays(dim, xyz, conn, kappa=kappa, mass_coeff=mass)
    for prec in (0, 1):
        s.set_params(solverType=1, tolerance=TAU, maxIters=500, precision=prec)
        x, info = s.solve(b)
        assert info["status"] == 0
        r_gpu = np.linalg.norm(b - A @ x)
        r_ref = np.linalg.norm(b - A @ x_ref)
        gap = np.linalg.norm(x - x_ref)
        bound = (r_gpu + r_ref) / lmin
        rel = gap / np.linalg.norm(x_ref)
        report(f"      {name} precision={prec}: gpu {info['iterations']} it, reference restatement "
               f"{ref['iterations']} it; |x_gpu - x_ref|/|x_ref| = {rel:.2e} (bound {bound / np.linalg.norm(x_ref):.2e}); "
               f"|x_gpu - x_direct|/|x_direct| = {np.linalg.norm(x - xd) / np.linalg.norm(xd):.2e}")
        assert gap <= bound * 1.001
        assert np.linalg.norm(x - xd) / np.linalg.norm(xd) <= 1e-8
    s.close()


def test_c1_square65_vs_reference_restatement(fem, gpu, oracle, meshgen):
    report("[gpu] C1 (65^2 square, sine load) vs the restated reference AMG-PCG")
    c = meshgen.CONFIGS["c1_square65_fp64"]
    xyz, tris = meshgen.square_mesh(c["N"], seed=0)
    b = meshgen.sine_load_2d(xyz, tris)
    _compare(fem, oracle, 2, xyz, tris, None, c["mass"], b, "C1")


def test_cube21_vs_reference_restatement(fem, gpu, oracle, meshgen):
    report("[gpu] 21^3 cube (sigma 1) vs the restated reference AMG-PCG")
    xyz, tets = meshgen.cube_mesh(21, seed=6)
    kappa = meshgen.lognormal_kappa(6, tets.shape[0], 1.0)
    b = meshgen.gaussian_rhs(6, xyz.shape[0])
    _compare(fem, oracle, 3, xyz, tets, kappa, 1e-2, b, "21^3")
