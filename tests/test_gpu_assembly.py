"""GPU parity of the device pattern + assembly kernels against the oracle.

Bar: pattern bit-exact; values within 1e-13 of max|A| (FP64 sums in a
different order); A bitwise symmetric and bitwise reproducible run to run."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import REFDATA, report
from matfile import read_mat_v5

pytestmark = pytest.mark.gpu


def _gpu_matrix(fem, dim, xyz, conn, kappa=None, mass=1.0, flags=0):
    s = fem.FEMSolver.from_arrays(dim, xyz, conn, kappa=kappa, mass_coeff=mass, flags=flags)
    rp, ci, v = s.get_matrix_csr()
    return s, rp, ci, v


@pytest.mark.parametrize("dim,N,sigma,mass", [(3, 6, 1.0, 1e-2), (3, 9, 2.0, 1.0), (2, 9, None, 1.0),
                                              (2, 17, 1.5, 0.3)])
def test_pattern_and_values_match_oracle(fem, gpu, oracle, meshgen, dim, N, sigma, mass):
    report(f"[gpu] assembly dim={dim} N={N} sigma={sigma}")
    xyz, conn = meshgen.cube_mesh(N, seed=11) if dim == 3 else meshgen.square_mesh(N, seed=11)
    kappa = None if sigma is None else meshgen.lognormal_kappa(11, conn.shape[0], sigma)
    s, rp, ci, v = _gpu_matrix(fem, dim, xyz, conn, kappa, mass)
    orp, oci, ov = oracle.assemble(xyz, conn, kappa, mass)
    assert np.array_equal(rp, orp)
    assert np.array_equal(ci, oci)
    scale = np.abs(ov).max()
    assert np.abs(v - ov).max() <= 1e-13 * scale
    s.close()


def test_symmetric_and_deterministic(fem, gpu, meshgen):
    report("[gpu] assembly symmetry/determinism N=8")
    xyz, conn = meshgen.cube_mesh(8, seed=4)
    kappa = meshgen.lognormal_kappa(4, conn.shape[0], 1.0)
    s1, rp, ci, v1 = _gpu_matrix(fem, 3, xyz, conn, kappa, 1e-2)
    s2, _, _, v2 = _gpu_matrix(fem, 3, xyz, conn, kappa, 1e-2)
    assert np.array_equal(v1.view(np.uint64), v2.view(np.uint64))
    import scipy.sparse as sp
    A = sp.csr_matrix((v1, ci, rp), shape=(xyz.shape[0],) * 2)
    D = (A - A.T).tocoo()
    assert np.count_nonzero(D.data) == 0  # bitwise symmetric
    s1.close()
    s2.close()


def test_tetvol_fixture_pattern(fem, gpu):
    report("[gpu] tetVol fixture pattern")
    s = fem.FEMSolver(str(REFDATA / "tetVol"), True, False)
    rp, ci, v = s.get_matrix_csr()
    n = s.getMatrixRows()
    (nr, nc), r, c, _ = read_mat_v5(REFDATA / "tetVolA.mat")
    keys = np.sort(np.repeat(np.arange(n, dtype=np.int64), np.diff(rp)) * n + ci)
    assert np.array_equal(keys, np.unique(r.astype(np.int64) * n + c))
    assert rp[-1] == 129844
    s.close()


def test_file_mesh_assembly_matches_oracle(fem, gpu, oracle):
    """Reference-parity path: float-parsed TetGen mesh, label coefficients, mass 1."""
    report("[gpu] CubeMesh file assembly")
    base = str(REFDATA / "CubeMesh_size256step16_correct")
    s = fem.FEMSolver(base, True, False)
    xyz, tets, lab = fem.read_tetgen(base)
    rp, ci, v = s.get_matrix_csr()
    orp, oci, ov = oracle.assemble(xyz, tets, None, 1.0)
    assert np.array_equal(ci, oci)
    assert np.abs(v - ov).max() <= 1e-13 * np.abs(ov).max()
    s.close()


def test_apply_operator_matches_scipy(fem, gpu, oracle, meshgen):
    report("[gpu] apply_operator k=1,3")
    xyz, conn = meshgen.cube_mesh(7, seed=2)
    kappa = meshgen.lognormal_kappa(2, conn.shape[0], 1.0)
    s = fem.FEMSolver.from_arrays(3, xyz, conn, kappa=kappa, mass_coeff=1e-2)
    A = oracle.scipy_matrix(xyz, conn, kappa, 1e-2)
    rng = np.random.default_rng(0)
    x = rng.normal(size=xyz.shape[0])
    assert np.allclose(s.apply_operator(x), A @ x, rtol=1e-12, atol=1e-12 * np.abs(A @ x).max())
    X = rng.normal(size=(xyz.shape[0], 3))
    assert np.allclose(s.apply_operator(X), A @ X, rtol=1e-12, atol=1e-12 * np.abs(A @ X).max())
    s.close()


def test_element_grad_matches_oracle(fem, gpu, oracle, meshgen):
    report("[gpu] element gradient")
    for dim, N in ((3, 6), (2, 10)):
        xyz, conn = meshgen.cube_mesh(N, seed=8) if dim == 3 else meshgen.square_mesh(N, seed=8)
        s = fem.FEMSolver.from_arrays(dim, xyz, conn)
        rng = np.random.default_rng(3)
        a = rng.normal(size=xyz.shape[0])
        u = rng.normal(size=xyz.shape[0])
        g = s.element_grad(a, u)
        og = oracle.element_grad(xyz, conn, a, u)
        assert np.abs(g - og).max() <= 1e-12 * np.abs(og).max()
        s.close()


def test_invalid_mesh_inputs_rejected(fem, gpu, meshgen):
    """Out-of-range vertex indices and NaN inputs are rejected with a status
    (checked on the device before anything is built), never a crash."""
    report("[gpu] invalid mesh inputs")
    xyz, tets = meshgen.cube_mesh(5, seed=1)
    kappa = meshgen.lognormal_kappa(1, tets.shape[0], 1.0)
    bad_conn = tets.copy()
    bad_conn[7, 2] = xyz.shape[0]
    with pytest.raises(fem.FemError) as e:
        fem.FEMSolver.from_arrays(3, xyz, bad_conn, kappa=kappa, mass_coeff=1e-2)
    assert e.value.status == 1 and "out of range" in str(e.value)
    bad_k = kappa.copy()
    bad_k[3] = np.nan
    with pytest.raises(fem.FemError) as e:
        fem.FEMSolver.from_arrays(3, xyz, tets, kappa=bad_k, mass_coeff=1e-2)
    assert e.value.status == 6
    bad_x = xyz.copy()
    bad_x[11, 1] = np.nan
    with pytest.raises(fem.FemError) as e:
        fem.FEMSolver.from_arrays(3, bad_x, tets, kappa=kappa, mass_coeff=1e-2)
    assert e.value.status == 6
    s = fem.FEMSolver.from_arrays(3, xyz, tets, kappa=kappa, mass_coeff=1e-2)
    with pytest.raises(fem.FemError) as e:
        s.update_kappa(bad_k)
    assert e.value.status == 6
    s.close()
