"""GPU AMG-PCG parity: solutions against an FP64 direct solve (scipy SuperLU on
the oracle's matrix), the true-residual stop test, mixed precision, batched
right-hand sides, preconditioner symmetry and the reference's own
end-to-end fixtures (src/test/sanity2D.cc, sanity3D.cc, tetVol.cc)."""
from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from conftest import REFDATA, report

pytestmark = pytest.mark.gpu


def _problem(meshgen, oracle, N, sigma, mass, seed=0):
    xyz, conn = meshgen.cube_mesh(N, seed=seed)
    kappa = meshgen.lognormal_kappa(seed, conn.shape[0], sigma)
    A = oracle.scipy_matrix(xyz, conn, kappa, mass)
    b = meshgen.gaussian_rhs(seed, xyz.shape[0])
    return xyz, conn, kappa, A, b


@pytest.mark.parametrize("precision", [0, 1, 2])
@pytest.mark.parametrize("smoother", [0, 1])
def test_pcg_matches_direct_solve(fem, gpu, oracle, meshgen, precision, smoother):
    report(f"[gpu] pcg N=13 precision={precision} smoother={smoother}")
    xyz, conn, kappa, A, b = _problem(meshgen, oracle, 13, 1.0, 1e-2, seed=1)
    s = fem.FEMSolver.from_arrays(3, xyz, conn, kappa=kappa, mass_coeff=1e-2)
    # tolerance 1e-12: |x - x_direct| / |x_direct| <= cond(K) * 1e-12, well
    # inside the 1e-8 bar for the solution error against the direct solve
    s.set_params(solverType=1, tolerance=1e-12, maxIters=500, precision=precision, smootherType=smoother,
                 preInnerIters=2, postInnerIters=2, topSize=64)
    x, info = s.solve(b)
    xd = spla.spsolve(A.tocsc(), b)
    # the residual on the library's own assembled matrix (the oracle's differs
    # by ~1e-16 |A| per entry, which alone leaves ~5e-12 of residual here)
    Ag = s.get_matrix_scipy()
    rel_res = np.linalg.norm(b - Ag @ x) / np.linalg.norm(b)
    err = np.linalg.norm(x - xd) / np.linalg.norm(xd)
    report(f"      iters={info['iterations']} levels={info['num_levels']} res={rel_res:.2e} err={err:.2e}")
    assert info["status"] == 0 and info["converged"] == 1
    assert info["rel_residual"] <= 1e-12
    # device and host evaluate b - A x with different rounding
    assert abs(info["rel_residual"] - rel_res) <= 0.2 * rel_res + 1e-13
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) <= 1e-10
    assert err <= 1e-8
    assert info["num_levels"] >= 2
    s.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_zero_rhs_gives_zero_solution(fem, gpu, oracle, meshgen, precision):
    """A x = 0 has the solution x = 0 whatever the initial guess, alone and as
    one column of a batch (the other columns solve normally)."""
    report(f"[gpu] zero right-hand side, precision={precision}")
    xyz, conn, kappa, A, b = _problem(meshgen, oracle, 9, 1.0, 1e-2, seed=4)
    s = fem.FEMSolver.from_arrays(3, xyz, conn, kappa=kappa, mass_coeff=1e-2)
    s.set_params(solverType=1, tolerance=1e-10, maxIters=300, precision=precision, smootherType=1,
                 preInnerIters=2, postInnerIters=2, topSize=64)
    n = xyz.shape[0]
    x0 = meshgen.noise(3, n, 1.0)
    x, info = s.solve(np.zeros(n), x0=x0)
    assert info["status"] == 0 and np.all(x == 0.0)
    B = np.zeros((n, 3))
    B[:, 0] = b
    B[:, 2] = 2.0 * b
    X0 = np.stack([meshgen.noise(5 + c, n, 1.0) for c in range(3)], axis=1)
    X, info = s.solve(B, x0=X0)
    assert info["status"] == 0
    assert np.all(X[:, 1] == 0.0)
    xd = spla.spsolve(A.tocsc(), b)
    for c, f in ((0, 1.0), (2, 2.0)):
        assert np.linalg.norm(X[:, c] - f * xd) / np.linalg.norm(f * xd) <= 1e-7
    s.close()


def test_batched_rhs_equal_single(fem, gpu, oracle, meshgen):
    report("[gpu] batched k=4 vs single")
    xyz, conn, kappa, A, _ = _problem(meshgen, oracle, 11, 1.5, 1e-2, seed=2)
    s = fem.FEMSolver.from_arrays(3, xyz, conn, kappa=kappa, mass_coeff=1e-2)
    s.set_params(solverType=1, tolerance=1e-9, maxIters=300, precision=1, preInnerIters=2, postInnerIters=2,
                 topSize=64)
    B = meshgen.gaussian_rhs(5, xyz.shape[0], 4)
    X, info = s.solve(B)
    assert info["status"] == 0
    for c in range(4):
        x1, _ = s.solve(np.ascontiguousarray(B[:, c]))
        r = np.linalg.norm(B[:, c] - A @ X[:, c]) / np.linalg.norm(B[:, c])
        assert r <= 1e-9
        assert np.linalg.norm(X[:, c] - x1) / np.linalg.norm(x1) <= 1e-7
    s.close()


@pytest.mark.parametrize("smoother", [0, 1])
@pytest.mark.parametrize("precision", [0, 1, 2])
def test_preconditioner_symmetric_positive(fem, gpu, meshgen, smoother, precision):
    """<M u, v> = <u, M v> and <M u, u> > 0 for the V-cycle as PCG applies it:
    exactly in FP64, up to the rounding of the FP32 (and FP16) level-0
    operator in the mixed modes."""
    report(f"[gpu] V-cycle symmetry smoother={smoother} precision={precision}")
    xyz, conn = meshgen.cube_mesh(10, seed=3)
    kappa = meshgen.lognormal_kappa(3, conn.shape[0], 1.0)
    s = fem.FEMSolver.from_arrays(3, xyz, conn, kappa=kappa, mass_coeff=1e-2)
    s.set_params(precision=precision, smootherType=smoother, preInnerIters=2, postInnerIters=2, topSize=64)
    rng = np.random.default_rng(7)
    u = rng.normal(size=xyz.shape[0])
    v = rng.normal(size=xyz.shape[0])
    Mu, Mv = s.apply_preconditioner(u), s.apply_preconditioner(v)
    tol = 1e-10 if precision == 0 else 1e-5
    assert abs(Mu @ v - u @ Mv) <= tol * np.linalg.norm(Mu) * np.linalg.norm(v)
    assert Mu @ u > 0 and Mv @ v > 0
    s.close()


def test_one_vcycle_mode(fem, gpu, oracle, meshgen):
    """solverType 0 applies exactly one V-cycle (src/core/cuda/amg.cu:193)."""
    report("[gpu] solverType 0")
    xyz, conn, kappa, A, b = _problem(meshgen, oracle, 9, 0.5, 1.0, seed=4)
    s = fem.FEMSolver.from_arrays(3, xyz, conn, kappa=kappa, mass_coeff=1.0)
    s.set_params(solverType=0, tolerance=1e-6, topSize=32)
    x, info = s.solve(b)
    assert info["iterations"] == 1
    r = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
    assert r < 0.5 and abs(info["rel_residual"] - r) < 1e-9
    assert np.allclose(x, s.apply_preconditioner(b), rtol=1e-12, atol=1e-14)
    s.close()


def test_square_c1_like(fem, gpu, oracle, meshgen):
    report("[gpu] 2D sine load N=33")
    xyz, tris = meshgen.square_mesh(33, seed=0)
    b = meshgen.sine_load_2d(xyz, tris)
    A = oracle.scipy_matrix(xyz, tris, None, 1.0)
    s = fem.FEMSolver.from_arrays(2, xyz, tris, mass_coeff=1.0)
    s.set_params(solverType=1, tolerance=1e-10, maxIters=200, topSize=64)
    x, info = s.solve(b)
    xd = spla.spsolve(A.tocsc(), b)
    assert info["status"] == 0
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) <= 1e-8
    s.close()


def test_not_converged_status(fem, gpu, meshgen):
    report("[gpu] maxIters exhausted")
    xyz, conn = meshgen.cube_mesh(9, seed=5)
    s = fem.FEMSolver.from_arrays(3, xyz, conn, mass_coeff=1e-2)
    s.set_params(solverType=1, tolerance=1e-14, maxIters=2, topSize=64)
    b = meshgen.gaussian_rhs(1, xyz.shape[0])
    x, info = s.solve(b)
    assert info["status"] == fem._lib.FEM_ERR_NOT_CONVERGED and info["converged"] == 0
    assert info["iterations"] == 2 and np.all(np.isfinite(x))
    with pytest.raises(fem.NotConverged):
        s.solve(b, raise_on_nonconvergence=True)
    s.close()


# ---- the reference's end-to-end tests (same tolerances), plus tight versions ----

def test_reference_sanity3d(fem, gpu):
    """src/test/sanity3D.cc: simple mesh + simple.mat (79.9568 I), ||x - ans|| < 1."""
    report("[gpu] sanity3D")
    cfg = fem.FEMSolver(str(REFDATA / "simple"), True, False)
    assert cfg.readMatlabSparseMatrix(str(REFDATA / "simple.mat")) == 0
    b = fem.read_matlab_array(str(REFDATA / "simpleb.mat"))
    x = np.zeros(cfg.getMatrixRows())
    cfg.solveFEM(x, b)
    ans = fem.read_matlab_array(str(REFDATA / "simpleAns.mat"))
    assert np.linalg.norm(x - ans) < 1.0
    # tight: PCG to 1e-10 reproduces the fixture answer (itself A^-1 b to 2.3e-8)
    cfg.solverType_ = 1
    cfg.tolerance_ = 1e-10
    x2 = np.zeros_like(x)
    cfg.solveFEM(x2, b)
    assert np.linalg.norm(x2 - ans) / np.linalg.norm(ans) < 1e-6
    cfg.close()


def test_reference_tetvol(fem, gpu):
    """src/test/tetVol.cc: SCIRun matrix on the tetVol mesh, ||x - ans|| < 25."""
    report("[gpu] tetVol")
    cfg = fem.FEMSolver(str(REFDATA / "tetVol"), True, False)
    assert cfg.readMatlabSparseMatrix(str(REFDATA / "tetVolA.mat")) == 0
    b = fem.read_matlab_array(str(REFDATA / "tetVolb.mat"))
    x = np.zeros(cfg.getMatrixRows())
    cfg.solveFEM(x, b)
    ans = fem.read_matlab_array(str(REFDATA / "tetVolAns.mat"))
    assert np.linalg.norm(x - ans) < 25.0
    cfg.solverType_ = 1
    cfg.tolerance_ = 1e-10
    cfg.maxIters_ = 500
    x2 = np.zeros_like(x)
    info = cfg.solveFEM(x2, b)
    report(f"      tetVol PCG iters={info['iterations']} res={info['rel_residual']:.2e}")
    assert info["status"] == 0
    # the fixture answer itself has relative residual 6.7e-6
    assert np.linalg.norm(x2 - ans) / np.linalg.norm(ans) < 1e-4
    cfg.close()


def test_reference_sanity2d(fem, gpu):
    """src/test/sanity2D.cc: simple.ply + simpleTri.mat, ||x - ans|| < 100."""
    report("[gpu] sanity2D")
    cfg = fem.FEMSolver(str(REFDATA / "simple.ply"), False, False)
    assert cfg.readMatlabSparseMatrix(str(REFDATA / "simpleTri.mat")) == 0
    b = fem.read_matlab_array(str(REFDATA / "simpleTrib.mat"))
    x = np.zeros(cfg.getMatrixRows())
    cfg.solveFEM(x, b)
    ans = fem.read_matlab_array(str(REFDATA / "simpleTriAns.mat"))
    assert np.linalg.norm(x - ans) < 100.0
    cfg.close()


def test_graph_replay_matches_eager(fem, gpu, oracle, meshgen):
    """launchMode 1 replays one captured PCG iteration; it must give bitwise the
    results of launching every kernel (launchMode 0): across repeated solves,
    a tolerance change (re-capture), k = 2 columns and a re-setup."""
    report("[gpu] captured PCG iteration vs eager launches")
    xyz, conn, kappa, A, b = _problem(meshgen, oracle, 17, 1.5, 1e-2, seed=4)
    B = meshgen.gaussian_rhs(9, xyz.shape[0], 2)
    runs = {}
    for mode in (0, 1):
        s = fem.FEMSolver.from_arrays(3, xyz, conn, kappa=kappa, mass_coeff=1e-2)
        s.set_params(solverType=1, tolerance=1e-9, maxIters=300, precision=1, smootherType=1, preInnerIters=2,
                     postInnerIters=2, topSize=64, launchMode=mode)
        out = []
        for _ in range(3):
            x, info = s.solve(b)
            out.append((x, info["iterations"]))
        s.set_params(tolerance=1e-11)
        x, info = s.solve(b)
        out.append((x, info["iterations"]))
        X, info = s.solve(B)
        out.append((X, info["iterations"]))
        s.set_params(precision=0)  # new hierarchy
        x, info = s.solve(b)
        out.append((x, info["iterations"]))
        runs[mode] = out
        s.close()
    for (x0, i0), (x1, i1) in zip(runs[0], runs[1]):
        assert i0 == i1
        assert np.array_equal(x0, x1)
    # tolerance 1e-11 on the device residual; the host re-evaluation of b - A x
    # carries ~|A||x| eps of rounding of its own (x ~ 1e4 here)
    assert np.linalg.norm(b - A @ runs[1][3][0]) / np.linalg.norm(b) <= 1e-10


def test_queued_ahead_iteration_leaves_solution_unchanged(fem, gpu, oracle, meshgen):
    """The next PCG iteration may be queued before the host has read the
    previous one's convergence flags; an iteration queued after convergence
    must not change x.  A solve capped at exactly the reported iteration count
    (which cannot queue beyond the cap) gives bitwise the same x."""
    report("[gpu] iteration queued past convergence")
    xyz, conn, kappa, A, b = _problem(meshgen, oracle, 21, 2.0, 1e-2, seed=6)
    s = fem.FEMSolver.from_arrays(3, xyz, conn, kappa=kappa, mass_coeff=1e-2)
    base = dict(solverType=1, tolerance=1e-9, precision=1, smootherType=1, preInnerIters=2, postInnerIters=2,
                topSize=64)
    s.set_params(maxIters=500, **base)
    x1, i1 = s.solve(b)
    assert i1["status"] == 0
    s.set_params(maxIters=i1["iterations"], **base)
    x2, i2 = s.solve(b)
    assert i2["status"] == 0 and i2["iterations"] == i1["iterations"]
    assert np.array_equal(x1, x2)
    s.close()


@pytest.mark.parametrize("k,smoother", [(3, 1), (8, 1), (12, 1), (16, 1), (32, 1), (64, 1), (8, 0)])
@pytest.mark.parametrize("precision", [0, 1])
def test_batched_widths(fem, gpu, oracle, meshgen, k, smoother, precision):
    """k columns take the one-column-per-lane kernels (k = 3), the 4-column
    group kernels on the sliced-ELL copies (fine operator, prolongators; FP32
    and FP64, operator + dot included) or the CSR column-group kernels with
    split entries (long coarse rows, restrictions); in mixed precision the
    last level-0 smoothing step also writes the r.z partials (k a multiple of
    4: 1, 2, 4, 8 or 16 lanes per row, k = 12 with idle lanes; Chebyshev and
    Jacobi) and the PCG update takes the first one.  Every column must match
    its single-column solve."""
    report(f"[gpu] batched k={k} smoother={smoother} precision={precision}")
    xyz, conn, kappa, A, _ = _problem(meshgen, oracle, 11, 1.5, 1e-2, seed=8)
    s = fem.FEMSolver.from_arrays(3, xyz, conn, kappa=kappa, mass_coeff=1e-2)
    s.set_params(solverType=1, tolerance=1e-9, maxIters=300, precision=precision, preInnerIters=2,
                 postInnerIters=2, topSize=64, smootherType=smoother)
    B = meshgen.gaussian_rhs(11, xyz.shape[0], k)
    X, info = s.solve(B)
    assert info["status"] == 0
    Y = s.apply_operator(X)
    # rounding of A X is relative to |A| |X| (X ~ 1e4 with A X ~ 1 here)
    assert np.max(np.abs(Y - A @ X)) <= 1e-14 * np.max(abs(A) @ np.abs(X))
    for c in range(k):
        x1, _ = s.solve(np.ascontiguousarray(B[:, c]))
        assert np.linalg.norm(B[:, c] - A @ X[:, c]) / np.linalg.norm(B[:, c]) <= 1e-9
        assert np.linalg.norm(X[:, c] - x1) / np.linalg.norm(x1) <= 1e-7
    s.close()


def test_handles_reuse_the_device_reserve(fem, gpu, oracle, meshgen):
    """Three handles in turn on a 41^3 mesh (10.8 MB of host arrays: the
    threaded upload): the second draws its temporaries from the blocks the
    first's mesh build and setup released, the third runs after
    release_cached_memory() freed them.  Stale contents of reused blocks must
    not matter: the three solutions are bitwise equal and solve the oracle's
    system."""
    report("[gpu] device reserve: three handles, release in between")
    xyz, conn, kappa, A, b = _problem(meshgen, oracle, 41, 1.0, 1e-2, seed=6)
    assert xyz.nbytes + conn.nbytes + kappa.nbytes >= 8 << 20
    xs = []
    for k in range(3):
        if k == 2:
            fem.release_cached_memory()
        s = fem.FEMSolver.from_arrays(3, xyz, conn, kappa=kappa, mass_coeff=1e-2)
        s.set_params(solverType=1, tolerance=1e-10, maxIters=300, precision=1, smootherType=1,
                     preInnerIters=2, postInnerIters=2, topSize=64)
        x, info = s.solve(b)
        assert info["converged"] == 1
        xs.append(x)
        s.close()
    assert np.array_equal(xs[0], xs[1]) and np.array_equal(xs[0], xs[2])
    assert np.linalg.norm(b - A @ xs[0]) / np.linalg.norm(b) <= 1e-9
