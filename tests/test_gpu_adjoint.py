"""Adjoint gradient dL/dkappa (no reference counterpart): against central finite
differences along seeded random directions and against an FP64 CPU gradient
(SuperLU solves on the oracle's matrix + the oracle's element gradient).

Loss L(kappa) = 0.5 ||u(kappa) - u_obs||^2, K(kappa) u = b,
dL/dkappa_e = -|T_e| grad(lambda) . grad(u) with K lambda = u - u_obs."""
from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from conftest import report

pytestmark = pytest.mark.gpu


def _setup(meshgen, N=9, seed=3):
    xyz, tets = meshgen.cube_mesh(N, seed=seed)
    kappa = meshgen.lognormal_kappa(seed, tets.shape[0], 1.0)
    kappa2 = meshgen.lognormal_kappa(seed + 101, tets.shape[0], 1.0)
    b = meshgen.gaussian_rhs(seed, xyz.shape[0])
    return xyz, tets, kappa, kappa2, b


def _tight(s, precision=0):
    # 1e-12: the FP64 true residual of these systems floors near 1e-14..1e-13
    s.set_params(solverType=1, tolerance=1e-12, maxIters=1000, precision=precision, smootherType=1,
                 preInnerIters=2, postInnerIters=2, topSize=64)


def test_adjoint_matches_fp64_cpu_gradient(fem, gpu, oracle, meshgen):
    report("[gpu] adjoint vs FP64 CPU gradient N=9")
    xyz, tets, kappa, kappa2, b = _setup(meshgen)
    A2 = oracle.scipy_matrix(xyz, tets, kappa2, 1e-2)
    u_obs = spla.spsolve(A2.tocsc(), b) + meshgen.noise(7, xyz.shape[0], 1e-3)
    s = fem.FEMSolver.from_arrays(3, xyz, tets, kappa=kappa, mass_coeff=1e-2)
    _tight(s)
    u, info = s.solve(b)
    assert info["status"] == 0
    g, lam, ainfo = s.adjoint_grad(u, u - u_obs)
    assert ainfo["status"] == 0
    # FP64 CPU reference
    A = oracle.scipy_matrix(xyz, tets, kappa, 1e-2)
    lu = spla.splu(A.tocsc())
    u_ref = lu.solve(b)
    lam_ref = lu.solve(u_ref - u_obs)
    g_ref = oracle.element_grad(xyz, tets, lam_ref, u_ref)
    assert np.linalg.norm(u - u_ref) / np.linalg.norm(u_ref) < 1e-8
    # error bound ~ cond(K) x tolerance for u and lambda
    assert np.linalg.norm(lam - lam_ref) / np.linalg.norm(lam_ref) < 1e-7
    assert np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref) < 1e-7
    s.close()


def test_adjoint_matches_finite_differences(fem, gpu, oracle, meshgen):
    """GPU adjoint gradient against central differences of the loss computed
    with FP64 direct solves (h = 1e-6, 16 seeded directions)."""
    report("[gpu] adjoint vs central differences (16 directions, h=1e-6)")
    xyz, tets, kappa, kappa2, b = _setup(meshgen, N=7, seed=5)
    A2 = oracle.scipy_matrix(xyz, tets, kappa2, 1e-2)
    u_obs = spla.spsolve(A2.tocsc(), b) + meshgen.noise(9, xyz.shape[0], 1e-3)

    def loss(k):
        u = spla.splu(oracle.scipy_matrix(xyz, tets, k, 1e-2).tocsc()).solve(b)
        return 0.5 * float(np.dot(u - u_obs, u - u_obs))

    s = fem.FEMSolver.from_arrays(3, xyz, tets, kappa=kappa, mass_coeff=1e-2)
    _tight(s)
    u, _ = s.solve(b)
    g, _, info = s.adjoint_grad(u, u - u_obs)
    assert info["status"] == 0
    rng = np.random.default_rng(16)
    dirs = [rng.normal(size=kappa.size) * kappa for _ in range(16)]  # relative: kappa stays > 0
    # h = 1e-6: FD rounding (~1e-12 L / h) bounds the agreement near 1e-4;
    # h = 1e-4: truncation ~h^2, rounding 100x smaller -> a much tighter check
    # error scaled by |g| |d| (directional derivatives can cancel to ~0)
    for h, tol in ((1e-6, 1e-4), (1e-4, 1e-6)):
        errs = []
        for d in dirs:
            fd = (loss(kappa + h * d) - loss(kappa - h * d)) / (2 * h)
            ad = float(np.dot(g, d))
            errs.append(abs(fd - ad) / (np.linalg.norm(g) * np.linalg.norm(d)))
        report(f"      h={h:g}: max rel FD error {max(errs):.2e}")
        assert max(errs) < tol
    s.close()


def test_torch_autograd_function(fem, gpu, oracle, meshgen):
    import torch
    report("[gpu] torch.autograd.Function backward")
    xyz, tets, kappa, kappa2, b = _setup(meshgen, N=7, seed=11)
    s = fem.FEMSolver.from_arrays(3, xyz, tets, kappa=kappa, mass_coeff=1e-2)
    _tight(s, precision=0)
    s.set_params(tolerance=1e-11)
    dev = torch.device("cuda:0")
    k_t = torch.tensor(kappa, dtype=torch.float64, device=dev, requires_grad=True)
    b_t = torch.tensor(b, dtype=torch.float64, device=dev, requires_grad=True)
    u_obs = torch.tensor(meshgen.noise(3, xyz.shape[0], 1.0), dtype=torch.float64, device=dev)
    Fn = fem.FEMSolveFunction
    u = Fn.apply(k_t, b_t, s)
    L = 0.5 * ((u - u_obs) ** 2).sum()
    L.backward()
    # FP64 CPU reference of both gradients
    A = oracle.scipy_matrix(xyz, tets, kappa, 1e-2)
    lu = spla.splu(A.tocsc())
    u_ref = lu.solve(b)
    lam = lu.solve(u_ref - u_obs.cpu().numpy())
    g_ref = oracle.element_grad(xyz, tets, lam, u_ref)
    gk = k_t.grad.cpu().numpy()
    gb = b_t.grad.cpu().numpy()
    assert np.linalg.norm(u.detach().cpu().numpy() - u_ref) / np.linalg.norm(u_ref) < 1e-8
    assert np.linalg.norm(gk - g_ref) / np.linalg.norm(g_ref) < 1e-7
    assert np.linalg.norm(gb - lam) / np.linalg.norm(lam) < 1e-7
    s.close()


def test_torch_autograd_two_applies_different_kappa(fem, gpu, oracle, meshgen):
    """Two forward applies on one solver with different kappa and one backward
    through both: each node's adjoint must use its own K (the solver holds the
    second kappa when backward starts)."""
    import torch
    report("[gpu] torch.autograd: two applies, different kappa, one backward")
    xyz, tets, kappa, kappa2, b = _setup(meshgen, N=7, seed=13)
    s = fem.FEMSolver.from_arrays(3, xyz, tets, kappa=kappa, mass_coeff=1e-2)
    _tight(s, precision=0)
    s.set_params(tolerance=1e-11)
    dev = torch.device("cuda:0")
    k1 = torch.tensor(kappa, dtype=torch.float64, device=dev, requires_grad=True)
    k2 = torch.tensor(kappa2, dtype=torch.float64, device=dev, requires_grad=True)
    b_t = torch.tensor(b, dtype=torch.float64, device=dev)
    w1 = torch.tensor(meshgen.noise(21, xyz.shape[0], 1.0), dtype=torch.float64, device=dev)
    w2 = torch.tensor(meshgen.noise(22, xyz.shape[0], 1.0), dtype=torch.float64, device=dev)
    Fn = fem.FEMSolveFunction
    u1 = Fn.apply(k1, b_t, s)
    u2 = Fn.apply(k2, b_t, s)
    L = (w1 * u1).sum() + 0.5 * ((u2 - w2) ** 2).sum()
    L.backward()
    for kap, kt, dLdu_fn in ((kappa, k1, lambda u: w1.cpu().numpy()),
                             (kappa2, k2, lambda u: u - w2.cpu().numpy())):
        lu = spla.splu(oracle.scipy_matrix(xyz, tets, kap, 1e-2).tocsc())
        u_ref = lu.solve(b)
        lam = lu.solve(dLdu_fn(u_ref))
        g_ref = oracle.element_grad(xyz, tets, lam, u_ref)
        g = kt.grad.cpu().numpy()
        assert np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref) < 1e-7
    s.close()


def test_torch_inputs_from_async_kernels_on_default_stream(fem, gpu, oracle, meshgen):
    """b and the initial guess produced by asynchronous torch kernels on the
    default stream immediately before the call: the library must see their
    final values (it runs on a stream of its own)."""
    import torch
    report("[gpu] torch default-stream producers before solve / apply_operator")
    xyz, tets, kappa, _, b = _setup(meshgen, N=9, seed=17)
    s = fem.FEMSolver.from_arrays(3, xyz, tets, kappa=kappa, mass_coeff=1e-2)
    _tight(s, precision=0)
    A = oracle.scipy_matrix(xyz, tets, kappa, 1e-2)
    dev = torch.device("cuda:0")
    assert torch.cuda.current_stream().cuda_stream == 0
    n = xyz.shape[0]
    big = torch.randn(4096, 4096, dtype=torch.float64, device=dev)
    for trial in range(3):
        torch.cuda.synchronize()
        # a long queue of work ahead of the producer, then b = 0 + b_host * (trial + 1)
        for _ in range(4):
            big = big @ big
            big = big / big.norm()
        b_t = torch.zeros(n, dtype=torch.float64, device=dev) + big[0, 0] * 0
        b_t += torch.tensor(b, dtype=torch.float64, device=dev) * (trial + 1)
        x, info = s.solve(b_t)
        xd = spla.spsolve(A.tocsc(), b * (trial + 1))
        assert info["status"] == 0
        assert np.linalg.norm(x.cpu().numpy() - xd) / np.linalg.norm(xd) < 1e-8
        for _ in range(4):
            big = big @ big
            big = big / big.norm()
        xin = torch.zeros(n, dtype=torch.float64, device=dev) + big[0, 0] * 0
        xin += torch.tensor(b, dtype=torch.float64, device=dev)
        y = s.apply_operator(xin)
        yd = A @ b
        assert np.abs(y.cpu().numpy() - yd).max() <= 1e-13 * np.abs(yd).max()
    s.close()
