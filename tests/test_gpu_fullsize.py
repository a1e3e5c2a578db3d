"""GPU runs at the benchmark configurations' full sizes, checked through
properties that do not need a direct solve: the true FP64 residual evaluated
on the host from the exported matrix, agreement of the mixed and FP64 paths,
linearity, the closed form 1^T K 1 = mass_coeff * vol(domain), symmetry, and
the sliced-ELL operator (both its 16-bit-column and int32-column slices)
against a host CSR product, and every level matrix through the V-cycle's
kernels against host products."""
from __future__ import annotations

import sys

import numpy as np
import pytest
import scipy.sparse as sp

from conftest import ROOT, report

pytestmark = pytest.mark.gpu

sys.path.insert(0, str(ROOT))
import bench  # noqa: E402  (the benchmark's V-cycle settings are the tested ones)

# bench.py's V-cycle settings (precision set per test)
PARAMS = {k: v for k, v in bench.solver_params(1).items() if k != "precision"}
assert PARAMS["strengthThreshold"] == 0.01


def _config(meshgen, name):
    c = meshgen.CONFIGS[name]
    xyz, tets = meshgen.cube_mesh(c["N"], seed=0)
    kappa = meshgen.lognormal_kappa(0, tets.shape[0], c["sigma"])
    return xyz, tets, kappa, c["mass"]


def _true_residual(A, x, b):
    return np.linalg.norm(b - A @ x) / np.linalg.norm(b)


def _check_mass_closed_form(A, mass):
    """Stiffness rows sum to zero and the P1 mass matrix sums to vol = 1, so
    1^T K 1 = mass_coeff up to the rounding of the assembled entries."""
    one = np.ones(A.shape[0])
    scale = np.abs(A.data).sum()
    assert abs(one @ (A @ one) - mass) <= 1e-13 * scale + 1e-12 * mass


def _wide_slices(rp, ci, bits=13, windows=8):
    """Number of 64-row slices whose columns touch more than `windows`
    aligned blocks of 2^bits columns (stored with int32 columns)."""
    n = rp.size - 1
    wide = 0
    for s0 in range(0, n, 64):
        blk = np.unique(ci[rp[s0]:rp[min(n, s0 + 64)]] >> bits)
        wide += blk.size > windows
    return wide, (n + 63) // 64


@pytest.mark.parametrize("name", ["c2_cube49_fp64", "c3_cube97_mixed"])
def test_config_solves(fem, gpu, meshgen, name):
    report(f"[gpu] full-size {name}")
    xyz, tets, kappa, mass = _config(meshgen, name)
    s = fem.FEMSolver.from_arrays(3, xyz, tets, kappa=kappa, mass_coeff=mass)
    A = s.get_matrix_scipy()
    assert (A != A.T).nnz == 0
    _check_mass_closed_form(A, mass)
    b = meshgen.gaussian_rhs(0, xyz.shape[0])
    xs = {}
    for prec in (0, 1):
        s.set_params(precision=prec, **PARAMS)
        x, info = s.solve(b)
        res = _true_residual(A, x, b)
        report(f"      precision={prec}: {info['iterations']} iterations, host residual {res:.2e}, "
               f"{info['num_levels']} levels")
        assert info["status"] == 0 and info["converged"] == 1
        assert res <= 1e-8 * (1 + 1e-6)
        xs[prec] = x
    # both paths stop at 1e-8 on the same operator: the error bound is cond(A)
    # times the residual; the observed difference is far below that
    diff = np.linalg.norm(xs[1] - xs[0]) / np.linalg.norm(xs[0])
    report(f"      |x_mixed - x_fp64| / |x_fp64| = {diff:.2e}")
    assert diff <= 1e-6
    s.close()


def test_c3_operator_both_slice_kinds(fem, gpu, meshgen):
    report("[gpu] 97^3 sliced-ELL operator vs host CSR")
    xyz, tets, kappa, mass = _config(meshgen, "c3_cube97_mixed")
    s = fem.FEMSolver.from_arrays(3, xyz, tets, kappa=kappa, mass_coeff=mass)
    s.set_params(precision=1, **PARAMS)
    s.setup()
    rp, ci, _, _ = s.level_matrix(0)
    wide, ns = _wide_slices(rp, ci)
    report(f"      {ns - wide} compact / {wide} int32-column slices")
    assert 0 < wide < ns
    A = s.get_matrix_scipy()
    rng = np.random.default_rng(97)
    X = rng.normal(size=(A.shape[0], 2))
    for c in range(2):
        y = s.apply_operator(np.ascontiguousarray(X[:, c]))
        ref = A @ X[:, c]
        err = np.max(np.abs(y - ref)) / np.max(np.abs(ref))
        assert err <= 1e-13, err
    s.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_c3_level_kernels_match_host(fem, gpu, meshgen, precision):
    """Every level's operator, prolongator and restriction applied through the
    V-cycle's one-column kernels (sliced ELL with 1 to 8 lanes per row, or
    CSR) against the host product of the exported FP64 matrices.  Bound per
    row: (row length + 3) unit roundoffs of sum |m| |x| (values and x rounded
    to the working precision, then a sum of row length products)."""
    report(f"[gpu] 97^3 level kernels vs host products, precision={precision}")
    xyz, tets, kappa, mass = _config(meshgen, "c3_cube97_mixed")
    s = fem.FEMSolver.from_arrays(3, xyz, tets, kappa=kappa, mass_coeff=mass)
    s.set_params(precision=precision, **PARAMS)
    s.setup()
    u = 2.0 ** -53 if precision == 0 else 2.0 ** -24
    rng = np.random.default_rng(5)
    nl = s.num_levels()
    long_rows = 0
    for l in range(nl):
        rp, ci, v, _ = s.level_matrix(l)
        n = rp.size - 1
        mats = [(0, sp.csr_matrix((v, ci, rp), shape=(n, n)))]
        if l < nl - 1:
            nc = s.level_matrix(l + 1)[0].size - 1
            prp, pci, pv = s.level_prolongator(l)
            P = sp.csr_matrix((pv, pci, prp), shape=(n, nc))
            mats += [(1, P), (2, P.T.tocsr())]
        for which, M in mats:
            x = rng.normal(size=M.shape[1])
            y = s.apply_level(l, which, x)
            ref = M @ x
            bound = (np.diff(M.indptr) + 3) * u * (abs(M) @ abs(x)) + 1e-300
            worst = float(np.max(np.abs(y - ref) / bound))
            avg = M.nnz / max(M.shape[0], 1)
            long_rows += M.shape[0] >= 32768 and avg > 24
            report(f"      level {l} {'APR'[which]}: {M.shape[0]} x {M.shape[1]}, {avg:.1f} per row, "
                   f"error / bound {worst:.2e}")
            assert worst <= 1.0, (l, which, worst)
    # the multi-lane sliced kernels ran (long rows on a large level)
    assert long_rows >= 1
    s.close()


def test_c5_full_size_properties(fem, gpu, meshgen):
    report("[gpu] 161^3 (headline) mixed solve, linearity")
    xyz, tets, kappa, mass = _config(meshgen, "c5_cube161_mixed16")
    s = fem.FEMSolver.from_arrays(3, xyz, tets, kappa=kappa, mass_coeff=mass)
    s.set_params(precision=1, **PARAMS)
    A = s.get_matrix_scipy()
    _check_mass_closed_form(A, mass)
    B = meshgen.gaussian_rhs(0, xyz.shape[0], 2)
    b1, b2 = np.ascontiguousarray(B[:, 0]), np.ascontiguousarray(B[:, 1])
    x1, i1 = s.solve(b1)
    x2, _ = s.solve(b2)
    x3, _ = s.solve(b1 + 2.0 * b2)
    for x, b in ((x1, b1), (x2, b2), (x3, b1 + 2.0 * b2)):
        res = _true_residual(A, x, b)
        assert res <= 1e-8 * (1 + 1e-6), res
    lin = np.linalg.norm(x3 - (x1 + 2.0 * x2)) / np.linalg.norm(x3)
    report(f"      {i1['iterations']} iterations, {i1['num_levels']} levels, linearity error {lin:.2e}")
    assert lin <= 1e-6
    s.close()


def test_c1_square65(fem, gpu, oracle, meshgen):
    """C1: 65^2 square, P1 triangles, sine load, FP64 path against a direct
    solve of the oracle's matrix, and the mixed path against the FP64 one."""
    report("[gpu] full-size c1_square65_fp64")
    c = meshgen.CONFIGS["c1_square65_fp64"]
    xyz, tris = meshgen.square_mesh(c["N"], seed=0)
    b = meshgen.sine_load_2d(xyz, tris)
    A = oracle.scipy_matrix(xyz, tris, None, c["mass"])
    s = fem.FEMSolver.from_arrays(2, xyz, tris, mass_coeff=c["mass"])
    xs = {}
    for prec in (0, 1):
        s.set_params(precision=prec, **dict(PARAMS, tolerance=1e-10, topSize=64))
        x, info = s.solve(b)
        assert info["status"] == 0
        xs[prec] = x
    import scipy.sparse.linalg as spla
    xd = spla.spsolve(A.tocsc(), b)
    assert np.linalg.norm(xs[0] - xd) / np.linalg.norm(xd) <= 1e-8
    assert np.linalg.norm(xs[1] - xs[0]) / np.linalg.norm(xs[0]) <= 1e-8
    s.close()


def _secant_fd(s, sd, kappa, d, h, b, u_obs):
    """Central difference quotient (L(kappa + h d) - L(kappa - h d)) / (2 h) of
    L = 0.5 |u - u_obs|^2, evaluated through the exact secant identity instead
    of subtracting two nearly equal losses: with K(kappa) = sum kappa_e S_e + M,
    K_+ - K_- = 2 h S(d), so u_+ - u_- = -2 h w with w = K_+^-1 S(d) u_- and
        (L_+ - L_-) / (2 h) = -w.(u_- - u_obs) + h w.w
    exactly.  The solves' errors then enter relative to w, not to the 1e-6-sized
    difference of two losses.  sd: a handle on the same mesh assembled with
    kappa = d and no mass (S(d) u by its operator)."""
    s.update_kappa(kappa - h * d)
    um, im = s.solve(b)
    assert im["status"] == 0
    sd.update_kappa(d)
    rhs = sd.apply_operator(um)
    s.update_kappa(kappa + h * d)
    w, iw = s.solve(rhs)
    assert iw["status"] == 0
    return float(-np.dot(w, um - u_obs) + h * np.dot(w, w))


def test_c4_adjoint_full_size(fem, gpu, meshgen):
    """C4: 49^3, sigma 1, L = 0.5 |u - u_obs|^2 with u_obs from a second seeded
    coefficient plus N(0, 1e-3) noise.  The GPU adjoint gradient against
    (1) the FP64 CPU reference gradient (SuperLU solves of the oracle's
    matrices and the oracle's element gradient, tests/golden/c4_gradient.npz
    from tests/golden/make_c4_gradient.py) at 4096 seeded elements and in
    norm; (2) central differences along 16 seeded directions at h = 1e-6
    (exact secant identity, GPU FP64 solves at 1e-12); (3) plain central
    differences of the loss along 4 directions at h = 1e-4."""
    report("[gpu] full-size c4_cube49_adjoint")
    xyz, tets, kappa, mass = _config(meshgen, "c4_cube49_adjoint")
    kappa2 = meshgen.lognormal_kappa(101, tets.shape[0], 1.0)
    b = meshgen.gaussian_rhs(0, xyz.shape[0])
    tight = dict(PARAMS, precision=0, tolerance=1e-12, maxIters=1000)
    s2 = fem.FEMSolver.from_arrays(3, xyz, tets, kappa=kappa2, mass_coeff=mass)
    s2.set_params(**tight)
    u2, _ = s2.solve(b)
    s2.close()
    u_obs = u2 + meshgen.noise(7, xyz.shape[0], 1e-3)
    s = fem.FEMSolver.from_arrays(3, xyz, tets, kappa=kappa, mass_coeff=mass)
    s.set_params(**tight)
    u, info = s.solve(b)
    g, _, ainfo = s.adjoint_grad(u, u - u_obs)
    assert info["status"] == 0 and ainfo["status"] == 0
    # (1) FP64 CPU reference gradient
    fx = np.load(ROOT / "tests" / "golden" / "c4_gradient.npz")
    idx, g_ref = fx["idx"], fx["g"]
    err_s = np.linalg.norm(g[idx] - g_ref) / np.linalg.norm(g_ref)
    err_n = abs(np.linalg.norm(g) - float(fx["g_norm"])) / float(fx["g_norm"])
    loss = 0.5 * float(np.dot(u - u_obs, u - u_obs))
    report(f"      vs FP64 CPU gradient: sampled rel error {err_s:.2e}, norm rel error {err_n:.2e}, "
           f"loss rel error {abs(loss - float(fx['loss'])) / float(fx['loss']):.2e}")
    assert err_s <= 1e-7 and err_n <= 1e-8
    # (2) 16 directions, h = 1e-6
    sd = fem.FEMSolver.from_arrays(3, xyz, tets, kappa=np.ones_like(kappa), mass_coeff=0.0)
    rng = np.random.default_rng(1649)
    errs = []
    for _ in range(16):
        d = rng.normal(size=kappa.size) * kappa
        fd = _secant_fd(s, sd, kappa, d, 1e-6, b, u_obs)
        ad = float(np.dot(g, d))
        errs.append(abs(fd - ad) / (np.linalg.norm(g) * np.linalg.norm(d)))
    report(f"      16 directions h=1e-6: max scaled error {max(errs):.1e}")
    assert max(errs) < 1e-7
    sd.close()

    # (3) plain central differences of the loss, h = 1e-4
    def loss_at(k):
        s.update_kappa(k)
        uk, ik = s.solve(b)
        assert ik["status"] == 0
        return 0.5 * float(np.dot(uk - u_obs, uk - u_obs))

    rng = np.random.default_rng(49)
    h = 1e-4
    for _ in range(4):
        d = rng.normal(size=kappa.size) * kappa
        fd = (loss_at(kappa + h * d) - loss_at(kappa - h * d)) / (2 * h)
        ad = float(np.dot(g, d))
        err = abs(fd - ad) / (np.linalg.norm(g) * np.linalg.norm(d))
        report(f"      FD {fd:.6e} adjoint {ad:.6e} scaled error {err:.1e}")
        assert err < 1e-6
    s.close()


def test_c5_batched_and_adjoint_full_size(fem, gpu, meshgen):
    """C5's defining legs at 161^3: the 16 right-hand sides through one batched
    V-cycle (true FP64 residual <= 1e-8 per column on the host, and two
    columns against their single-right-hand-side solves), and the adjoint
    gradient for one right-hand side against central differences along 2
    seeded directions (h = 1e-6, exact secant identity) on the mixed path."""
    report("[gpu] full-size c5_cube161_mixed16: 16-RHS solve and adjoint")
    xyz, tets, kappa, mass = _config(meshgen, "c5_cube161_mixed16")
    n = xyz.shape[0]
    s = fem.FEMSolver.from_arrays(3, xyz, tets, kappa=kappa, mass_coeff=mass)
    s.set_params(precision=1, **PARAMS)
    A = s.get_matrix_scipy()
    B = meshgen.gaussian_rhs(0, n, 16)
    X, info = s.solve(B)
    assert info["status"] == 0
    R = B - A @ X
    res = np.linalg.norm(R, axis=0) / np.linalg.norm(B, axis=0)
    report(f"      16 RHS: {info['iterations']} iterations, max column residual {res.max():.2e}")
    assert np.all(res <= 1e-8 * (1 + 1e-6))
    for c in (0, 11):
        x1, i1 = s.solve(np.ascontiguousarray(B[:, c]))
        assert i1["status"] == 0
        diff = np.linalg.norm(X[:, c] - x1) / np.linalg.norm(x1)
        report(f"      column {c}: |x_batched - x_single| / |x_single| = {diff:.2e}")
        assert diff <= 1e-6
    del R, X
    # adjoint for one right-hand side (tolerance 1e-10 so the FD quotients'
    # solve errors stay far below the gradient)
    b = np.ascontiguousarray(B[:, 0])
    del B
    s.set_params(precision=1, **dict(PARAMS, tolerance=1e-10, maxIters=1000))
    u, info = s.solve(b)
    # dL/du = -noise: a rough right-hand side (a smooth one is dominated by the
    # lowest, mass-scaled modes, whose FP64 residual floor on this 7e9-condition
    # operator sits above 1e-10)
    u_obs = u + meshgen.noise(5, n, 1e-3)
    g, _, ainfo = s.adjoint_grad(u, u - u_obs)
    report(f"      forward {info['iterations']} it (res {info['rel_residual']:.1e}), adjoint "
           f"{ainfo['iterations']} it (res {ainfo['rel_residual']:.1e})")
    assert info["status"] == 0 and ainfo["status"] == 0
    sd = fem.FEMSolver.from_arrays(3, xyz, tets, kappa=np.ones_like(kappa), mass_coeff=0.0)
    rng = np.random.default_rng(161)
    for _ in range(2):
        d = rng.normal(size=kappa.size) * kappa
        fd = _secant_fd(s, sd, kappa, d, 1e-6, b, u_obs)
        ad = float(np.dot(g, d))
        err = abs(fd - ad) / (np.linalg.norm(g) * np.linalg.norm(d))
        report(f"      FD {fd:.6e} adjoint {ad:.6e} scaled error {err:.1e}")
        assert err < 1e-6
    sd.close()
    s.close()
