"""The V-cycle's kernels against a host statement of the same V-cycle
(tests/vcycle_host.py, FP64 numpy on the exported hierarchy): one smoothing
step per level (Jacobi and Chebyshev, FP64 and FP32 kernels, per-row
rounding bounds), the coarsest-level solve (applied in FP64 in every
precision), a whole preconditioner application (the fused FP64-in / FP64-out
level-0 steps of the mixed path included), and the PCG's fused first
smoothing step (computed by the PCG update) through two PCG iterations
driven on the host."""
from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from conftest import report
from vcycle_host import HostVCycle, cheb_coefficients

pytestmark = pytest.mark.gpu

# bench.py's V-cycle settings at a size the host statement runs in seconds
PARAMS = dict(solverType=1, tolerance=1e-8, maxIters=500, smootherType=1, preInnerIters=3, postInnerIters=3,
              postRelaxes=1, topSize=256, aggregationRule=1, strengthThreshold=0.01)


def _solver(fem, meshgen, N, sigma, precision, smoother, seed=0, **extra):
    xyz, tets = meshgen.cube_mesh(N, seed=seed)
    kappa = meshgen.lognormal_kappa(seed, tets.shape[0], sigma)
    s = fem.FEMSolver.from_arrays(3, xyz, tets, kappa=kappa, mass_coeff=1e-2)
    s.set_params(**dict(PARAMS, precision=precision, smootherType=smoother, **extra))
    s.setup()
    return s


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("mode", [0, 1])
def test_smoother_step_every_level(fem, gpu, meshgen, precision, mode):
    """One Jacobi / Chebyshev step on every smoothed level against
    y = x + c1 d + c2 D^-1 (b - A x) (Chebyshev) or x + c1 D^-1 (b - A x).
    Bound per row i: (len_i + 8) u (|x_i| + |c1 d_i| + |c2 dinv_i| (|b_i| + (|A||x|)_i))
    with u the unit roundoff of the kernel's precision (inputs rounded to it)."""
    report(f"[gpu] smoother step per level, precision={precision} mode={mode}")
    s = _solver(fem, meshgen, 25, 2.0, precision, 1, topSize=64)
    u = 2.0 ** -53 if precision == 0 else 2.0 ** -24
    rng = np.random.default_rng(31 + mode)
    nl = s.num_levels()
    assert nl >= 3
    for l in range(nl - 1):
        rp, ci, v, _ = s.level_matrix(l)
        n = rp.size - 1
        A = sp.csr_matrix((v, ci, rp), shape=(n, n))
        dinv = 1.0 / A.diagonal()
        x, b, d = rng.normal(size=n), rng.normal(size=n), rng.normal(size=n)
        c1, c2 = (0.37, 0.81) if mode == 1 else (0.66, 0.0)
        y, d_new = s.smoother_step(l, mode, c1, c2, x, b, d)
        res = b - A @ x
        absAx = abs(A) @ abs(x)
        ln = np.diff(rp)
        if mode == 1:
            d_ref = c1 * d + c2 * dinv * res
            y_ref = x + d_ref
            bound = (ln + 8) * u * (abs(x) + abs(c1 * d) + abs(c2 * dinv) * (abs(b) + absAx))
            err_d = float(np.max(np.abs(d_new - d_ref) / (bound + 1e-300)))
        else:
            y_ref = x + c1 * dinv * res
            bound = (ln + 8) * u * (abs(x) + abs(c1 * dinv) * (abs(b) + absAx))
            err_d = 0.0
            assert np.array_equal(d_new, d.astype(np.float32).astype(np.float64) if precision else d)
        err_y = float(np.max(np.abs(y - y_ref) / (bound + 1e-300)))
        report(f"      level {l}: n={n} error/bound y {err_y:.2e} d {err_d:.2e}")
        assert err_y <= 1.0 and err_d <= 1.0, (l, err_y, err_d)
    s.close()


def test_smoother_step_level0_full_size(fem, gpu, meshgen):
    """The level-0 Chebyshev step (the LDS-staged FP32 kernel the bench times)
    at the 97^3 cube against the host statement, with the per-row bound."""
    report("[gpu] level-0 Chebyshev step at 97^3 vs host")
    c = meshgen.CONFIGS["c3_cube97_mixed"]
    s = _solver(fem, meshgen, c["N"], c["sigma"], 1, 1)
    rp, ci, v, _ = s.level_matrix(0)
    n = rp.size - 1
    A = sp.csr_matrix((v, ci, rp), shape=(n, n))
    dinv = 1.0 / A.diagonal()
    S = s.level_smoother(0)
    coef = cheb_coefficients(S["lmin"], S["lmax"], 3)
    rng = np.random.default_rng(97)
    x, b, d = rng.normal(size=n), rng.normal(size=n), rng.normal(size=n)
    u = 2.0 ** -24
    for c1, c2 in coef[1:]:
        y, d_new = s.smoother_step(0, 1, c1, c2, x, b, d)
        d_ref = c1 * d + c2 * dinv * (b - A @ x)
        bound = (np.diff(rp) + 8) * u * (abs(x) + abs(c1 * d) + abs(c2 * dinv) * (abs(b) + abs(A) @ abs(x)))
        err = float(np.max(np.abs(d_new - d_ref) / bound))
        erry = float(np.max(np.abs(y - (x + d_ref)) / bound))
        report(f"      c1={c1:.4f} c2={c2:.4e}: error/bound d {err:.2e} y {erry:.2e}")
        assert err <= 1.0 and erry <= 1.0
    s.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_coarse_solve_in_fp64(fem, gpu, meshgen, precision):
    """The coarsest level's direct solve: FP64 inverse and FP64 sums in every
    precision.  Mixed: y = fl32(A_c^-1 fl32(x)) to one FP32 unit in the last
    place (an FP32 inverse or FP32 sums miss this by cond(A_c) ulps);
    FP64: y = A_c^-1 x to 1e-12 relative."""
    report(f"[gpu] coarsest-level solve, precision={precision}")
    s = _solver(fem, meshgen, 21, 2.0, precision, 1, topSize=256)
    lc = s.num_levels() - 1
    rp, ci, v, _ = s.level_matrix(lc)
    nc = rp.size - 1
    Ac = sp.csr_matrix((v, ci, rp), shape=(nc, nc)).toarray()
    cond = np.linalg.cond(Ac)
    rng = np.random.default_rng(7)
    x = rng.normal(size=nc)
    y = s.apply_level(lc, 3, x)
    if precision == 0:
        ref = np.linalg.solve(Ac, x)
        err = np.linalg.norm(y - ref) / np.linalg.norm(ref)
        report(f"      nc={nc} cond={cond:.1e} rel error {err:.1e}")
        assert err <= 1e-12 * max(1.0, cond / 1e3)
    else:
        xf = x.astype(np.float32).astype(np.float64)
        ref = np.linalg.solve(Ac, xf)
        # one FP32 ulp of each entry (floored at the ulp of 1e-6 max|ref|, where
        # the FP64 inverse's own error, ~cond * 1e-16 max|ref|, can matter)
        mag = np.maximum(np.abs(ref), 1e-6 * np.abs(ref).max())
        ulp = np.spacing(mag.astype(np.float32)).astype(np.float64)
        worst = float(np.max(np.abs(y - ref) / ulp))
        report(f"      nc={nc} cond={cond:.1e} max error {worst:.2f} ulp(fp32)")
        assert worst <= 1.0
        assert np.array_equal(y, y.astype(np.float32).astype(np.float64))
    s.close()


@pytest.mark.parametrize("precision,smoother,N", [(0, 1, 21), (0, 0, 21), (1, 1, 21), (1, 0, 21), (2, 1, 21),
                                                (1, 1, 97)])
def test_preconditioner_matches_host_vcycle(fem, gpu, meshgen, precision, smoother, N):
    """z = M r through fem_apply_preconditioner (mixed one-column path: the
    first level-0 step reads the FP64 r, the last writes z in FP64) against
    the FP64 host V-cycle on the same hierarchy, for a random r and for a
    smooth r = A v.  Relative 2-norm bounds:
    - FP64 hierarchy: 1e-10 (FP64 rounding, amplified by the coarse solve);
    - FP32 hierarchy: 20 e_in + 1e-6, where e_in = |M r - M fl32(r)| / |M r|
      is what rounding the input alone to FP32 changes in the host V-cycle
      (the kernels round ~10 intermediate vectors the same way; for smooth r
      the coarse correction amplifies such rounding to ~1e-4..1e-3);
    - FP16 level-0 smoother matrix (precision 2): 2e-2 (off-diagonals rounded
      to 11 bits).
    A wrong smoothing coefficient or a dropped term changes z by O(0.1)."""
    report(f"[gpu] preconditioner vs host V-cycle, precision={precision} smoother={smoother} N={N}")
    sigma = 2.0 if N > 21 else 1.5
    s = _solver(fem, meshgen, N, sigma, precision, smoother, topSize=256 if N > 21 else 64)
    host = HostVCycle(s)
    perm = s.permutation()
    rng = np.random.default_rng(N + smoother)
    n = perm.size
    for trial in range(2):
        r_user = rng.normal(size=n) if trial == 0 else host.A[0] @ rng.normal(size=n)
        if trial == 1:
            r_user = r_user[np.argsort(perm)]  # internal -> user order
        z_user = s.apply_preconditioner(r_user)
        r_int = r_user[perm]
        z_int = host.vcycle(0, r_int)
        z_ref = np.empty(n)
        z_ref[perm] = z_int
        err = np.linalg.norm(z_user - z_ref) / np.linalg.norm(z_ref)
        e_in = np.linalg.norm(host.vcycle(0, r_int.astype(np.float32).astype(np.float64)) - z_int) / \
            np.linalg.norm(z_int)
        tol = {0: 1e-10, 1: 20 * e_in + 1e-6, 2: 2e-2}[precision]
        report(f"      trial {trial}: {s.num_levels()} levels, rel error {err:.2e} (input-rounding effect "
               f"{e_in:.1e}, bound {tol:.1e})")
        assert err <= tol
    s.close()


def test_pcg_first_step_fused_into_update(fem, gpu, meshgen):
    """From the second PCG iteration on, the mixed path takes the level-0
    smoother's first step inside the PCG update (x0 = c D^-1 fl32(r), d0 = x0).
    Two GPU PCG iterations against two host PCG iterations whose
    preconditioner is fem_apply_preconditioner (first step by its own kernel):
    equal to FP64 rounding of the host/device dot products."""
    report("[gpu] PCG iteration 2 (first smoothing step in the update) vs host-driven PCG")
    xyz, tets = meshgen.cube_mesh(21, seed=4)
    kappa = meshgen.lognormal_kappa(4, tets.shape[0], 1.5)
    s = fem.FEMSolver.from_arrays(3, xyz, tets, kappa=kappa, mass_coeff=1e-2)
    for smoother in (1, 0):
        s.set_params(**dict(PARAMS, precision=1, smootherType=smoother, topSize=64, maxIters=2, tolerance=1e-30))
        A = s.get_matrix_scipy()
        b = meshgen.gaussian_rhs(4, xyz.shape[0])
        x_gpu, info = s.solve(b)
        assert info["iterations"] == 2
        # host PCG, two iterations from x0 = 0
        r = b.copy()
        z = s.apply_preconditioner(r)
        p = z.copy()
        rz = r @ z
        x = np.zeros_like(b)
        for it in range(2):
            q = A @ p
            alpha = rz / (p @ q)
            x = x + alpha * p
            r = r - alpha * q
            if it == 1:
                break
            z = s.apply_preconditioner(r)
            rz_new = r @ z
            p = z + (rz_new / rz) * p
            rz = rz_new
        err = np.linalg.norm(x_gpu - x) / np.linalg.norm(x)
        report(f"      smoother={smoother}: |x_gpu - x_host| / |x_host| = {err:.2e}")
        assert err <= 1e-9
    s.close()


@pytest.mark.parametrize("ctype", [1, 2])
@pytest.mark.parametrize("precision", [0, 1])
def test_w_and_f_cycles(fem, gpu, meshgen, ctype, precision):
    """cycleType 1 (W) and 2 (F): the preconditioner against the host
    statement of the same cycle (vcycle_host.HostVCycle.cycle).
    FP64 hierarchy: to 1e-10, and the V-cycle differs from it by more than
    100x that (a cycle type that were ignored would give the V-cycle's z).
    FP32 hierarchy: 20 (e_in + e_h) + 1e-6, with e_in what rounding the input
    alone and e_h what rounding the stored operators, prolongators and D^-1
    alone to FP32 change in the host cycle: the second coarse cycle works on
    the residual of a nearly solved coarse problem, where the FP32 operator's
    rounding (relative to |A||x|, not to |A x|) is of the order of the W-V
    difference itself.  Then PCG with the cycle (F: flexible beta) reaches the
    direct solution, in no more iterations than with the V-cycle."""
    report(f"[gpu] cycleType={ctype} precision={precision}")
    s = _solver(fem, meshgen, 25, 2.0, precision, 1, topSize=32, cycleType=ctype)
    host = HostVCycle(s)
    assert s.num_levels() >= 3  # level 1 is then cycled twice per coarse correction of level 0
    perm = s.permutation()
    n = perm.size
    rng = np.random.default_rng(7 + ctype)
    r_user = rng.normal(size=n)
    z_user = s.apply_preconditioner(r_user)
    r_int = r_user[perm]
    z_int = host.cycle(0, r_int, ctype)
    z_ref = np.empty(n)
    z_ref[perm] = z_int
    nz = np.linalg.norm(z_int)
    err = np.linalg.norm(z_user - z_ref) / nz
    dv = np.linalg.norm(host.vcycle(0, r_int) - z_int) / nz
    if precision == 0:
        tol = 1e-10
        report(f"      rel error vs host cycle {err:.2e} (bound {tol:.1e}); V-cycle differs by {dv:.2e}")
        assert err <= tol and dv > 100 * tol
    else:
        e_in = np.linalg.norm(host.cycle(0, r_int.astype(np.float32).astype(np.float64), ctype) - z_int) / nz
        e_h = np.linalg.norm(host.rounded_to_fp32().cycle(0, r_int, ctype) - z_int) / nz
        tol = 20 * (e_in + e_h) + 1e-6
        report(f"      rel error vs host cycle {err:.2e} (e_in {e_in:.1e}, e_h {e_h:.1e}, bound {tol:.1e}); "
               f"V-cycle differs by {dv:.2e}")
        assert err <= tol
    # PCG with this cycle against the direct solve
    A = s.get_matrix_scipy().tocsc()
    b = rng.normal(size=n)
    s.set_params(tolerance=1e-10)
    x, info = s.solve(b)
    xd = spla.spsolve(A, b)
    e = np.linalg.norm(x - xd) / np.linalg.norm(xd)
    s.set_params(cycleType=0)
    _, info_v = s.solve(b)
    report(f"      PCG: {info['iterations']} iterations (V-cycle: {info_v['iterations']}), rel error {e:.1e}")
    assert info["status"] == 0 and e <= 1e-8
    assert info["iterations"] <= info_v["iterations"]
    s.close()
