"""fem_apply_preconditioner_noalloc (the entry point that allocates no device
memory; its work lives in one static device array per GPU) against
fem_apply_preconditioner on the same handle: bitwise the same z, on two
operators in a row on the shared array (the second one smaller), and twice on
one handle.  Also its refusals: an operator above the row limit, a handle
without a hierarchy."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from conftest import report

pytestmark = pytest.mark.gpu

PARAMS = dict(solverType=1, tolerance=1e-8, maxIters=500, smootherType=1, preInnerIters=3, postInnerIters=3,
              postRelaxes=1, topSize=256, aggregationRule=1, strengthThreshold=0.01)


def _solver(fem, meshgen, N, precision, seed, setup=True):
    xyz, tets = meshgen.cube_mesh(N, seed=seed)
    kappa = meshgen.lognormal_kappa(seed, tets.shape[0], 2.0)
    s = fem.FEMSolver.from_arrays(3, xyz, tets, kappa=kappa, mass_coeff=1e-2)
    s.set_params(**dict(PARAMS, precision=precision))
    if setup:
        s.setup()
    return s


def _noalloc(fem, s, r):
    """z = M r through the allocation-free entry point (device pointers)."""
    import torch
    rd = torch.from_numpy(np.ascontiguousarray(r, dtype=np.float64)).cuda()
    zd = torch.full_like(rd, float("nan"))
    torch.cuda.synchronize()  # r is on the device before the library's stream reads it
    st = fem.lib().fem_apply_preconditioner_noalloc(s._h, C.c_void_p(rd.data_ptr()), C.c_void_p(zd.data_ptr()))
    assert st == 0, fem.lib().fem_last_error(s._h).decode()
    torch.cuda.synchronize()  # the entry point returns without waiting for its stream
    return zd.cpu().numpy()


@pytest.mark.parametrize("precision", [1, 0, 2])
def test_noalloc_preconditioner_matches_allocating_entry(fem, gpu, meshgen, precision):
    assert int(fem.lib().fem_noalloc_max_rows()) == 1 << 21
    for N, seed in ((41, 3), (21, 4)):  # the second operator is smaller: leftovers of the first must not matter
        report(f"[gpu] noalloc preconditioner precision={precision} N={N}")
        s = _solver(fem, meshgen, N, precision, seed)
        n = s.getMatrixRows()
        rng = np.random.default_rng(100 + N + precision)
        for rep in range(2):
            r = rng.normal(size=n)
            z_ref = s.apply_preconditioner(r)
            z = _noalloc(fem, s, r)
            assert np.all(np.isfinite(z))
            assert np.array_equal(z, z_ref), (N, rep, float(np.abs(z - z_ref).max()))
        s.close()


def test_noalloc_preconditioner_refusals(fem, gpu, meshgen):
    import torch
    L = fem.lib()
    # no hierarchy yet: the entry point builds nothing
    report("[gpu] noalloc preconditioner refusals: no setup")
    s = _solver(fem, meshgen, 9, 1, 0, setup=False)
    n = s.getMatrixRows()
    rd = torch.zeros(n, dtype=torch.float64, device="cuda")
    zd = torch.zeros_like(rd)
    st = L.fem_apply_preconditioner_noalloc(s._h, C.c_void_p(rd.data_ptr()), C.c_void_p(zd.data_ptr()))
    assert st == 7 and b"fem_setup" in L.fem_last_error(s._h)  # FEM_ERR_STATE
    s.close()
    # above the row limit (130^3 = 2,197,000 > 2^21 rows): refused before any work
    report("[gpu] noalloc preconditioner refusals: 130^3 rows")
    s = _solver(fem, meshgen, 130, 1, 0, setup=False)
    n = s.getMatrixRows()
    assert n > (1 << 21)
    rd = torch.zeros(n, dtype=torch.float64, device="cuda")
    zd = torch.zeros_like(rd)
    torch.cuda.synchronize()
    st = L.fem_apply_preconditioner_noalloc(s._h, C.c_void_p(rd.data_ptr()), C.c_void_p(zd.data_ptr()))
    assert st == 8 and b"limit" in L.fem_last_error(s._h)  # FEM_ERR_CAPACITY
    s.close()
