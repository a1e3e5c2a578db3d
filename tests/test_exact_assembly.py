"""Pins of the extended-precision assembly reference (tests/exact_assembly.py),
which the GPU assembly is checked against at sizes where the FP64 oracle's
absolute-coordinate cofactor inverse (the reference's formula) loses digits.
CPU only."""
from __future__ import annotations

import numpy as np
import pytest

import exact_assembly as ea


def _mesh(meshgen, N, sigma, seed=3):
    xyz, tets = meshgen.cube_mesh(N, seed=seed)
    kappa = meshgen.lognormal_kappa(seed, tets.shape[0], sigma)
    return xyz, tets, kappa


def test_matches_oracle_on_a_coarse_mesh(oracle, meshgen):
    """At h = 1/4 the oracle's FP64 formula loses at most a few bits, so the
    two agree to 1e-13 of max|K| (a wrong sign, a transposed inverse or a
    dropped mass term would differ at O(1))."""
    xyz, tets, kappa = _mesh(meshgen, 5, 1.0)
    rp, ci, ov = oracle.assemble(xyz, tets, kappa, 0.3)
    v = ea.assemble(xyz, tets, kappa, 0.3, rp, ci).astype(np.float64)
    assert np.abs(v - ov).max() <= 1e-13 * np.abs(ov).max()


def test_linear_fields_are_harmonic_at_interior_rows(oracle, meshgen):
    """Stiffness only (mass 0, constant coefficient): K u = 0 at every
    interior row for u linear in x, y, z (the gradients of the basis
    reproduce a constant gradient)."""
    N = 7
    xyz, tets, _ = _mesh(meshgen, N, 1.5)
    kappa = np.full(tets.shape[0], 2.5)
    rp, ci, _ = oracle.assemble(xyz, tets, kappa, 0.0)
    v = ea.assemble(xyz, tets, kappa, 0.0, rp, ci)
    X = xyz.astype(ea.LD)  # u exactly linear in the coordinates the assembly used
    u = ea.LD(0.3) * X[:, 0] - ea.LD(1.7) * X[:, 1] + ea.LD(0.9) * X[:, 2] + ea.LD(2.0)
    n = xyz.shape[0]
    rows = np.repeat(np.arange(n), np.diff(rp))
    Ku = np.zeros(n, dtype=ea.LD)
    np.add.at(Ku, rows, v * u[ci])
    interior = np.all((xyz > 1e-12) & (xyz < 1 - 1e-12), axis=1)
    assert interior.sum() == (N - 2) ** 3
    scale = np.abs(v).max() * np.abs(u).max()
    assert float(np.abs(Ku[interior]).max()) <= 1e-17 * float(scale)
    # boundary rows carry the flux and do not vanish
    assert float(np.abs(Ku[~interior]).max()) > 1e-3 * float(scale)


def test_mass_sums_to_the_volume_and_matrix_is_symmetric(oracle, meshgen):
    """kappa = 0, mass 1: the entries sum to the volume of the unit cube
    (sum_ab (1 + delta_ab) / 20 = 1 per element), and K = K^T to extended
    rounding."""
    xyz, tets, _ = _mesh(meshgen, 6, 1.0)
    rp, ci, _ = oracle.assemble(xyz, tets, None, 1.0)
    v = ea.assemble(xyz, tets, np.zeros(tets.shape[0]), 1.0, rp, ci)
    assert abs(float(v.sum()) - 1.0) <= 1e-15
    n = xyz.shape[0]
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(rp))
    keys = rows * n + ci
    tpos = np.searchsorted(keys, ci.astype(np.int64) * n + rows)  # position of (j, i)
    assert np.array_equal(keys[tpos], ci.astype(np.int64) * n + rows)
    # the element sums of (i, j) and (j, i) run in different orders: equal to extended rounding
    assert float(np.abs(v - v[tpos]).max()) <= 1e-18 * float(np.abs(v).max())


@pytest.mark.parametrize("N", [9, 17])
def test_oracle_formula_loses_digits_with_refinement(oracle, meshgen, N):
    """The reference's absolute-coordinate cofactor inverse, restated by the
    oracle in FP64, drifts from the extended-precision values as h shrinks
    (the error grows like eps (|x| / h)^2 or faster), while an FP64 sum of
    the extended-precision element values stays at rounding: the documented
    reason the GPU-oracle bar is a multiple of max|K| rather than a few ulps."""
    xyz, tets, kappa = _mesh(meshgen, N, 2.0)
    rp, ci, ov = oracle.assemble(xyz, tets, kappa, 1e-2)
    v = ea.assemble(xyz, tets, kappa, 1e-2, rp, ci)
    scale = float(np.abs(v).max())
    err_oracle = float(np.abs(ov.astype(ea.LD) - v).max()) / scale
    err_rounded = float(np.abs(v.astype(np.float64).astype(ea.LD) - v).max()) / scale
    assert err_rounded <= 1.2e-16
    assert err_oracle > 4 * err_rounded


def _square(meshgen, N, sigma, seed=3):
    xy, tris = meshgen.square_mesh(N, seed=seed)
    return xy, tris, meshgen.lognormal_kappa(seed, tris.shape[0], sigma)


def test_triangles_match_oracle_on_a_coarse_mesh(oracle, meshgen):
    xy, tris, kappa = _square(meshgen, 5, 1.0)
    rp, ci, ov = oracle.assemble(xy, tris, kappa, 0.3)
    v = ea.assemble(xy, tris, kappa, 0.3, rp, ci).astype(np.float64)
    assert np.abs(v - ov).max() <= 1e-13 * np.abs(ov).max()


def test_triangles_harmonic_and_area(oracle, meshgen):
    """Constant coefficient, mass 0: K u = 0 at interior rows for linear u;
    coefficient 0, mass 1: the entries sum to the area of the unit square."""
    N = 9
    xy, tris, _ = _square(meshgen, N, 1.0)
    rp, ci, _ = oracle.assemble(xy, tris, None, 0.0)
    v = ea.assemble(xy, tris, np.full(tris.shape[0], 1.5), 0.0, rp, ci)
    X = xy.astype(ea.LD)
    u = ea.LD(0.7) * X[:, 0] - ea.LD(1.3) * X[:, 1] + ea.LD(0.5)
    n = xy.shape[0]
    rows = np.repeat(np.arange(n), np.diff(rp))
    Ku = np.zeros(n, dtype=ea.LD)
    np.add.at(Ku, rows, v * u[ci])
    interior = np.all((xy[:, :2] > 1e-12) & (xy[:, :2] < 1 - 1e-12), axis=1)
    assert interior.sum() == (N - 2) ** 2
    scale = float(np.abs(v).max() * np.abs(u).max())
    assert float(np.abs(Ku[interior]).max()) <= 1e-17 * scale
    assert float(np.abs(Ku[~interior]).max()) > 1e-3 * scale
    m = ea.assemble(xy, tris, np.zeros(tris.shape[0]), 1.0, rp, ci)
    assert abs(float(m.sum()) - 1.0) <= 1e-15
