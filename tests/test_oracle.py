"""Pins of the CPU oracle (oracle/fem_oracle.c) against closed forms, invariants
and the reference's own fixtures (tests/golden/reference_data, copied from
src/test/test_data).  CPU only."""
from __future__ import annotations

import itertools
import math

import numpy as np
import pytest
import scipy.special as ssp

from conftest import REFDATA
from matfile import read_mat_v5


# ---- quadrature ---------------------------------------------------------------

@pytest.mark.parametrize("q", [3, 4, 6])
def test_lobatto_exact_to_degree(oracle, q):
    z, w = oracle.gauss("lobatto", q, 0, 0)
    for k in range(2 * q - 2):  # exact up to degree 2q-3
        exact = (1 - (-1) ** (k + 1)) / (k + 1)
        assert abs(np.dot(w, z ** k) - exact) < 1e-13
    # interior nodes are the zeros of P^(1,1)_{q-2}
    ref = np.sort(ssp.roots_jacobi(q - 2, 1, 1)[0])
    assert np.allclose(z[1:-1], ref, atol=1e-14)


@pytest.mark.parametrize("q,a", [(4, 1), (4, 2), (6, 1)])
def test_radau_exact_to_degree(oracle, q, a):
    z, w = oracle.gauss("radau", q, a, 0)
    for k in range(2 * q - 1):  # exact up to degree 2q-2 with weight (1-x)^a
        # int_{-1}^{1} (1-x)^a x^k dx via the binomial expansion of x = 1 - t
        exact = sum(math.comb(k, j) * (-1) ** j * 2 ** (a + j + 1) / (a + j + 1) for j in range(k + 1))
        assert abs(np.dot(w, z ** k) - exact) < 1e-12 * max(1.0, abs(exact))
    ref = np.sort(ssp.roots_jacobi(q - 1, a, 1)[0])
    assert np.allclose(z[1:], ref, atol=1e-14)


def test_mass_integrals_closed_form(oracle):
    # 8 * int_ref phi_k phi_g = 8 (1 + delta) / 120 (tet), 8 (1 + delta) / 24 (triangle)
    t = oracle.tet_mass_integrals()
    exp = [(2 if k == g else 1) / 15 for k in range(4) for g in range(k, 4)]
    assert np.allclose(t, exp, rtol=1e-14, atol=0)
    s = oracle.tri_mass_integrals()
    exp = [(2 if k == g else 1) / 3 for k in range(3) for g in range(k, 3)]
    assert np.allclose(s, exp, rtol=1e-14, atol=0)


# ---- element matrices ------------------------------------------------------------

def _random_tet(rng):
    while True:
        X = rng.normal(size=(4, 3))
        vol = np.linalg.det(np.stack([X[1] - X[0], X[2] - X[0], X[3] - X[0]])) / 6
        if abs(vol) > 0.05:
            return X, abs(vol)


def test_tet_element_against_barycentric_gradients(oracle):
    rng = np.random.default_rng(1)
    for _ in range(20):
        X, vol = _random_tet(rng)
        co = rng.uniform(0.5, 3.0)
        K, M = oracle.tet_element(X, co)
        # independent: barycentric gradients from a numpy inverse of [x y z 1]
        Bary = np.linalg.inv(np.hstack([X, np.ones((4, 1))]))  # column s = coefficients of phi_s
        G = Bary[:3, :].T
        assert np.allclose(K, co * vol * G @ G.T, rtol=1e-12, atol=1e-12)
        assert np.allclose(K.sum(axis=1), 0, atol=1e-12 * np.abs(K).max())
        assert np.allclose(M, vol / 20 * (np.ones((4, 4)) + np.eye(4)), rtol=1e-12)
        # a linear field has zero element "stiffness load" against constants
        u = X @ rng.normal(size=3) + 0.3
        assert abs(np.ones(4) @ K @ u) < 1e-11 * np.abs(K).max()


def test_tri_element_signed_mass(oracle):
    X = np.array([[0, 0, 0], [1.0, 0, 0], [0, 2.0, 0]])
    K, M = oracle.tri_element(X, signed_mass=True)
    area = 1.0
    assert np.allclose(M, area / 12 * (np.ones((3, 3)) + np.eye(3)), rtol=1e-14)
    G = np.array([[-1, -0.5], [1, 0], [0, 0.5]])
    assert np.allclose(K, area * G @ G.T, rtol=1e-14)
    # clockwise ordering flips the sign of the mass with the reference's signed Jacobian
    Xcw = X[[0, 2, 1]]
    K2, M2 = oracle.tri_element(Xcw, signed_mass=True)
    assert np.allclose(M2, -M, rtol=1e-14)
    K3, M3 = oracle.tri_element(Xcw, signed_mass=False)
    assert np.allclose(M3, M, rtol=1e-14)
    assert np.allclose(np.sort(K2.ravel()), np.sort(K.ravel()), rtol=1e-14)


# ---- pattern ---------------------------------------------------------------------

def _read_tetgen_py(base):
    """Independent parser for the fixture meshes (whitespace tables)."""
    node = [l.split() for l in open(str(base) + ".node") if l.strip() and not l.startswith("#")]
    nv = int(node[0][0])
    xyz = np.array([[float(t) for t in r[1:4]] for r in node[1:1 + nv]])
    ele = [l.split() for l in open(str(base) + ".ele") if l.strip() and not l.startswith("#")]
    ne = int(ele[0][0])
    tets = np.array([[int(t) for t in r[1:5]] for r in ele[1:1 + ne]], dtype=np.int64)
    if tets.min() != 0:
        tets -= 1
    return xyz, tets.astype(np.int32)


def _brute_pattern(conn, nv):
    rows = [set([i]) for i in range(nv)]
    for t in conn:
        for a, b in itertools.product(t, t):
            rows[a].add(int(b))
    return rows


def test_pattern_cubemesh_fixture_counts(oracle):
    xyz, tets = _read_tetgen_py(REFDATA / "CubeMesh_size256step16_correct")
    rp, ci = oracle.pattern(tets, xyz.shape[0])
    assert rp[-1] == 66961  # Kuhn 6-tet split of a 17^3 grid
    xyz2, tets2 = _read_tetgen_py(REFDATA / "CubeMesh_size256step16")
    rp2, _ = oracle.pattern(tets2, xyz2.shape[0])
    assert rp2[-1] == 82321


def test_pattern_matches_tetvol_matrix_file(oracle):
    """tetVolA.mat (the reference's SCIRun matrix) has exactly the mesh pattern."""
    xyz, tets = _read_tetgen_py(REFDATA / "tetVol")
    rp, ci = oracle.pattern(tets, xyz.shape[0])
    (nr, nc), rows, cols, vals = read_mat_v5(REFDATA / "tetVolA.mat")
    assert nr == nc == xyz.shape[0]
    file_keys = np.unique(rows.astype(np.int64) * nr + cols)
    ours = np.repeat(np.arange(nr, dtype=np.int64), np.diff(rp)) * nr + ci
    assert rp[-1] == 129844
    assert np.array_equal(np.sort(ours), file_keys)


def test_pattern_bruteforce_small(oracle, meshgen):
    for dim, N in ((3, 4), (2, 5)):
        xyz, conn = meshgen.cube_mesh(N, seed=3) if dim == 3 else meshgen.square_mesh(N, seed=3)
        rp, ci = oracle.pattern(conn, xyz.shape[0])
        rows = _brute_pattern(conn, xyz.shape[0])
        for i, r in enumerate(rows):
            assert list(ci[rp[i]:rp[i + 1]]) == sorted(r)


# ---- assembly invariants ------------------------------------------------------------

def test_assembly_invariants_3d(oracle, meshgen):
    xyz, tets = meshgen.cube_mesh(7, seed=5)
    kappa = meshgen.lognormal_kappa(5, tets.shape[0], 1.0)
    A = oracle.scipy_matrix(xyz, tets, kappa, mass_coeff=0.0)
    M = oracle.scipy_matrix(xyz, tets, np.zeros(tets.shape[0]), mass_coeff=1.0)
    # symmetry, zero row sums of the stiffness, mass integrates to vol = 1
    assert abs(A - A.T).max() < 1e-14 * abs(A).max()
    assert np.abs(A @ np.ones(xyz.shape[0])).max() < 1e-12 * abs(A).max()
    assert abs(M.sum() - 1.0) < 1e-13
    assert abs(np.ones(xyz.shape[0]) @ (M @ xyz[:, 0]) - 0.5) < 1e-13  # int x dV
    # P1 reproduces linear fields: interior rows of S u_lin vanish (constant kappa)
    S = oracle.scipy_matrix(xyz, tets, np.full(tets.shape[0], 2.5), mass_coeff=0.0)
    u = xyz @ np.array([0.3, -1.2, 0.7]) + 2.0
    r = S @ u
    interior = np.all((xyz > 1e-12) & (xyz < 1 - 1e-12), axis=1)
    assert np.abs(r[interior]).max() < 1e-11 * abs(A).max()
    assert np.abs(r[~interior]).max() > 1e-6  # boundary rows carry the flux


def test_assembly_invariants_2d(oracle, meshgen):
    xyz, tris = meshgen.square_mesh(9, seed=2)
    S = oracle.scipy_matrix(xyz, tris, None, mass_coeff=0.0)
    M = oracle.scipy_matrix(xyz, tris, np.zeros(tris.shape[0]), mass_coeff=1.0)
    assert abs(M.sum() - 1.0) < 1e-13
    u = xyz[:, :2] @ np.array([1.5, -0.4])
    interior = np.all((xyz[:, :2] > 1e-12) & (xyz[:, :2] < 1 - 1e-12), axis=1)
    assert np.abs((S @ u)[interior]).max() < 1e-11
    # mass + stiffness combine linearly
    A = oracle.scipy_matrix(xyz, tris, None, mass_coeff=1.0)
    assert abs(A - S - M).max() < 1e-15


def test_element_grad_is_kappa_derivative(oracle, meshgen):
    xyz, tets = meshgen.cube_mesh(4, seed=1)
    rng = np.random.default_rng(0)
    a = rng.normal(size=xyz.shape[0])
    u = rng.normal(size=xyz.shape[0])
    g = oracle.element_grad(xyz, tets, a, u)
    # -a^T (dK/dkappa_e) u by assembling with kappa = e_e (finite element by element)
    for e in (0, 7, 31):
        ke = np.zeros(tets.shape[0])
        ke[e] = 1.0
        Ke = oracle.scipy_matrix(xyz, tets, ke, mass_coeff=0.0)
        assert abs(g[e] + a @ (Ke @ u)) < 1e-12 * (1 + abs(g[e]))


def test_meshgen_closed_forms(meshgen):
    xyz, tets = meshgen.cube_mesh(5, seed=9)
    assert tets.shape == (6 * 4 ** 3, 4)
    p = xyz[tets]
    vol = np.einsum("ij,ij->i", p[:, 1] - p[:, 0], np.cross(p[:, 2] - p[:, 0], p[:, 3] - p[:, 0])) / 6
    assert np.all(vol > 0) and abs(vol.sum() - 1.0) < 1e-13
    xyz2, tris = meshgen.square_mesh(5, seed=9)
    q = xyz2[tris][:, :, :2]
    ar = 0.5 * ((q[:, 1, 0] - q[:, 0, 0]) * (q[:, 2, 1] - q[:, 0, 1]) - (q[:, 1, 1] - q[:, 0, 1]) * (q[:, 2, 0] - q[:, 0, 0]))
    assert np.all(ar > 0) and abs(ar.sum() - 1.0) < 1e-13
    # reproducible streams
    assert np.array_equal(meshgen.normal(1, 3, 10), meshgen.normal(1, 3, 10))
    g = meshgen.normal(7, 3, 200000)
    assert abs(g.mean()) < 0.01 and abs(g.std() - 1) < 0.01


def test_meshgen_streams_pinned(meshgen):
    """Generated inputs are a fixed function of the seed (SHA-256 prefixes of
    the arrays' bytes), and the blocked stream evaluation does not depend on
    where a block boundary falls."""
    import hashlib

    def digest(*arrays):
        h = hashlib.sha256()
        for a in arrays:
            h.update(np.ascontiguousarray(a).tobytes())
        return h.hexdigest()[:16]

    assert digest(*meshgen.cube_mesh(9, 0)) == "e695c7879396b64e"
    assert digest(*meshgen.square_mesh(9, 0)) == "8b29eef556f20bbf"
    assert digest(meshgen.lognormal_kappa(0, 1000, 2.0)) == "8b5e28dfad1a1b40"
    assert digest(meshgen.gaussian_rhs(0, 300, 3)) == "0d88381bf3f7a8ce"
    c = meshgen._CHUNK
    long = meshgen.splitmix64(4, 2, c + 10)
    assert np.array_equal(meshgen.splitmix64(4, 2, 20, offset=c - 10), long[c - 10:c + 10])
    g = meshgen.normal(3, 1, c + 7)
    u = meshgen.uniform(3, 1, 2 * (c + 7))
    ref = np.sqrt(-2.0 * np.log(1.0 - u[:c + 7])) * np.cos(2.0 * np.pi * u[c + 7:])
    assert np.array_equal(g, ref)
