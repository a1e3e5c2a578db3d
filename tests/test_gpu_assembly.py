"""GPU parity of the device pattern + assembly kernels against the oracle.

Bar: pattern bit-exact; values within 1e-13 of max|A| of the oracle (whose
FP64 restatement of the reference's element formula loses digits as h
shrinks) and within 1e-15 of max|A| of an extended-precision assembly; A
bitwise symmetric and bitwise reproducible run to run."""
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


@pytest.mark.parametrize("dim,N,sigma", [(3, 25, 2.0), (3, 41, 2.0), (2, 257, 1.5)])
def test_values_match_extended_precision(fem, gpu, oracle, meshgen, dim, N, sigma):
    """Against the assembly in extended precision (tests/exact_assembly.py):
    the GPU's FP64 values are within 1e-15 of max|K| (measured 1e-16 to 3e-16,
    a few FP64 roundings of the element terms an entry sums).  The oracle's
    FP64 restatement of the reference's formula (cofactors of absolute
    coordinates) is 4e-14 to 7e-14 of max|K| from the same values on these
    tet meshes and 7e-12 on the 257^2 triangles, and drifts further as the
    mesh is refined; the oracle comparisons above run on meshes coarse enough
    for their 1e-13 bar."""
    import exact_assembly as ea
    xyz, conn = meshgen.cube_mesh(N, seed=5) if dim == 3 else meshgen.square_mesh(N, seed=5)
    kappa = meshgen.lognormal_kappa(5, conn.shape[0], sigma)
    s, rp, ci, v = _gpu_matrix(fem, dim, xyz, conn, kappa, 1e-2)
    ref = ea.assemble(xyz, conn, kappa, 1e-2, rp, ci)
    scale = float(np.abs(ref).max())
    err_gpu = float(np.abs(v.astype(ea.LD) - ref).max()) / scale
    _, _, ov = oracle.assemble(xyz, conn, kappa, 1e-2)
    err_oracle = float(np.abs(ov.astype(ea.LD) - ref).max()) / scale
    report(f"[gpu] dim={dim} N={N} sigma={sigma}: |A_gpu - K_ext| / max|K| = {err_gpu:.2e}, oracle {err_oracle:.2e}")
    assert err_gpu <= 1e-15
    s.close()


def _hub_meshes(meshgen):
    """Meshes whose hub vertex has a row longer than 255 entries (the slices
    holding it sum in place in the CSR values): a 300^2 grid with a hub
    joined to its 300 bottom-edge vertices by a fan of triangles, and the
    cone of tets from an apex over every triangle of a planar grid (the apex
    row holds all 22^2 grid vertices)."""
    xy, tris = meshgen.square_mesh(300, seed=2)
    bottom = np.flatnonzero(xy[:, 1] < 1e-12)
    bottom = bottom[np.argsort(xy[bottom, 0])]
    hub = xy.shape[0]
    xy2 = np.vstack([xy, [[0.5, -0.5, 0.0]]])
    fan = np.stack([np.full(bottom.size - 1, hub), bottom[1:], bottom[:-1]], 1)
    tris2 = np.ascontiguousarray(np.concatenate([tris, fan]).astype(np.int32))
    g, t = meshgen.square_mesh(22, seed=4)
    apex = g.shape[0]
    xyz3 = np.vstack([g, [[0.5, 0.5, 1.0]]])
    tets3 = np.ascontiguousarray(np.concatenate([t, np.full((t.shape[0], 1), apex)], 1).astype(np.int32))
    return [(2, xy2, tris2), (3, xyz3, tets3)]


def test_long_rows_sum_in_place(fem, gpu, oracle, meshgen):
    """Rows over 255 entries: pattern bit-exact, values within 1e-15 of
    max|A| of the extended-precision assembly and within 1e-12 of the oracle
    (whose FP64 formula is 6.3e-13 of max|A| off the extended-precision
    values on the 300^2 grid), A bitwise symmetric."""
    import exact_assembly as ea
    import scipy.sparse as sp
    for dim, xyz, conn in _hub_meshes(meshgen):
        report(f"[gpu] assembly with a hub row, dim={dim}, {conn.shape[0]} elements")
        kappa = meshgen.lognormal_kappa(9, conn.shape[0], 1.0)
        s, rp, ci, v = _gpu_matrix(fem, dim, xyz, conn, kappa, 0.5)
        assert np.diff(rp).max() > 255
        orp, oci, ov = oracle.assemble(xyz, conn, kappa, 0.5)
        assert np.array_equal(rp, orp) and np.array_equal(ci, oci)
        assert np.abs(v - ov).max() <= 1e-12 * np.abs(ov).max()
        ref = ea.assemble(xyz, conn, kappa, 0.5, rp, ci)
        assert float(np.abs(v.astype(ea.LD) - ref).max()) <= 1e-15 * float(np.abs(ref).max())
        A = sp.csr_matrix((v, ci, rp), shape=(xyz.shape[0],) * 2)
        assert np.count_nonzero((A - A.T).tocoo().data) == 0
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


def _label_table(labels):
    """Material label -> coefficient of perform_element_loop_3D.cuh:585-610:
    labels 0 and 1 give 1, labels 2..6 give the label itself."""
    co = np.ones(labels.shape, dtype=np.float64)
    hi = (labels >= 2) & (labels <= 6)
    co[hi] = labels[hi]
    return co


def test_labelled_tetgen_file_matches_oracle(fem, gpu, oracle, meshgen, tmp_path):
    """A TetGen mesh whose .ele carries material labels 0..6 (1-based
    indices, as TetGen writes them), through the file constructor: the
    element coefficients follow the reference's label table and the mass
    coefficient is 1; coordinates are parsed as float as the reference does."""
    report("[gpu] labelled TetGen file (labels 0..6)")
    xyz, tets = meshgen.cube_mesh(6, seed=21)
    rng = np.random.default_rng(21)
    labels = rng.integers(0, 7, size=tets.shape[0])
    assert set(np.unique(labels)) == set(range(7))
    base = tmp_path / "labelled"
    with open(str(base) + ".node", "w") as f:
        f.write(f"{xyz.shape[0]} 3 0 0\n")
        for i, p in enumerate(xyz):
            # FP32 values written exactly (9 significant digits identify a float)
            q = p.astype(np.float32)
            f.write(f"{i + 1} {q[0]:.9g} {q[1]:.9g} {q[2]:.9g}\n")
    with open(str(base) + ".ele", "w") as f:
        f.write(f"{tets.shape[0]} 4 1\n")
        for e, t in enumerate(tets):
            f.write(f"{e + 1} {t[0] + 1} {t[1] + 1} {t[2] + 1} {t[3] + 1} {labels[e]}\n")
    xr, tr, lr = fem.read_tetgen(str(base))
    assert np.array_equal(tr, tets) and np.array_equal(lr, labels)
    assert np.array_equal(xr, xyz.astype(np.float32).astype(np.float64))
    s = fem.FEMSolver(str(base), True, False)
    rp, ci, v = s.get_matrix_csr()
    orp, oci, ov = oracle.assemble(xr, tets, _label_table(labels), 1.0)
    assert np.array_equal(rp, orp) and np.array_equal(ci, oci)
    assert np.abs(v - ov).max() <= 1e-13 * np.abs(ov).max()
    # the same mesh with every label 1 differs: the coefficients were applied
    orp1, _, ov1 = oracle.assemble(xr, tets, None, 1.0)
    assert np.abs(ov1 - ov).max() > 0.1 * np.abs(ov).max()
    s.close()


def test_signed_mass_2d_matches_oracle(fem, gpu, oracle, meshgen, tmp_path):
    """2D mass with the signed Jacobian (perform_element_loop_2D.cuh:92):
    clockwise triangles get negative mass.  Through fem_create_from_arrays
    with flags bit 0 and through a triangle-mesh file (OFF), against the
    oracle's signed element matrices; without the flag the mass is unsigned."""
    report("[gpu] 2D signed mass")
    xyz, tris = meshgen.square_mesh(9, seed=5)
    rng = np.random.default_rng(5)
    flip = rng.random(tris.shape[0]) < 0.3
    tris = tris.copy()
    tris[flip] = tris[flip][:, [0, 2, 1]]  # clockwise
    for signed in (True, False):
        s, rp, ci, v = _gpu_matrix(fem, 2, xyz, tris, None, 1.0, flags=1 if signed else 0)
        orp, oci, ov = oracle.assemble(xyz, tris, None, 1.0, signed_mass=signed)
        assert np.array_equal(ci, oci)
        assert np.abs(v - ov).max() <= 1e-13 * np.abs(ov).max()
        s.close()
    orp, oci, ov_s = oracle.assemble(xyz, tris, None, 1.0, signed_mass=True)
    _, _, ov_u = oracle.assemble(xyz, tris, None, 1.0, signed_mass=False)
    assert np.abs(ov_s - ov_u).max() > 1e-6  # the flag matters on this mesh
    # file path: TriMesh-style float points, signed mass, mass coefficient 1
    p = tmp_path / "flip.off"
    lines = ["OFF", f"{xyz.shape[0]} {tris.shape[0]} 0"]
    lines += [f"{q[0]:.9g} {q[1]:.9g} 0" for q in xyz.astype(np.float32)]
    lines += ["3 " + " ".join(str(int(i)) for i in t) for t in tris]
    p.write_text("\n".join(lines) + "\n")
    s = fem.FEMSolver(str(p), False, False)
    rp, ci, v = s.get_matrix_csr()
    xf = xyz.astype(np.float32).astype(np.float64)
    xf[:, 2] = 0.0
    orp, oci, ov = oracle.assemble(xf, tris, None, 1.0, signed_mass=True)
    assert np.array_equal(ci, oci)
    assert np.abs(v - ov).max() <= 1e-13 * np.abs(ov).max()
    s.close()
