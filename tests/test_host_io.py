"""Host-side checks of libfem_mi355x.so that need no GPU: the C ABI exports,
MATLAB / TetGen / PLY / VTK IO against the reference fixtures and independent
parsers, and the error codes of malformed inputs."""
from __future__ import annotations

import re
import struct

import numpy as np
import pytest
import scipy.io as sio

from conftest import REFDATA, ROOT
from matfile import read_mat_v5


def test_library_exports_every_header_symbol(fem):
    lib = fem.lib()
    header = (ROOT / "include" / "fem_mi355x.h").read_text()
    names = set(re.findall(r"\b(fem_[a-z0-9_]+)\s*\(", header))
    assert len(names) >= 30
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    # the ctypes table covers exactly the header
    assert names == set(fem._lib.SIGNATURES)


def test_params_defaults_match_reference(fem):
    import ctypes as C
    p = fem.FemParams()
    fem.lib().fem_params_default(C.byref(p))
    # src/FEMSolver.cu:11-33
    assert (p.maxLevels, p.maxIters, p.preInnerIters, p.postInnerIters, p.postRelaxes) == (100, 100, 5, 5, 1)
    assert (p.topSize, p.randMisParameters, p.partitionMaxSize, p.solverType) == (256, 90102, 512, 0)
    assert (p.tolerance, p.smootherWeight, p.proOmega, p.blockSize) == (1e-6, 1.0, 0.67, 256)


def test_read_matlab_array_matches_scipy(fem):
    for f in ("simpleb.mat", "simpleAns.mat", "simpleTrib.mat", "tetVolb.mat", "tetVolAns.mat"):
        ours = fem.read_matlab_array(str(REFDATA / f))
        d = {k: v for k, v in sio.loadmat(str(REFDATA / f)).items() if not k.startswith("__")}
        ref = np.asarray(next(iter(d.values()))).ravel()
        assert np.array_equal(ours, ref), f


def test_read_matlab_sparse_matches_independent_parsers(fem):
    for f in ("simple.mat", "simpleTri.mat", "tetVolA.mat"):
        (nr, nc), r, c, v = fem.read_matlab_sparse_coo(str(REFDATA / f))
        (nr2, nc2), r2, c2, v2 = read_mat_v5(REFDATA / f)
        assert (nr, nc) == (nr2, nc2)
        assert np.array_equal(r, r2) and np.array_equal(c, c2) and np.array_equal(v, v2)
        if f != "tetVolA.mat":  # scipy cannot read tetVolA's name element
            S = sio.loadmat(str(REFDATA / f))
            M = next(v for k, v in S.items() if not k.startswith("__")).tocoo()
            assert M.nnz == v.size
    # fixture facts: simple.mat = 79.9568 I, simpleTri.mat = 26.1327 I
    (_, _), r, c, v = fem.read_matlab_sparse_coo(str(REFDATA / "simple.mat"))
    assert np.all(r == c) and np.allclose(v, v[0]) and abs(v[0] - 79.9568) < 1e-3
    # float rounding as in src/FEMSolver.cu:309
    (_, _), _, _, vf = fem.read_matlab_sparse_coo(str(REFDATA / "tetVolA.mat"), round_to_float=True)
    (_, _), _, _, vd = fem.read_matlab_sparse_coo(str(REFDATA / "tetVolA.mat"))
    assert np.array_equal(vf, vd.astype(np.float32).astype(np.float64))


def test_write_matlab_array_layout(fem, tmp_path):
    x = np.linspace(-1, 1, 37)
    p = tmp_path / "out.mat"
    assert fem.write_matlab_array(str(p), x) == 0
    raw = p.read_bytes()
    assert raw[:116].decode().rstrip() == "MATLAB 5.0 MAT-file, Platform: GLNXA64, Created by SCI-Solver_FEM."
    assert raw[124:128] == b"\x00\x01IM"
    # header size field = bytes after it (src/FEMSolver.cu:509-513)
    assert struct.unpack_from("<ii", raw, 128) == (14, len(raw) - 136)
    assert len(raw) == 136 + 48 + 8 * x.size
    d = sio.loadmat(str(p))
    assert np.array_equal(d["x_h"].ravel(), x)
    assert np.array_equal(fem.read_matlab_array(str(p)), x)


def test_read_tetgen_fixture(fem):
    xyz, tets, lab = fem.read_tetgen(str(REFDATA / "simple"))
    assert xyz.shape == (5000, 3) and tets.shape == (14406, 4)
    assert tets.min() == 0 and tets.max() == 4999
    # coordinates are parsed as float like the reference (tetmesh.cu:285-292)
    assert xyz[1, 0] == float(np.float32(0.0204082))
    xyz2, tets2, _ = fem.read_tetgen(str(REFDATA / "tetVol"))
    assert tets2.shape == (46305, 4) and xyz2.shape[0] == 10648


def test_read_trimesh_fixtures(fem):
    xyz, tris = fem.read_trimesh(str(REFDATA / "simple.ply"))
    assert xyz.shape == (2500, 3) and tris.shape == (4802, 3)
    xyz, tris = fem.read_trimesh(str(REFDATA / "sphere_290verts.ply"))
    assert xyz.shape == (290, 3) and tris.shape == (576, 3)


def test_malformed_inputs_return_codes(fem, tmp_path):
    with pytest.raises(fem.FemError):
        fem.read_tetgen(str(tmp_path / "missing"))
    # MAT header announcing compression (type 15) -> reference code 1 / -1
    bad = tmp_path / "z.mat"
    bad.write_bytes(b" " * 128 + struct.pack("<ii", 15, 0))
    with pytest.raises(fem.FemError) as e:
        fem.read_matlab_sparse_coo(str(bad))
    assert e.value.status == 1
    with pytest.raises(fem.FemError) as e:
        fem.read_matlab_array(str(bad))
    assert e.value.status == -1
    # a dense file is not a sparse matrix
    with pytest.raises(fem.FemError):
        fem.read_matlab_sparse_coo(str(REFDATA / "simpleb.mat"))


def test_write_vtk(fem, tmp_path):
    import ctypes as C
    xyz, tets, _ = fem.read_tetgen(str(REFDATA / "CubeMesh_size256step16_correct"))
    vals = np.arange(xyz.shape[0], dtype=np.float64)
    base = str(tmp_path / "cube")
    st = fem.lib().fem_write_vtk(base.encode(), 3, xyz.shape[0], np.ascontiguousarray(xyz).ctypes.data, tets.shape[0],
                                 np.ascontiguousarray(tets).ctypes.data, vals.ctypes.data, vals.size)
    assert st == 0
    txt = open(base + ".vtk").read().split("\n")
    assert txt[4] == f"POINTS {xyz.shape[0]} float"
    assert f"CELLS {tets.shape[0]} {tets.shape[0] * 5}" in txt
    assert txt.count("10") == tets.shape[0]


def test_bench_algorithmic_bytes_of_the_chebyshev_step():
    """The bench line prices a level-0 Chebyshev step as a CSR SpMV (4-byte values
    and columns, x read and x_new written) plus D^-1, b and the direction d read
    and written.  At the 161^3 cube the SpMV part is the 545.9 MB quoted for an
    FP32 CSR SpMV there, and the nnz is the Kuhn-mesh closed form."""
    import sys
    sys.path.insert(0, str(ROOT))
    import bench
    N = 161
    n = N ** 3
    # P1 pattern of the Kuhn split: each interior vertex couples to itself and
    # 14 neighbours; edges = 3 N^2 (N-1) axis + 3 N (N-1)^2 face + (N-1)^3 body
    edges = 3 * N * N * (N - 1) + 3 * N * (N - 1) ** 2 + (N - 1) ** 3
    nnz = n + 2 * edges
    assert nnz == 61_979_041
    spmv = nnz * 8 + 4 * (n + 1) + 2 * n * 4
    assert round(spmv / 1e6, 1) == 545.9
    assert bench.cheb_step_bytes(n, nnz) == spmv + 4 * n * 4 == 612_684_200


def test_bench_algorithmic_bytes_of_assembly():
    """Assembly at the 161^3 cube: connectivity and kappa per tet (589.8 MB),
    coordinates (100.2 MB) and one FP64 value per entry (495.8 MB)."""
    import sys
    sys.path.insert(0, str(ROOT))
    import bench
    n, ne, nnz = 161 ** 3, 6 * 160 ** 3, 61_979_041
    assert ne == 24_576_000
    assert round(bench.assembly_bytes(n, ne, nnz) / 1e6, 1) == 1185.8
