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


def test_corrupt_sizes_return_codes_not_aborts(fem, tmp_path):
    """Header counts and element sizes taken from a file never drive an
    allocation by themselves: an absurd or negative count is a status code
    (the reference's 1 / -1 for MAT files), not an abort through the C ABI."""
    # .node announcing 2e9 vertices but holding two
    (tmp_path / "huge.node").write_text("2000000000 3 0 0\n1 0 0 0\n2 1 0 0\n")
    (tmp_path / "huge.ele").write_text("1 4 0\n1 1 2 1 2\n")
    with pytest.raises(fem.FemError):
        fem.read_tetgen(str(tmp_path / "huge"))
    # .ele announcing more elements than it holds
    (tmp_path / "short.node").write_text("4 3 0 0\n1 0 0 0\n2 1 0 0\n3 0 1 0\n4 0 0 1\n")
    (tmp_path / "short.ele").write_text("900000000 4 0\n1 1 2 3 4\n")
    with pytest.raises(fem.FemError):
        fem.read_tetgen(str(tmp_path / "short"))
    # PLY / OFF with absurd counts
    (tmp_path / "big.off").write_text("OFF\n-5 1 0\n")
    with pytest.raises(fem.FemError):
        fem.read_trimesh(str(tmp_path / "big.off"))
    (tmp_path / "big2.off").write_text("OFF\n3000000000 1 0\n0 0 0\n")
    with pytest.raises(fem.FemError):
        fem.read_trimesh(str(tmp_path / "big2.off"))
    (tmp_path / "poly.off").write_text("OFF\n3 1 0\n0 0 0\n1 0 0\n0 1 0\n2000000000 0 1 2\n")
    with pytest.raises(fem.FemError):
        fem.read_trimesh(str(tmp_path / "poly.off"))

    # MAT v5 sparse: a valid header whose row-index element claims 2 GB
    good = (REFDATA / "simple.mat").read_bytes()
    (_, _), r, c, v = fem.read_matlab_sparse_coo(str(REFDATA / "simple.mat"))
    ir_at = _mat_sparse_ir_offset(good)
    for size in (0x7FFFFFF0, -8):
        bad = bytearray(good)
        struct.pack_into("<i", bad, ir_at + 4, size)
        p = tmp_path / f"bad_ir_{size & 0xffff}.mat"
        p.write_bytes(bytes(bad))
        with pytest.raises(fem.FemError) as e:
            fem.read_matlab_sparse_coo(str(p))
        assert e.value.status == 1
    # truncated copies of a valid sparse and a valid dense file
    for cut in (140, 200, len(good) // 2, len(good) - 3):
        p = tmp_path / f"trunc_{cut}.mat"
        p.write_bytes(good[:cut])
        with pytest.raises(fem.FemError) as e:
            fem.read_matlab_sparse_coo(str(p))
        assert e.value.status == 1
    dense = (REFDATA / "simpleb.mat").read_bytes()
    for cut in (150, len(dense) - 5):
        p = tmp_path / f"trunc_dense_{cut}.mat"
        p.write_bytes(dense[:cut])
        with pytest.raises(fem.FemError) as e:
            fem.read_matlab_array(str(p))
        assert e.value.status == -1
    # dense data element claiming 4 GB
    bad = bytearray(dense)
    struct.pack_into("<I", bad, len(dense) - 8 * 5000 - 4, 0xFFFFFFF8)
    p = tmp_path / "dense_huge.mat"
    p.write_bytes(bytes(bad))
    with pytest.raises(fem.FemError) as e:
        fem.read_matlab_array(str(p))
    assert e.value.status == -1


def _mat_sparse_ir_offset(raw: bytes) -> int:
    """Byte offset of the row-index data element tag in a MAT v5 sparse file
    (after the 128-byte header, the matrix tag, array flags, dimensions and
    the name element)."""
    off = 128 + 8          # matrix tag
    off += 16              # array flags element
    off += 16              # dimensions element (two int32)
    t, n = struct.unpack_from("<HH", raw, off)
    if n:                  # small data element: 4-byte tag + data padded to 4
        off += 4 + ((n + 3) // 4) * 4
    else:
        t, n = struct.unpack_from("<ii", raw, off)
        off += 8 + ((n + 7) // 8) * 8
    return off


def test_read_off_triangles_and_polygons(fem, tmp_path):
    """TriMesh's OFF reader (TriMesh_io.cu:692-708): counts on the header or
    the next line, float vertices, polygon faces split into a fan."""
    p = tmp_path / "quad.off"
    p.write_text("OFF\n# comment line\n5 2 0\n0 0 0\n1 0 0\n1 1 0\n0 1 0\n0.5 0.5 0.25\n"
                 "4 0 1 2 3\n3 0 1 4\n")
    xyz, tris = fem.read_trimesh(str(p))
    assert xyz.shape == (5, 3) and xyz[4, 2] == 0.25
    assert tris.tolist() == [[0, 1, 2], [0, 2, 3], [0, 1, 4]]
    p2 = tmp_path / "inline.off"
    p2.write_text("OFF 3 1 0\n0 0 0\n1 0 0\n0 1 0\n3 0 1 2\n")
    xyz, tris = fem.read_trimesh(str(p2))
    assert xyz.shape == (3, 3) and tris.tolist() == [[0, 1, 2]]
    # the same triangle mesh through PLY and OFF reads identically
    xs, ts = fem.read_trimesh(str(REFDATA / "sphere_290verts.ply"))
    lines = ["OFF", f"{xs.shape[0]} {ts.shape[0]} 0"]
    lines += [" ".join(f"{c:.9g}" for c in row) for row in xs]
    lines += ["3 " + " ".join(str(int(i)) for i in t) for t in ts]
    p3 = tmp_path / "sphere.off"
    p3.write_text("\n".join(lines) + "\n")
    xo, to = fem.read_trimesh(str(p3))
    assert np.array_equal(to, ts) and np.array_equal(xo, xs)
    # out-of-range face index
    p4 = tmp_path / "badidx.off"
    p4.write_text("OFF\n3 1 0\n0 0 0\n1 0 0\n0 1 0\n3 0 1 7\n")
    with pytest.raises(fem.FemError):
        fem.read_trimesh(str(p4))
