"""The C++ drop-in (include/FEMSolver.h) through the example drivers that keep
the reference examples' command lines (src/examples/example1.cu:12-69,
example2.cu:8-65)."""
from __future__ import annotations

import subprocess

import numpy as np
import pytest

from conftest import REFDATA, ROOT, report

LIB = ROOT / "sci-solver_fem_amd" / "lib"


def test_header_compiles(tmp_path):
    src = tmp_path / "t.cpp"
    src.write_text('#include "FEMSolver.h"\nint main(){ return sizeof(FEMSolver) > 0 ? 0 : 1; }\n')
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-fsyntax-only", "-I", str(ROOT / "include"), str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
def test_example1_cubemesh(fem, gpu, oracle, tmp_path):
    report("[gpu] example1 CubeMesh PCG")
    base = str(REFDATA / "CubeMesh_size256step16_correct")
    r = subprocess.run([str(LIB / "example1"), "-i", base, "-pcg"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    x = fem.read_matlab_array(str(tmp_path / "output.mat"))
    xyz, tets, _ = fem.read_tetgen(base)
    A = oracle.scipy_matrix(xyz, tets, None, 1.0)
    b = np.ones(xyz.shape[0])
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) < 1e-9
    vtk = (tmp_path / "CubeMesh_size256step16_correct.vtk").read_text().split("\n")
    assert vtk[4] == f"POINTS {xyz.shape[0]} float"


@pytest.mark.gpu
def test_example1_tetvol_matrix_files(fem, gpu, tmp_path):
    report("[gpu] example1 tetVol -A -b")
    r = subprocess.run([str(LIB / "example1"), "-i", str(REFDATA / "tetVol"), "-A", str(REFDATA / "tetVolA.mat"),
                        "-b", str(REFDATA / "tetVolb.mat")], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    x = fem.read_matlab_array(str(tmp_path / "output.mat"))
    ans = fem.read_matlab_array(str(REFDATA / "tetVolAns.mat"))
    assert np.linalg.norm(x - ans) < 25.0  # src/test/tetVol.cc:24


@pytest.mark.gpu
def test_example2_simple_ply(fem, gpu, tmp_path):
    """src/test/sanity2D.cc through the C++ driver: ||x - ans|| < 100."""
    report("[gpu] example2 simple.ply -A -b")
    r = subprocess.run([str(LIB / "example2"), "-i", str(REFDATA / "simple.ply"), "-A",
                        str(REFDATA / "simpleTri.mat"), "-b", str(REFDATA / "simpleTrib.mat")], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    x = fem.read_matlab_array(str(tmp_path / "output.mat"))
    ans = fem.read_matlab_array(str(REFDATA / "simpleTriAns.mat"))
    assert np.linalg.norm(x - ans) < 100.0
    # the reference strips the folder and the extension (example2.cu:56-63)
    assert (tmp_path / "simple.vtk").exists() and not (tmp_path / "simple.ply.vtk").exists()
