"""Shared fixtures: package loader, oracle loader, GPU marker."""
from __future__ import annotations

import ctypes as C
import importlib.util
import os
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
GOLDEN = ROOT / "tests" / "golden"
REFDATA = GOLDEN / "reference_data"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libfem_mi355x.so)")
    config.addinivalue_line("markers", "slow: larger problem sizes")


def load_package():
    name = "sci_solver_fem_amd"
    if name in sys.modules:
        return sys.modules[name]
    pkg_dir = ROOT / "sci-solver_fem_amd"
    spec = importlib.util.spec_from_file_location(name, pkg_dir / "__init__.py",
                                                  submodule_search_locations=[str(pkg_dir)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def fem():
    return load_package()


@pytest.fixture(scope="session")
def meshgen(fem):
    return fem.meshgen


def load_oracle():
    """Oracle shared library (builds it with make if missing)."""
    so = ROOT / "oracle" / "liboracle.so"
    if not so.exists():
        subprocess.run(["make", "-C", str(ROOT / "oracle")], check=True, capture_output=True)
    return Oracle(C.CDLL(str(so)))


class Oracle:
    """Thin numpy front end of oracle/fem_oracle.c."""

    def __init__(self, lib):
        self.lib = lib
        V = C.c_void_p
        lib.orc_pattern.restype = C.c_int64
        lib.orc_pattern.argtypes = [C.c_int, C.c_int64, C.c_int64, V, V, V]
        lib.orc_assemble.restype = C.c_int
        lib.orc_assemble.argtypes = [C.c_int, C.c_int64, V, C.c_int64, V, V, C.c_double, C.c_int, V, V, V]
        lib.orc_tet_mass_integrals.argtypes = [V]
        lib.orc_tri_mass_integrals.argtypes = [V]
        lib.orc_tet_element.argtypes = [V, C.c_double, V, V, V]
        lib.orc_tri_element.argtypes = [V, C.c_int, V, V, V]
        lib.orc_gauss_lobatto_jacobi.argtypes = [C.c_int, C.c_double, C.c_double, V, V]
        lib.orc_gauss_radau_jacobi.argtypes = [C.c_int, C.c_double, C.c_double, V, V]
        lib.orc_element_grad.argtypes = [C.c_int, V, C.c_int64, V, V, V, V]
        lib.orc_mis_seed.restype = C.c_uint32
        lib.orc_mis_seed.argtypes = [C.c_uint64, C.c_uint32]
        lib.orc_mis.argtypes = [C.c_int64, V, V, C.c_int, C.c_uint64, V]
        lib.orc_aggregate.restype = C.c_int64
        lib.orc_aggregate.argtypes = [C.c_int64, V, V, C.c_int, C.c_int, C.c_int, C.c_uint64, V]
        lib.orc_induced_graph.argtypes = [C.c_int64, V, V, V, C.c_int64, C.POINTER(C.POINTER(C.c_int32)),
                                          C.POINTER(C.POINTER(C.c_int32))]
        lib.orc_free.argtypes = [V]
        lib.orc_condition.restype = C.c_int64
        lib.orc_condition.argtypes = [C.c_int64, V, V, V, C.c_int64, C.c_int, C.c_int]
        lib.orc_flood_fill_mis.argtypes = [C.c_int64, V, V, C.c_int, C.c_int64, V]
        lib.orc_aggregate_to_nearest.restype = C.c_int64
        lib.orc_aggregate_to_nearest.argtypes = [C.c_int64, V, V, V, V]
        lib.orc_amg_pcg.restype = C.c_int
        lib.orc_amg_pcg.argtypes = [C.c_int64, V, V, V, V, V, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int,
                                    C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, C.c_uint64, V]

    @staticmethod
    def _p(a):
        return a.ctypes.data if a is not None else None

    def pattern(self, conn, nv):
        conn = np.ascontiguousarray(conn, dtype=np.int32)
        npe = conn.shape[1]
        rp = np.zeros(nv + 1, dtype=np.int64)
        nnz = self.lib.orc_pattern(npe, nv, conn.shape[0], conn.ctypes.data, rp.ctypes.data, None)
        ci = np.zeros(nnz, dtype=np.int32)
        self.lib.orc_pattern(npe, nv, conn.shape[0], conn.ctypes.data, rp.ctypes.data, ci.ctypes.data)
        return rp, ci

    def assemble(self, xyz, conn, kappa=None, mass_coeff=1.0, signed_mass=False):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        conn = np.ascontiguousarray(conn, dtype=np.int32)
        dim = conn.shape[1] - 1
        nv = xyz.shape[0]
        rp, ci = self.pattern(conn, nv)
        vals = np.zeros(ci.size)
        k = None if kappa is None else np.ascontiguousarray(kappa, dtype=np.float64)
        st = self.lib.orc_assemble(dim, nv, xyz.ctypes.data, conn.shape[0], conn.ctypes.data, self._p(k),
                                   float(mass_coeff), int(signed_mass), rp.ctypes.data, ci.ctypes.data,
                                   vals.ctypes.data)
        assert st == 0
        return rp, ci, vals

    def scipy_matrix(self, xyz, conn, kappa=None, mass_coeff=1.0, signed_mass=False):
        import scipy.sparse as sp
        rp, ci, v = self.assemble(xyz, conn, kappa, mass_coeff, signed_mass)
        n = xyz.shape[0]
        return sp.csr_matrix((v, ci, rp), shape=(n, n))

    def tet_mass_integrals(self):
        o = np.zeros(10)
        self.lib.orc_tet_mass_integrals(o.ctypes.data)
        return o

    def tri_mass_integrals(self):
        o = np.zeros(6)
        self.lib.orc_tri_mass_integrals(o.ctypes.data)
        return o

    def tet_element(self, X, co=1.0):
        X = np.ascontiguousarray(X, dtype=np.float64)
        K = np.zeros(16)
        M = np.zeros(16)
        integ = self.tet_mass_integrals()
        self.lib.orc_tet_element(X.ctypes.data, float(co), integ.ctypes.data, K.ctypes.data, M.ctypes.data)
        return K.reshape(4, 4), M.reshape(4, 4)

    def tri_element(self, X, signed_mass=False):
        X = np.ascontiguousarray(X, dtype=np.float64)
        K = np.zeros(9)
        M = np.zeros(9)
        integ = self.tri_mass_integrals()
        self.lib.orc_tri_element(X.ctypes.data, int(signed_mass), integ.ctypes.data, K.ctypes.data, M.ctypes.data)
        return K.reshape(3, 3), M.reshape(3, 3)

    def gauss(self, kind, q, a, b):
        z = np.zeros(q)
        w = np.zeros(q)
        fn = self.lib.orc_gauss_lobatto_jacobi if kind == "lobatto" else self.lib.orc_gauss_radau_jacobi
        fn(q, float(a), float(b), z.ctypes.data, w.ctypes.data)
        return z, w

    def mis(self, ap, aj, depth, seed):
        ap = np.ascontiguousarray(ap, dtype=np.int32)
        aj = np.ascontiguousarray(aj, dtype=np.int32)
        out = np.empty(ap.size - 1, dtype=np.int32)
        self.lib.orc_mis(ap.size - 1, ap.ctypes.data, aj.ctypes.data, depth, seed, out.ctypes.data)
        return out

    def aggregate(self, ap, aj, depth=2, min_size=9, rule=0, seed=0x5EED):
        ap = np.ascontiguousarray(ap, dtype=np.int32)
        aj = np.ascontiguousarray(aj, dtype=np.int32)
        out = np.empty(ap.size - 1, dtype=np.int32)
        na = self.lib.orc_aggregate(ap.size - 1, ap.ctypes.data, aj.ctypes.data, depth, min_size, rule, seed,
                                    out.ctypes.data)
        return out, int(na)

    def condition(self, ap, aj, agg, nagg, min_size, max_size):
        """AggMIS merge-split conditioning (orc_condition); returns (agg, nagg)."""
        ap = np.ascontiguousarray(ap, dtype=np.int32)
        aj = np.ascontiguousarray(aj, dtype=np.int32)
        out = np.array(agg, dtype=np.int32, copy=True)
        na = self.lib.orc_condition(ap.size - 1, ap.ctypes.data, aj.ctypes.data, out.ctypes.data, nagg, min_size,
                                    max_size)
        return out, int(na)

    def flood_fill(self, ap, aj, k, starter):
        """Flood-fill distance-k MIS (orc_flood_fill_mis) and aggregation to the
        nearest root (orc_aggregate_to_nearest); returns (roots, agg, nagg)."""
        ap = np.ascontiguousarray(ap, dtype=np.int32)
        aj = np.ascontiguousarray(aj, dtype=np.int32)
        n = ap.size - 1
        roots = np.empty(n, dtype=np.int32)
        agg = np.empty(n, dtype=np.int32)
        self.lib.orc_flood_fill_mis(n, ap.ctypes.data, aj.ctypes.data, k, starter, roots.ctypes.data)
        na = self.lib.orc_aggregate_to_nearest(n, ap.ctypes.data, aj.ctypes.data, roots.ctypes.data, agg.ctypes.data)
        return roots, agg, int(na)

    def induced_graph(self, ap, aj, agg, nc):
        ap = np.ascontiguousarray(ap, dtype=np.int32)
        aj = np.ascontiguousarray(aj, dtype=np.int32)
        agg = np.ascontiguousarray(agg, dtype=np.int32)
        cp = C.POINTER(C.c_int32)()
        cj = C.POINTER(C.c_int32)()
        self.lib.orc_induced_graph(ap.size - 1, ap.ctypes.data, aj.ctypes.data, agg.ctypes.data, nc, C.byref(cp),
                                   C.byref(cj))
        P = np.ctypeslib.as_array(cp, shape=(nc + 1,)).copy()
        J = np.ctypeslib.as_array(cj, shape=(max(int(P[-1]), 1),))[:P[-1]].copy()
        self.lib.orc_free(C.cast(cp, C.c_void_p))
        self.lib.orc_free(C.cast(cj, C.c_void_p))
        return P, J

    def amg_pcg(self, xyz, conn, kappa, mass, b, tol=1e-8, max_iters=500, top_size=256, max_levels=100, pre=5,
                post=5, relaxes=1, smoother_weight=1.0, pro_omega=0.67, part_max=512, seed=0x5EED, A=None):
        """Reference AMG-PCG on the host; returns a dict with x and timings."""
        if A is None:
            rp, ci, v = self.assemble(xyz, conn, kappa, mass)
        else:
            rp, ci, v = A
            rp = np.ascontiguousarray(rp, dtype=np.int64)
            ci = np.ascontiguousarray(ci, dtype=np.int32)
            v = np.ascontiguousarray(v, dtype=np.float64)
        n = rp.size - 1
        bb = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(n)
        out = np.zeros(8)
        st = self.lib.orc_amg_pcg(n, rp.ctypes.data, ci.ctypes.data, v.ctypes.data, bb.ctypes.data, x.ctypes.data,
                                  tol, max_iters, top_size, max_levels, pre, post, relaxes, smoother_weight,
                                  pro_omega, part_max, seed, out.ctypes.data)
        return dict(x=x, status=st, iterations=int(out[0]), setup_s=out[1], solve_s=out[2], rel_residual=out[3],
                    threads=int(out[4]), levels=int(out[5]), operator_complexity=out[6], converged=bool(out[7]))

    def element_grad(self, xyz, conn, a, u):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        conn = np.ascontiguousarray(conn, dtype=np.int32)
        out = np.zeros(conn.shape[0])
        self.lib.orc_element_grad(conn.shape[1] - 1, xyz.ctypes.data, conn.shape[0], conn.ctypes.data,
                                  np.ascontiguousarray(a, dtype=np.float64).ctypes.data,
                                  np.ascontiguousarray(u, dtype=np.float64).ctypes.data, out.ctypes.data)
        return out


@pytest.fixture(scope="session")
def oracle():
    return load_oracle()


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu(fem):
    """Skip-free GPU gate: a gpu-marked test on a machine without a GPU fails loudly."""
    if not gpu_available():
        pytest.fail("gpu test run without a visible GPU")
    return True


def pytest_collection_modifyitems(config, items):
    # gpu tests are selected with -m gpu; nothing is silently skipped
    pass


def report(msg: str):
    print(msg, flush=True)
