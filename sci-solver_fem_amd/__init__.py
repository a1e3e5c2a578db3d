"""MI355X (gfx950) P1 FEM assembly + smoothed-aggregation AMG-PCG.

Python mirror of the reference's ``FEMSolver`` (src/FEMSolver.h:28-66): the
same public fields (``maxIters_``, ``tolerance_``, ...) with the same defaults,
``solveFEM(x_h, b_h)``, the MATLAB/VTK IO helpers, plus a
``torch.autograd.Function`` whose backward pass is an adjoint solve on the
reused hierarchy and a device element-gradient kernel.

All numerical work runs in ``lib/libfem_mi355x.so`` (HIP, gfx950) through the
C ABI in ``include/fem_mi355x.h``; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _lib
from ._lib import FemParams, FemSolveInfo
from . import meshgen  # noqa: F401  (re-exported)

__all__ = ["FEMSolver", "FemError", "NotConverged", "fem_solve", "FEMSolveFunction", "meshgen",
           "release_cached_memory", "read_matlab_array", "write_matlab_array", "read_tetgen", "read_trimesh",
           "read_matlab_sparse_coo", "device_aggregate", "lib"]


def lib():
    return _lib.load()


class FemError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"[status {status}] {msg}")
        self.status = status


class NotConverged(FemError):
    pass


def _ptr(a) -> int:
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return int(a.data_ptr())  # torch tensor


def _check(h, st):
    if st == _lib.FEM_OK:
        return
    msg = lib().fem_last_error(h).decode(errors="replace")
    if st == _lib.FEM_ERR_NOT_CONVERGED:
        raise NotConverged(st, msg)
    raise FemError(st, msg)


# reference field name -> fem_params field
_FIELDS = {
    "verbose_": "verbose", "maxLevels_": "maxLevels", "maxIters_": "maxIters",
    "preInnerIters_": "preInnerIters", "postInnerIters_": "postInnerIters",
    "postRelaxes_": "postRelaxes", "cycleIters_": "cycleIters", "dsType_": "dsType",
    "topSize_": "topSize", "randMisParameters_": "randMisParameters",
    "partitionMaxSize_": "partitionMaxSize", "aggregatorType_": "aggregatorType",
    "convergeType_": "convergeType", "tolerance_": "tolerance", "cycleType_": "cycleType",
    "solverType_": "solverType", "smootherWeight_": "smootherWeight", "proOmega_": "proOmega",
    "device_": "device", "blockSize_": "blockSize",
    # additions
    "precision_": "precision", "smootherType_": "smootherType",
    "aggregationRule_": "aggregationRule", "seed_": "seed", "strengthThreshold_": "strengthThreshold",
    "launchMode_": "launchMode",
}


class FEMSolver:
    """Drop-in for the reference ``FEMSolver`` (src/FEMSolver.h).

    ``FEMSolver(fname, isTetMesh=True, verbose=False)`` reads ``fname.node`` /
    ``fname.ele`` (or a PLY/OFF triangle mesh) and assembles A on the GPU, like
    src/FEMSolver.cu:9-44.  ``FEMSolver.from_arrays`` builds the same pipeline
    from in-memory meshes with a per-element conductivity.
    """

    def __init__(self, fname: Optional[str] = None, isTetMesh: bool = True, verbose: bool = False,
                 _handle=None):
        L = lib()
        object.__setattr__(self, "_h", None)
        if _handle is None:
            if fname is None:
                raise ValueError("fname is required")
            h = C.c_void_p()
            st = L.fem_create_from_files(fname.encode(), 1 if isTetMesh else 0, 1 if verbose else 0, C.byref(h))
            if st != _lib.FEM_OK:
                raise FemError(st, L.fem_last_error(None).decode(errors="replace"))
            _handle = h
        object.__setattr__(self, "_h", _handle)
        object.__setattr__(self, "filename_", fname or "")
        # bumped whenever the operator changes (update_kappa, a matrix file)
        object.__setattr__(self, "_operator_gen", 0)
        p = FemParams()
        L.fem_get_params(self._h, C.byref(p))
        object.__setattr__(self, "_p", p)

    @classmethod
    def from_arrays(cls, dim: int, xyz: np.ndarray, conn: np.ndarray, kappa: Optional[np.ndarray] = None,
                    mass_coeff: float = 1.0, flags: int = 0, verbose: bool = False) -> "FEMSolver":
        L = lib()
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        conn = np.ascontiguousarray(conn, dtype=np.int32)
        if xyz.ndim != 2 or xyz.shape[1] != 3:
            raise ValueError("xyz must be (nv, 3)")
        if conn.ndim != 2 or conn.shape[1] != dim + 1:
            raise ValueError(f"conn must be (ne, {dim + 1})")
        kp = None
        if kappa is not None:
            kappa = np.ascontiguousarray(kappa, dtype=np.float64)
            if kappa.shape != (conn.shape[0],):
                raise ValueError("kappa must have one entry per element")
            kp = kappa.ctypes.data
        h = C.c_void_p()
        st = L.fem_create_from_arrays(dim, xyz.shape[0], xyz.ctypes.data, conn.shape[0], conn.ctypes.data, kp,
                                      float(mass_coeff), int(flags), 1 if verbose else 0, C.byref(h))
        if st != _lib.FEM_OK:
            raise FemError(st, L.fem_last_error(None).decode(errors="replace"))
        return cls(_handle=h)

    # -- parameter fields ---------------------------------------------------
    def __getattr__(self, name):
        if name in _FIELDS:
            return getattr(self._p, _FIELDS[name])
        raise AttributeError(name)

    def __setattr__(self, name, value):
        if name in _FIELDS:
            setattr(self._p, _FIELDS[name], value)
            _check(self._h, lib().fem_set_params(self._h, C.byref(self._p)))
        else:
            object.__setattr__(self, name, value)

    def set_params(self, **kw):
        for k, v in kw.items():
            key = k if k in _FIELDS else k + "_"
            if key not in _FIELDS:
                raise AttributeError(k)
            setattr(self._p, _FIELDS[key], v)
        _check(self._h, lib().fem_set_params(self._h, C.byref(self._p)))

    def close(self):
        if self._h:
            lib().fem_destroy(self._h)
            object.__setattr__(self, "_h", None)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- reference API ------------------------------------------------------
    def getMatrixRows(self) -> int:
        return int(lib().fem_get_matrix_rows(self._h))

    def nnz(self) -> int:
        return int(lib().fem_get_matrix_nnz(self._h))

    def num_elements(self) -> int:
        return int(lib().fem_get_num_elements(self._h))

    def readMatlabSparseMatrix(self, filename: str) -> int:
        object.__setattr__(self, "_operator_gen", self._operator_gen + 1)
        return int(lib().fem_read_matlab_sparse(self._h, filename.encode()))

    @staticmethod
    def readMatlabArray(filename: str):
        return read_matlab_array(filename)

    @staticmethod
    def writeMatlabArray(filename: str, array) -> int:
        return write_matlab_array(filename, array)

    def solveFEM(self, x_h: np.ndarray, b_h: np.ndarray) -> dict:
        """x_h: initial guess in, solution out (float64, modified in place)."""
        if x_h.dtype != np.float64 or not x_h.flags.c_contiguous:
            raise ValueError("x_h must be a contiguous float64 array")
        b = np.ascontiguousarray(b_h, dtype=np.float64)
        n = self.getMatrixRows()
        k = 1 if b.ndim == 1 else b.shape[1]
        if b.shape[0] != n or x_h.shape != b.shape:
            raise ValueError("b_h and x_h must have getMatrixRows() rows")
        info = FemSolveInfo()
        st = lib().fem_solve(self._h, b.ctypes.data, x_h.ctypes.data, k, 0, C.byref(info))
        d = info.as_dict()
        if st not in (_lib.FEM_OK, _lib.FEM_ERR_NOT_CONVERGED):
            _check(self._h, st)
        d["status"] = st
        return d

    # -- additions ----------------------------------------------------------
    def setup(self):
        _check(self._h, lib().fem_setup(self._h))

    def solve(self, b, x0=None, raise_on_nonconvergence: bool = False):
        """Solve A x = b. numpy in -> numpy out; torch.cuda tensors stay on the device."""
        L = lib()
        info = FemSolveInfo()
        if isinstance(b, np.ndarray):
            bb = np.ascontiguousarray(b, dtype=np.float64)
            x = np.zeros_like(bb) if x0 is None else np.array(x0, dtype=np.float64, copy=True, order="C")
            k = 1 if bb.ndim == 1 else bb.shape[1]
            st = L.fem_solve(self._h, bb.ctypes.data, x.ctypes.data, k, 0, C.byref(info))
        else:
            import torch
            bb = b.contiguous().to(torch.float64)
            x = torch.zeros_like(bb) if x0 is None else x0.detach().to(torch.float64).contiguous().clone()
            k = 1 if bb.dim() == 1 else bb.shape[1]
            self._bind_torch_stream()
            st = L.fem_solve(self._h, _ptr(bb), _ptr(x), k, 1, C.byref(info))
        if st != _lib.FEM_OK and (raise_on_nonconvergence or st != _lib.FEM_ERR_NOT_CONVERGED):
            _check(self._h, st)
        d = info.as_dict()
        d["status"] = st
        return x, d

    def _bind_torch_stream(self):
        """Order the next library call after torch's pending work on its inputs.

        Called after the input and output tensors are produced and right
        before the C call.  On a torch side stream the library queues its work
        on that stream (in order after every producer).  On torch's default
        (null) stream the library keeps its own non-blocking stream, which
        does not wait on the null stream, so the null stream is drained first.
        Every library call returns after its own work is done, so torch's
        later consumers of the outputs need no wait."""
        import torch
        cur = torch.cuda.current_stream()
        s = cur.cuda_stream
        if not s:
            cur.synchronize()
        _check(self._h, lib().fem_set_stream(self._h, C.c_void_p(s)))

    def apply_operator(self, x):
        if isinstance(x, np.ndarray):
            xx = np.ascontiguousarray(x, dtype=np.float64)
            y = np.empty_like(xx)
            k = 1 if xx.ndim == 1 else xx.shape[1]
            _check(self._h, lib().fem_apply_operator(self._h, xx.ctypes.data, y.ctypes.data, k, 0))
            return y
        import torch
        xx = x.contiguous().to(torch.float64)
        y = torch.empty_like(xx)
        k = 1 if xx.dim() == 1 else xx.shape[1]
        self._bind_torch_stream()
        _check(self._h, lib().fem_apply_operator(self._h, _ptr(xx), _ptr(y), k, 1))
        return y

    def apply_preconditioner(self, r: np.ndarray) -> np.ndarray:
        rr = np.ascontiguousarray(r, dtype=np.float64)
        z = np.empty_like(rr)
        _check(self._h, lib().fem_apply_preconditioner(self._h, rr.ctypes.data, z.ctypes.data, 0))
        return z

    def update_kappa(self, kappa):
        if isinstance(kappa, np.ndarray):
            k = np.ascontiguousarray(kappa, dtype=np.float64)
        else:
            k = kappa.detach().to("cpu", dtype=__import__("torch").float64).contiguous().numpy()
        object.__setattr__(self, "_operator_gen", self._operator_gen + 1)
        _check(self._h, lib().fem_update_kappa(self._h, k.ctypes.data))

    def adjoint_grad(self, u, dLdu):
        """dL/dkappa for L(u) with K(kappa) u = b, given dL/du (adjoint on the same hierarchy)."""
        info = FemSolveInfo()
        if isinstance(u, np.ndarray):
            uu = np.ascontiguousarray(u, dtype=np.float64)
            gg = np.ascontiguousarray(dLdu, dtype=np.float64)
            g = np.empty(self.num_elements())
            lam = np.empty_like(uu)
            st = lib().fem_adjoint_grad(self._h, uu.ctypes.data, gg.ctypes.data, g.ctypes.data, lam.ctypes.data, 0,
                                        C.byref(info))
        else:
            import torch
            uu = u.contiguous().to(torch.float64)
            gg = dLdu.contiguous().to(torch.float64)
            g = torch.empty(self.num_elements(), dtype=torch.float64, device=uu.device)
            lam = torch.empty_like(uu)
            self._bind_torch_stream()
            st = lib().fem_adjoint_grad(self._h, _ptr(uu), _ptr(gg), _ptr(g), _ptr(lam), 1, C.byref(info))
        if st not in (_lib.FEM_OK, _lib.FEM_ERR_NOT_CONVERGED):
            _check(self._h, st)
        d = info.as_dict()
        d["status"] = st
        return g, lam, d

    def element_grad(self, a, u):
        """-|T_e| grad(a).grad(u) per element (user element order)."""
        if isinstance(a, np.ndarray):
            aa = np.ascontiguousarray(a, dtype=np.float64)
            uu = np.ascontiguousarray(u, dtype=np.float64)
            g = np.empty(self.num_elements())
            _check(self._h, lib().fem_element_grad(self._h, aa.ctypes.data, uu.ctypes.data, g.ctypes.data, 0))
            return g
        import torch
        aa = a.contiguous().to(torch.float64)
        uu = u.contiguous().to(torch.float64)
        g = torch.empty(self.num_elements(), dtype=torch.float64, device=aa.device)
        self._bind_torch_stream()
        _check(self._h, lib().fem_element_grad(self._h, _ptr(aa), _ptr(uu), _ptr(g), 1))
        return g

    def get_matrix_csr(self):
        """Assembled operator, user numbering: (row_ptr int64, cols int32, vals float64)."""
        n, nnz = self.getMatrixRows(), self.nnz()
        rp = np.empty(n + 1, dtype=np.int64)
        ci = np.empty(nnz, dtype=np.int32)
        v = np.empty(nnz, dtype=np.float64)
        _check(self._h, lib().fem_get_matrix_csr(self._h, rp.ctypes.data, ci.ctypes.data, v.ctypes.data))
        return rp, ci, v

    def get_matrix_scipy(self):
        import scipy.sparse as sp
        rp, ci, v = self.get_matrix_csr()
        n = self.getMatrixRows()
        return sp.csr_matrix((v, ci, rp), shape=(n, n))

    def num_levels(self) -> int:
        return int(lib().fem_get_num_levels(self._h))

    def level_matrix(self, lvl: int):
        """Level operator in internal numbering: (row_ptr, cols, vals, aggregates or None)."""
        n, nnz = C.c_int64(), C.c_int64()
        _check(self._h, lib().fem_get_level_info(self._h, lvl, C.byref(n), C.byref(nnz)))
        rp = np.empty(n.value + 1, dtype=np.int32)
        ci = np.empty(nnz.value, dtype=np.int32)
        v = np.empty(nnz.value, dtype=np.float64)
        agg = np.full(n.value, -1, dtype=np.int32)
        _check(self._h, lib().fem_get_level_matrix(self._h, lvl, rp.ctypes.data, ci.ctypes.data, v.ctypes.data,
                                                   agg.ctypes.data))
        return rp, ci, v, (agg if agg.size and agg.max() >= 0 else None)

    def level_prolongator(self, lvl: int):
        rp_n = C.c_int64()
        n, nnz_a = C.c_int64(), C.c_int64()
        _check(self._h, lib().fem_get_level_info(self._h, lvl, C.byref(n), C.byref(nnz_a)))
        _check(self._h, lib().fem_get_level_prolongator(self._h, lvl, C.byref(rp_n), None, None, None))
        rp = np.empty(n.value + 1, dtype=np.int32)
        ci = np.empty(rp_n.value, dtype=np.int32)
        v = np.empty(rp_n.value, dtype=np.float64)
        _check(self._h, lib().fem_get_level_prolongator(self._h, lvl, C.byref(rp_n), rp.ctypes.data, ci.ctypes.data,
                                                        v.ctypes.data))
        return rp, ci, v

    def apply_level(self, lvl: int, which: int, x: np.ndarray) -> np.ndarray:
        """y = M x for level `lvl`'s operator (which 0), prolongator (1) or
        restriction (2), through the V-cycle's one-column kernels in the
        hierarchy's precision (internal numbering)."""
        n, nc = C.c_int64(), C.c_int64()
        _check(self._h, lib().fem_get_level_info(self._h, lvl, C.byref(n), None))
        if which in (1, 2):
            _check(self._h, lib().fem_get_level_info(self._h, lvl + 1, C.byref(nc), None))
        rows, cols = {0: (n.value, n.value), 1: (n.value, nc.value), 2: (nc.value, n.value),
                      3: (n.value, n.value)}.get(which, (0, 0))
        xx = np.ascontiguousarray(x, dtype=np.float64)
        if xx.shape != (cols,):
            raise ValueError(f"x must have {cols} entries")
        y = np.empty(rows, dtype=np.float64)
        _check(self._h, lib().fem_apply_level(self._h, lvl, which, xx.ctypes.data, y.ctypes.data))
        return y

    def level_smoother(self, lvl: int) -> dict:
        """Smoother data of level `lvl`: rho(D^-1 A) estimate, Chebyshev
        interval [lmin, lmax] and Jacobi weight omega."""
        v = [C.c_double() for _ in range(4)]
        _check(self._h, lib().fem_get_level_smoother(self._h, lvl, *[C.byref(t) for t in v]))
        return dict(zip(("rho", "lmin", "lmax", "omega"), (t.value for t in v)))

    def smoother_step(self, lvl: int, mode: int, c1: float, c2: float, x, b, d):
        """One smoothing step through the V-cycle's kernel (internal numbering):
        returns (y, d_new).  mode 0: y = x + c1 D^-1 (b - A x); mode 1:
        d_new = c1 d + c2 D^-1 (b - A x), y = x + d_new."""
        xx = np.ascontiguousarray(x, dtype=np.float64)
        bb = np.ascontiguousarray(b, dtype=np.float64)
        dd = np.array(d, dtype=np.float64, copy=True, order="C")
        n = C.c_int64()
        _check(self._h, lib().fem_get_level_info(self._h, lvl, C.byref(n), None))
        if not (xx.shape == bb.shape == dd.shape == (n.value,)):
            raise ValueError(f"x, b and d must be vectors of the level's {n.value} rows")
        y = np.empty_like(xx)
        _check(self._h, lib().fem_apply_smoother_step(self._h, lvl, mode, float(c1), float(c2), xx.ctypes.data,
                                                      bb.ctypes.data, dd.ctypes.data, y.ctypes.data))
        return y, dd

    def permutation(self) -> np.ndarray:
        p = np.empty(self.getMatrixRows(), dtype=np.int32)
        _check(self._h, lib().fem_get_permutation(self._h, p.ctypes.data))
        return p

    def profile(self, enable: bool = True):
        _check(self._h, lib().fem_profile_enable(self._h, 1 if enable else 0))

    def profile_stats(self, family: str):
        n = C.c_int64()
        ms = C.c_double()
        by = C.c_double()
        _check(self._h, lib().fem_profile_get(self._h, family.encode(), C.byref(n), C.byref(ms), C.byref(by)))
        return {"launches": n.value, "total_ms": ms.value, "bytes_per_launch": by.value}


# ---- module-level IO helpers (no GPU needed) ------------------------------

def release_cached_memory() -> None:
    """Free the library's reserve of device blocks (temporaries of mesh builds
    and setups kept for reuse, at most 8 GiB per device)."""
    lib().fem_release_cached_memory()


def read_matlab_array(filename: str) -> np.ndarray:
    L = lib()
    n = C.c_int64()
    st = L.fem_read_matlab_array(filename.encode(), None, C.byref(n))
    if st != 0:
        raise FemError(st, L.fem_last_error(None).decode(errors="replace"))
    out = np.empty(n.value)
    L.fem_read_matlab_array(filename.encode(), out.ctypes.data, C.byref(n))
    return out


def write_matlab_array(filename: str, array) -> int:
    a = np.ascontiguousarray(array, dtype=np.float64).ravel()
    return int(lib().fem_write_matlab_array(filename.encode(), a.ctypes.data, a.size))


def read_matlab_sparse_coo(filename: str, round_to_float: bool = False):
    L = lib()
    nr, nc, nnz = C.c_int64(), C.c_int64(), C.c_int64()
    st = L.fem_read_matlab_sparse_coo(filename.encode(), int(round_to_float), C.byref(nr), C.byref(nc), C.byref(nnz),
                                      None, None, None)
    if st != 0:
        raise FemError(st, L.fem_last_error(None).decode(errors="replace"))
    r = np.empty(nnz.value, dtype=np.int32)
    c = np.empty(nnz.value, dtype=np.int32)
    v = np.empty(nnz.value)
    L.fem_read_matlab_sparse_coo(filename.encode(), int(round_to_float), C.byref(nr), C.byref(nc), C.byref(nnz),
                                 r.ctypes.data, c.ctypes.data, v.ctypes.data)
    return (nr.value, nc.value), r, c, v


def read_tetgen(base: str):
    L = lib()
    nv, ne = C.c_int64(), C.c_int64()
    xyz = C.POINTER(C.c_double)()
    tets = C.POINTER(C.c_int32)()
    lab = C.POINTER(C.c_int32)()
    st = L.fem_read_tetgen(base.encode(), C.byref(nv), C.byref(xyz), C.byref(ne), C.byref(tets), C.byref(lab))
    if st != 0:
        raise FemError(st, L.fem_last_error(None).decode(errors="replace"))
    try:
        X = np.ctypeslib.as_array(xyz, shape=(nv.value * 3,)).reshape(-1, 3).copy()
        T = np.ctypeslib.as_array(tets, shape=(ne.value * 4,)).reshape(-1, 4).copy()
        Lb = np.ctypeslib.as_array(lab, shape=(ne.value,)).copy()
    finally:
        L.fem_free(C.cast(xyz, C.c_void_p))
        L.fem_free(C.cast(tets, C.c_void_p))
        L.fem_free(C.cast(lab, C.c_void_p))
    return X, T, Lb


def read_trimesh(filename: str):
    L = lib()
    nv, ne = C.c_int64(), C.c_int64()
    xyz = C.POINTER(C.c_double)()
    tris = C.POINTER(C.c_int32)()
    st = L.fem_read_trimesh(filename.encode(), C.byref(nv), C.byref(xyz), C.byref(ne), C.byref(tris))
    if st != 0:
        raise FemError(st, L.fem_last_error(None).decode(errors="replace"))
    try:
        X = np.ctypeslib.as_array(xyz, shape=(nv.value * 3,)).reshape(-1, 3).copy()
        T = np.ctypeslib.as_array(tris, shape=(ne.value * 3,)).reshape(-1, 3).copy()
    finally:
        L.fem_free(C.cast(xyz, C.c_void_p))
        L.fem_free(C.cast(tris, C.c_void_p))
    return X, T


def device_aggregate(adj_ptr: np.ndarray, adj: np.ndarray, depth: int = 2, min_size: int = 9, rule: int = 0,
                     seed: int = 0x5EED):
    """Run the device randomized-MIS aggregation on a host CSR graph (no diagonal)."""
    L = lib()
    ap = np.ascontiguousarray(adj_ptr, dtype=np.int32)
    aj = np.ascontiguousarray(adj, dtype=np.int32)
    n = ap.size - 1
    out = np.empty(max(n, 1), dtype=np.int32)
    na = C.c_int32()
    st = L.fem_device_aggregate(n, ap.ctypes.data, aj.ctypes.data if aj.size else None, depth, min_size, rule, seed,
                                out.ctypes.data, C.byref(na))
    if st != 0:
        raise FemError(st, L.fem_last_error(None).decode(errors="replace"))
    return out[:n], na.value


def fem_solve(dim, xyz, conn, b, kappa=None, mass_coeff=1.0, **params):
    """One-shot convenience: assemble, set up and solve; returns (x, info)."""
    s = FEMSolver.from_arrays(dim, xyz, conn, kappa=kappa, mass_coeff=mass_coeff)
    s.set_params(**params)
    try:
        return s.solve(b)
    finally:
        s.close()


def _torch_function():
    import torch

    class FEMSolveFunction(torch.autograd.Function):
        """u = K(kappa)^-1 b on the GPU; backward = adjoint solve + element gradient.

        forward(kappa, b, solver): reassembles K for kappa and rebuilds the
        hierarchy; backward reuses that hierarchy (K is symmetric):
            lambda = K^-1 dL/du,  dL/dkappa_e = -|T_e| grad(lambda).grad(u),
            dL/db = lambda.
        When the solver's operator changed after this node's forward (a later
        forward with another kappa on the same solver), backward reassembles
        K for this node's kappa first, so every node's adjoint uses its own K.
        """

        @staticmethod
        def forward(ctx, kappa, b, solver):
            solver.update_kappa(kappa)
            u, info = solver.solve(b, raise_on_nonconvergence=True)
            ctx.solver = solver
            ctx.operator_gen = solver._operator_gen
            ctx.save_for_backward(u, kappa)
            ctx.info = info
            return u

        @staticmethod
        def backward(ctx, grad_u):
            u, kappa = ctx.saved_tensors
            solver = ctx.solver
            if solver._operator_gen != ctx.operator_gen:
                solver.update_kappa(kappa)
                ctx.operator_gen = solver._operator_gen
            lam, _ = solver.solve(grad_u.contiguous(), raise_on_nonconvergence=True)
            g = solver.element_grad(lam, u)
            return g, lam, None

    return FEMSolveFunction


def __getattr__(name):
    if name == "FEMSolveFunction":
        return _torch_function()
    raise AttributeError(name)
