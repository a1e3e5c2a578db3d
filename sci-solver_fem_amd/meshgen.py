"""Seeded synthetic P1 meshes, coefficients and loads.

Every random quantity comes from a counter-based SplitMix64 stream written
here in numpy, so the same seed gives bit-identical inputs on every host.

Meshes (the reference ships only file meshes; these follow the Kuhn split of
its own `CubeMesh_size256step16_correct` fixture):

* ``cube_mesh(N)``   -- unit cube, N^3 vertex grid, each cube split into the
  6 Kuhn tetrahedra around the main diagonal, interior vertices jittered
  uniformly by +-jitter*h, vertex ids randomly permuted.
* ``square_mesh(N)`` -- unit square, N^2 grid, each cell split into 2
  counter-clockwise triangles along one diagonal, same jitter/permutation.

Tetrahedra and triangles are emitted with positive orientation; the
generator asserts every element measure is > 0.
"""
from __future__ import annotations

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)

# stream ids: one per random quantity, so changing one never shifts another
STREAM_JITTER = 1
STREAM_PERM = 2
STREAM_KAPPA = 3
STREAM_RHS = 4
STREAM_NOISE = 5


_CHUNK = 1 << 18  # elements per block: keeps the temporaries cache-resident


def _mix(z: np.ndarray) -> np.ndarray:
    """SplitMix64 finaliser, in place on z (uint64)."""
    t = np.empty_like(z)
    np.right_shift(z, np.uint64(30), out=t)
    z ^= t
    z *= _M1
    np.right_shift(z, np.uint64(27), out=t)
    z ^= t
    z *= _M2
    np.right_shift(z, np.uint64(31), out=t)
    z ^= t
    return z


def _stream_key(seed: int, stream: int) -> np.uint64:
    with np.errstate(over="ignore"):
        return _mix(np.array([(seed * 0x100000001B3 + stream * 0x632BE59BD9B4E019)
                              & 0xFFFFFFFFFFFFFFFF], dtype=np.uint64))[0]


def _words(key: np.uint64, lo: int, hi: int) -> np.ndarray:
    """Words with counters lo+1 .. hi of the stream with this key."""
    with np.errstate(over="ignore"):
        z = np.arange(lo + 1, hi + 1, dtype=np.uint64)
        z *= _GOLDEN
        z += key
        return _mix(z)


def splitmix64(seed: int, stream: int, n: int, offset: int = 0) -> np.ndarray:
    """n 64-bit words of the counter-based stream (seed, stream)."""
    key = _stream_key(seed, stream)
    out = np.empty(n, dtype=np.uint64)
    for a in range(0, n, _CHUNK):
        b = min(n, a + _CHUNK)
        out[a:b] = _words(key, offset + a, offset + b)
    return out


def _unit(w: np.ndarray) -> np.ndarray:
    w >>= np.uint64(11)
    u = w.astype(np.float64)
    u *= 1.0 / 9007199254740992.0
    return u


def uniform(seed: int, stream: int, n: int, offset: int = 0) -> np.ndarray:
    """U[0,1) doubles with 53 random bits."""
    key = _stream_key(seed, stream)
    out = np.empty(n, dtype=np.float64)
    for a in range(0, n, _CHUNK):
        b = min(n, a + _CHUNK)
        out[a:b] = _unit(_words(key, offset + a, offset + b))
    return out


def normal(seed: int, stream: int, n: int) -> np.ndarray:
    """N(0,1) by Box-Muller on pairs of uniforms (cosine branch only):
    g_i = sqrt(-2 ln(1 - u_i)) cos(2 pi u_{n+i}), u = uniform(seed, stream)."""
    key = _stream_key(seed, stream)
    out = np.empty(n, dtype=np.float64)
    for a in range(0, n, _CHUNK):
        b = min(n, a + _CHUNK)
        u1 = 1.0 - _unit(_words(key, a, b))  # (0, 1]
        u2 = _unit(_words(key, n + a, n + b))
        out[a:b] = np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)
    return out


def random_permutation(seed: int, stream: int, n: int) -> np.ndarray:
    keys = splitmix64(seed, stream, n)
    return np.argsort(keys, kind="stable").astype(np.int64)


# Kuhn split: for each ordering of the axes, walk corner 0 -> 7 along unit steps.
_AXIS_ORDERS = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]


def _kuhn_local():
    """6 tets as corner indices c = dx + 2*dy + 4*dz of the unit cube."""
    tets = []
    for order in _AXIS_ORDERS:
        c = [0, 0, 0]
        path = [0]
        for ax in order:
            c[ax] = 1
            path.append(c[0] + 2 * c[1] + 4 * c[2])
        tets.append(path)
    return np.array(tets, dtype=np.int64)


def _local_kuhn_oriented():
    """The 6 Kuhn tets of the unit cube with corners 1 and 2 swapped where the
    corner order is negatively oriented, so every tet has positive volume."""
    local = _kuhn_local()
    corner = np.array([[c & 1, (c >> 1) & 1, (c >> 2) & 1] for c in range(8)], dtype=np.float64)
    out = local.copy()
    for t in range(local.shape[0]):
        p = corner[local[t]]
        if np.linalg.det(p[1:] - p[0]) < 0:
            out[t, 1], out[t, 2] = local[t, 2], local[t, 1]
    return out


def cube_mesh(N: int, seed: int = 0, jitter: float = 0.2, permute: bool = True):
    """Unit-cube Kuhn tet mesh on an N^3 vertex grid.

    Returns (xyz float64 [nv,3], tets int32 [ne,4]) with nv = N^3 and
    ne = 6 (N-1)^3.  Grid vertex (i,j,k) is i*N*N + j*N + k before the
    permutation; tets are emitted cell by cell (cells in the same order), 6
    per cell.  Every tet is checked to have positive volume after the jitter.
    """
    h = 1.0 / (N - 1)
    nv = N * N * N
    g = np.arange(N, dtype=np.float64) * h
    X = np.repeat(g, N * N)
    Y = np.tile(np.repeat(g, N), N)
    Z = np.tile(g, N * N)
    if jitter:
        inner = np.zeros(N, dtype=bool)
        inner[1:N - 1] = True
        interior = (inner[:, None, None] & inner[None, :, None] & inner[None, None, :]).ravel()
        u = uniform(seed, STREAM_JITTER, 3 * nv).reshape(-1, 3)
        for d, C in enumerate((X, Y, Z)):
            C[interior] += ((2.0 * u[interior, d] - 1.0) * (jitter * h))
        del u, interior
    if permute:
        perm = random_permutation(seed, STREAM_PERM, nv)  # new id -> old id
        inv = np.empty(nv, dtype=np.int32)
        inv[perm] = np.arange(nv, dtype=np.int32)
    corner_off = np.array([dx * N * N + dy * N + dz
                           for dz in (0, 1) for dy in (0, 1) for dx in (0, 1)], dtype=np.int64)
    local = corner_off[_local_kuhn_oriented()]  # (6, 4) vertex offsets from the cell's base
    nc1 = N - 1
    ncell = nc1 ** 3
    tets = np.empty((6 * ncell, 4), dtype=np.int32)
    step = max(1, _CHUNK // 6)
    for a in range(0, ncell, step):
        b = min(ncell, a + step)
        c = np.arange(a, b, dtype=np.int64)
        ci, rem = np.divmod(c, nc1 * nc1)
        cj, ck = np.divmod(rem, nc1)
        base = ci * N * N + cj * N + ck
        t = (base[:, None, None] + local[None, :, :]).reshape(-1, 4)
        c0, c1, c2, c3 = (t[:, i] for i in range(4))
        ax, ay, az = X[c1] - X[c0], Y[c1] - Y[c0], Z[c1] - Z[c0]
        bx, by, bz = X[c2] - X[c0], Y[c2] - Y[c0], Z[c2] - Z[c0]
        cx, cy, cz = X[c3] - X[c0], Y[c3] - Y[c0], Z[c3] - Z[c0]
        det = ax * (by * cz - bz * cy) - ay * (bx * cz - bz * cx) + az * (bx * cy - by * cx)
        if not np.all(det > 0):
            raise ValueError("generated mesh has a degenerate or inverted tetrahedron")
        tets[6 * a:6 * b] = inv[t] if permute else t
    xyz = np.stack([X, Y, Z], axis=1)
    if permute:
        xyz = xyz[perm]
    return np.ascontiguousarray(xyz), tets


def square_mesh(N: int, seed: int = 0, jitter: float = 0.2, permute: bool = True):
    """Unit-square triangle mesh on an N^2 grid, 2 CCW triangles per cell.

    Returns (xyz float64 [nv,3] with z = 0, tris int32 [ne,3]).
    """
    h = 1.0 / (N - 1)
    g = np.arange(N, dtype=np.float64) * h
    X, Y = np.meshgrid(g, g, indexing="ij")  # vertex (i,j) -> i*N + j
    xyz = np.stack([X.ravel(), Y.ravel(), np.zeros(N * N)], axis=1)
    ij = np.stack(np.meshgrid(np.arange(N), np.arange(N), indexing="ij"), -1).reshape(-1, 2)
    interior = np.all((ij > 0) & (ij < N - 1), axis=1)
    if jitter:
        u = uniform(seed, STREAM_JITTER, 2 * xyz.shape[0]).reshape(-1, 2)
        xyz[interior, :2] += ((2.0 * u - 1.0) * (jitter * h))[interior]
    c = np.arange(N - 1)
    ci, cj = np.meshgrid(c, c, indexing="ij")
    v00 = (ci * N + cj).ravel()
    v10 = v00 + N
    v01 = v00 + 1
    v11 = v00 + N + 1
    tris = np.empty((2 * v00.size, 3), dtype=np.int64)
    tris[0::2] = np.stack([v00, v10, v11], 1)
    tris[1::2] = np.stack([v00, v11, v01], 1)
    a = xyz[tris[:, 1], :2] - xyz[tris[:, 0], :2]
    b = xyz[tris[:, 2], :2] - xyz[tris[:, 0], :2]
    det = a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0]
    if not np.all(det > 0):
        raise ValueError("generated mesh has a degenerate or clockwise triangle")
    if permute:
        perm = random_permutation(seed, STREAM_PERM, xyz.shape[0])
        inv = np.empty_like(perm)
        inv[perm] = np.arange(perm.size)
        xyz = xyz[perm]
        tris = inv[tris]
    return np.ascontiguousarray(xyz), np.ascontiguousarray(tris.astype(np.int32))


def lognormal_kappa(seed: int, ne: int, sigma: float) -> np.ndarray:
    """Per-element conductivity exp(sigma * g), g ~ N(0,1) i.i.d."""
    return np.exp(sigma * normal(seed, STREAM_KAPPA, ne))


def gaussian_rhs(seed: int, nv: int, nrhs: int = 1) -> np.ndarray:
    """N(0,1) load per vertex; shape (nv,) or (nv, nrhs) row-major."""
    g = normal(seed, STREAM_RHS, nv * nrhs)
    return g if nrhs == 1 else g.reshape(nv, nrhs)


def sine_load_2d(xyz: np.ndarray, tris: np.ndarray) -> np.ndarray:
    """b_i = sum_T sum_q w_q f(x_q) phi_i(x_q), f = sin(2 pi x) sin(2 pi y).

    3-point interior rule: reference points (1/6,1/6), (2/3,1/6), (1/6,2/3),
    weights area/3.
    """
    p0 = xyz[tris[:, 0], :2]
    e1 = xyz[tris[:, 1], :2] - p0
    e2 = xyz[tris[:, 2], :2] - p0
    area = 0.5 * np.abs(e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0])
    b = np.zeros(xyz.shape[0])
    for (xi, eta) in ((1 / 6, 1 / 6), (2 / 3, 1 / 6), (1 / 6, 2 / 3)):
        q = p0 + xi * e1 + eta * e2
        f = np.sin(2 * np.pi * q[:, 0]) * np.sin(2 * np.pi * q[:, 1]) * (area / 3.0)
        phis = (1.0 - xi - eta, xi, eta)
        for a in range(3):
            np.add.at(b, tris[:, a], f * phis[a])
    return b


def noise(seed: int, n: int, std: float) -> np.ndarray:
    return std * normal(seed, STREAM_NOISE, n)


# Named workloads (sizes and coefficients of the five benchmark problems).
CONFIGS = {
    "c1_square65_fp64": dict(dim=2, N=65, sigma=None, mass=1.0, rhs="sine", nrhs=1),
    "c2_cube49_fp64": dict(dim=3, N=49, sigma=1.0, mass=1e-2, rhs="gauss", nrhs=1),
    "c3_cube97_mixed": dict(dim=3, N=97, sigma=1.5, mass=1e-2, rhs="gauss", nrhs=1),
    "c4_cube49_adjoint": dict(dim=3, N=49, sigma=1.0, mass=1e-2, rhs="gauss", nrhs=1),
    "c5_cube161_mixed16": dict(dim=3, N=161, sigma=2.0, mass=1e-2, rhs="gauss", nrhs=16),
}


def make_problem(dim: int, N: int, sigma, mass: float, rhs: str, nrhs: int = 1, seed: int = 0):
    """Mesh + coefficient + load for one workload; returns a dict."""
    if dim == 2:
        xyz, conn = square_mesh(N, seed)
    else:
        xyz, conn = cube_mesh(N, seed)
    ne = conn.shape[0]
    kappa = np.ones(ne) if sigma is None else lognormal_kappa(seed, ne, sigma)
    if rhs == "sine":
        b = sine_load_2d(xyz, conn)
    else:
        b = gaussian_rhs(seed, xyz.shape[0], nrhs)
    return dict(dim=dim, xyz=xyz, conn=conn, kappa=kappa, mass=mass, b=b)
