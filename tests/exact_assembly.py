"""P1 tet and triangle assembly in extended precision (numpy longdouble: 64-bit
mantissa on x86-64), the accuracy reference for the FP64 assemblies.

The element matrices are those of the reference's element loops
(src/core/cuda/perform_element_loop_3D.cuh:481-627 and
perform_element_loop_2D.cuh:288-372, unsigned mass: stiffness
kappa |T| grad(phi_a) . grad(phi_b), mass |T| (1 + delta_ab) / 20 for tets and
/ 12 for triangles, scaled by the mass coefficient), computed from the edge
vectors p_i - p_0 with 11 more bits than FP64.  The reference forms the basis
gradients from cofactors of the absolute-coordinate matrix rows
(x_i, y_i, z_i, 1); in FP64 those minors cancel for elements small against
their distance from the origin (relative error about eps (|x| / h)^2 to
eps (|x| / h)^3), while in longdouble the cancellation leaves the result
accurate far below FP64 rounding, so this module is a reference for both FP64
formulations.  Test infrastructure.
"""
from __future__ import annotations

import numpy as np

LD = np.longdouble


def tet_gradients(xyz: np.ndarray, tets: np.ndarray):
    """Per element: basis gradients g[e, a, :] and volumes |T| (longdouble)."""
    P = xyz.astype(LD)[tets]  # (ne, 4, 3)
    E = P[:, 1:, :] - P[:, :1, :]  # rows e_i = p_i - p_0, i = 1..3
    # phi_i (i = 1..3) has gradient g_i with e_j . g_i = delta_ij: G = E^-T,
    # i.e. the rows of G are the columns of E^-1, by cofactors
    a, b, c = E[:, 0, :], E[:, 1, :], E[:, 2, :]
    cbc = np.cross(b, c)
    cca = np.cross(c, a)
    cab = np.cross(a, b)
    det = np.einsum("ij,ij->i", a, cbc)
    g = np.empty((tets.shape[0], 4, 3), dtype=LD)
    g[:, 1, :] = cbc / det[:, None]
    g[:, 2, :] = cca / det[:, None]
    g[:, 3, :] = cab / det[:, None]
    g[:, 0, :] = -(g[:, 1, :] + g[:, 2, :] + g[:, 3, :])
    return g, np.abs(det) / LD(6)


def tri_gradients(xy: np.ndarray, tris: np.ndarray):
    """Per triangle: basis gradients g[e, a, :] and areas |T| (longdouble)."""
    P = xy[:, :2].astype(LD)[tris]  # (ne, 3, 2)
    a = P[:, 1, :] - P[:, 0, :]
    b = P[:, 2, :] - P[:, 0, :]
    det = a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0]
    g = np.empty((tris.shape[0], 3, 2), dtype=LD)
    # e_j . g_i = delta_ij for the edges a, b: g_1 = (b_y, -b_x) / det, g_2 = (-a_y, a_x) / det
    g[:, 1, 0], g[:, 1, 1] = b[:, 1] / det, -b[:, 0] / det
    g[:, 2, 0], g[:, 2, 1] = -a[:, 1] / det, a[:, 0] / det
    g[:, 0, :] = -(g[:, 1, :] + g[:, 2, :])
    return g, np.abs(det) / LD(2)


def assemble(xyz: np.ndarray, tets: np.ndarray, kappa: np.ndarray, mass_coeff: float, rp: np.ndarray,
             ci: np.ndarray) -> np.ndarray:
    """Values of K = sum_e kappa_e S_e + mass_coeff M_e on the CSR pattern
    (rp, ci) (rows ascending, columns ascending within a row), longdouble.
    Elements with 4 vertices are tets, with 3 triangles (z ignored)."""
    n = rp.size - 1
    npe = tets.shape[1]
    g, vol = tet_gradients(xyz, tets) if npe == 4 else tri_gradients(xyz, tets)
    kv = kappa.astype(LD) * vol
    m = LD(mass_coeff) * vol / LD(20 if npe == 4 else 12)
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(rp))
    keys = rows * n + ci.astype(np.int64)
    vals = np.zeros(ci.size, dtype=LD)
    t = tets.astype(np.int64)
    for a in range(npe):
        for b in range(npe):
            s = kv * np.einsum("ij,ij->i", g[:, a, :], g[:, b, :]) + m * (2 if a == b else 1)
            slot = np.searchsorted(keys, t[:, a] * n + t[:, b])
            if not np.array_equal(keys[slot], t[:, a] * n + t[:, b]):
                raise ValueError("element entry outside the pattern")
            np.add.at(vals, slot, s)
    return vals
