"""Host (numpy / scipy, FP64) statement of the library's symmetric V-cycle,
built from the hierarchy the library exports (level operators, prolongators,
smoother data, internal numbering).  It is the checker of the device
smoother, residual, restriction, coarse-solve and prolongation kernels; it
does not run any library kernel.

V-cycle on level l for right-hand side b:
    x = S_pre(b, x = 0)              (nu_pre polynomial steps on D^-1 A)
    r = b - A x,  b_c = P^T r
    x_c = V-cycle(l + 1, b_c)        (coarsest level: x_c = A_c^-1 b_c)
    x = x + P x_c
    x = S_post(b, x)                 (nu_post = postInnerIters * postRelaxes)
Chebyshev on [lmin, lmax] with theta = (lmax + lmin) / 2, delta = (lmax - lmin) / 2,
sigma = theta / delta, rho_0 = 1 / sigma; step 0: d = D^-1 (b - A x) / theta
(from x = 0: x = d = D^-1 b / theta); step s > 0: rho_s = 1 / (2 sigma - rho_{s-1}),
d = rho_s rho_{s-1} d + (2 rho_s / delta) D^-1 (b - A x); every step x = x + d.
Jacobi: x = x + omega D^-1 (b - A x) (from x = 0: x = omega D^-1 b).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp


def cheb_coefficients(lmin, lmax, nu):
    """(c1, c2) of steps 0 .. nu-1 (step 0: c1 = 0, c2 = 1 / theta)."""
    theta = 0.5 * (lmax + lmin)
    delta = 0.5 * (lmax - lmin)
    sigma = theta / delta
    rho_old = 1.0 / sigma
    out = []
    for st in range(nu):
        if st == 0:
            out.append((0.0, 1.0 / theta))
        else:
            rho_new = 1.0 / (2.0 * sigma - rho_old)
            out.append((rho_new * rho_old, 2.0 * rho_new / delta))
            rho_old = rho_new
    return out


class HostVCycle:
    def __init__(self, solver):
        p = solver._p
        self.stype = int(p.smootherType)
        self.nu_pre = int(p.preInnerIters)
        self.nu_post = int(p.postInnerIters) * max(int(p.postRelaxes), 1)
        nl = solver.num_levels()
        self.A, self.P, self.dinv, self.sm = [], [], [], []
        for l in range(nl):
            rp, ci, v, _ = solver.level_matrix(l)
            n = rp.size - 1
            A = sp.csr_matrix((v, ci, rp), shape=(n, n))
            self.A.append(A)
            self.dinv.append(1.0 / A.diagonal())
        for l in range(nl - 1):
            prp, pci, pv = solver.level_prolongator(l)
            self.P.append(sp.csr_matrix((pv, pci, prp), shape=(self.A[l].shape[0], self.A[l + 1].shape[0])))
            self.sm.append(solver.level_smoother(l))
        self.coarse = nl - 1
        self.Cinv = np.linalg.inv(self.A[-1].toarray())

    def smooth(self, l, b, x, nu, zero):
        A, dinv, S = self.A[l], self.dinv[l], self.sm[l]
        if nu <= 0:
            return np.zeros_like(b) if zero else x
        if self.stype == 0:
            w = S["omega"]
            start = 0
            if zero:
                x = w * dinv * b
                start = 1
            for _ in range(start, nu):
                x = x + w * dinv * (b - A @ x)
            return x
        coef = cheb_coefficients(S["lmin"], S["lmax"], nu)
        start = 0
        if zero:
            x = coef[0][1] * dinv * b
            d = x.copy()
            start = 1
        else:
            d = np.zeros_like(b)
        for st in range(start, nu):
            c1, c2 = coef[st]
            d = c1 * d + c2 * dinv * (b - A @ x)
            x = x + d
        return x

    def vcycle(self, l, b):
        if l == self.coarse:
            return self.Cinv @ b
        A, P = self.A[l], self.P[l]
        x = self.smooth(l, b, None, self.nu_pre, True)
        r = b - A @ x
        xc = self.vcycle(l + 1, P.T @ r)
        x = x + P @ xc
        return self.smooth(l, b, x, self.nu_post, False)
