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

W- and F-cycles (cycleType 1, 2) differ only in the coarse correction:
    W: x_c = W(l + 1, b_c), then x_c = W(l + 1, b_c) from x_c
    F: x_c = F(l + 1, b_c), then x_c = V(l + 1, b_c) from x_c
(the coarsest level is solved once), where a cycle from a start x0 smooths
from x0 (Chebyshev step 0 from d = 0) instead of from zero.
"""
from __future__ import annotations

import copy

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

    def cycle(self, l, b, ctype=0, x0=None):
        """One cycle of type ctype (0 V, 1 W, 2 F) from zero or from x0."""
        if l == self.coarse:
            return self.Cinv @ b
        A, P = self.A[l], self.P[l]
        x = self.smooth(l, b, x0, self.nu_pre, x0 is None)
        bc = P.T @ (b - A @ x)
        xc = self.cycle(l + 1, bc, ctype)
        if ctype in (1, 2) and l + 1 != self.coarse:
            xc = self.cycle(l + 1, bc, 1 if ctype == 1 else 0, xc)
        x = x + P @ xc
        return self.smooth(l, b, x, self.nu_post, False)

    def rounded_to_fp32(self):
        """The same cycle with the level operators, prolongators and D^-1
        rounded to FP32 (FP64 arithmetic): what rounding the FP32 hierarchy's
        stored values alone changes."""
        h = copy.copy(self)
        h.A = [A.astype(np.float32).astype(np.float64) for A in self.A]
        h.P = [P.astype(np.float32).astype(np.float64) for P in self.P]
        h.dinv = [d.astype(np.float32).astype(np.float64) for d in self.dinv]
        return h
