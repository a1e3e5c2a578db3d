"""Writes tests/golden/c4_gradient.npz: the FP64 CPU reference gradient of the
C4 adjoint workload, at 4096 seeded elements, with its 2-norm and the loss.

C4: 49^3 jittered Kuhn tet cube (117,649 vertices, 663,552 tets), kappa =
exp(g), g ~ N(0, 1) (seed 0), K = kappa-stiffness + 1e-2 mass, b ~ N(0, 1)
(seed 0); u_obs = K(kappa2)^-1 b + N(0, 1e-3) noise (kappa2 seed 101, noise
seed 7); L = 0.5 |u - u_obs|^2, dL/dkappa_e = -|T_e| grad(lambda).grad(u)
with K lambda = u - u_obs.  Every solve is a SuperLU factorisation of the
oracle's matrix (FP64); the element gradient is the oracle's.  Runs in this
repository's CPU environment (about 7 minutes: two factorisations):

    python tests/golden/make_c4_gradient.py
"""
from __future__ import annotations

import sys
import time
from pathlib import Path

import numpy as np
import scipy.sparse.linalg as spla

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
from conftest import load_oracle, load_package  # noqa: E402

OUT = HERE / "c4_gradient.npz"
N_SAMPLE = 4096


def main():
    orc = load_oracle()
    mg = load_package().meshgen
    c = mg.CONFIGS["c4_cube49_adjoint"]
    xyz, tets = mg.cube_mesh(c["N"], seed=0)
    n, ne = xyz.shape[0], tets.shape[0]
    kappa = mg.lognormal_kappa(0, ne, c["sigma"])
    kappa2 = mg.lognormal_kappa(101, ne, c["sigma"])
    b = mg.gaussian_rhs(0, n)
    t = time.time()
    A2 = orc.scipy_matrix(xyz, tets, kappa2, c["mass"]).tocsc()
    u_obs = spla.splu(A2).solve(b) + mg.noise(7, n, 1e-3)
    del A2
    print(f"u_obs: {time.time() - t:.0f} s", flush=True)
    t = time.time()
    A = orc.scipy_matrix(xyz, tets, kappa, c["mass"]).tocsc()
    lu = spla.splu(A)
    u = lu.solve(b)
    lam = lu.solve(u - u_obs)
    print(f"u, lambda: {time.time() - t:.0f} s", flush=True)
    g = orc.element_grad(xyz, tets, lam, u)
    rng = np.random.default_rng(4049)
    idx = np.sort(rng.choice(ne, size=N_SAMPLE, replace=False)).astype(np.int64)
    res_u = np.linalg.norm(b - A @ u) / np.linalg.norm(b)
    res_l = np.linalg.norm((u - u_obs) - A @ lam) / np.linalg.norm(u - u_obs)
    np.savez(OUT, idx=idx, g=g[idx], g_norm=np.array(np.linalg.norm(g)),
             loss=np.array(0.5 * float(np.dot(u - u_obs, u - u_obs))),
             u_norm=np.array(np.linalg.norm(u)), lam_norm=np.array(np.linalg.norm(lam)),
             residuals=np.array([res_u, res_l]))
    print(f"wrote {OUT}: |g| = {np.linalg.norm(g):.6e}, residuals {res_u:.1e} {res_l:.1e}")


if __name__ == "__main__":
    main()
