"""CPU thick-restart Lanczos around any y = A x — TEST INFRASTRUCTURE ONLY.

Used by bench.py's cpu_baseline leg (full CPU solves of the small and medium sectors with the
oracle's row-gather matvec, a restatement of MomentumED/src/method/linear_map.jl:100-118) and
by tests/ as a checker. The product never imports it.

The reference's map solve is KrylovKit.eigsolve(A, vec0, N, :SR; ishermitian=true,
krylovdim, tol, maxiter) (linear_map.jl:179-191). KrylovKit 0.10 (not vendored) runs a
Lanczos iteration with full re-orthogonalisation and thick restarts that keep Ritz vectors,
stopping when the N lowest Ritz pairs have residual norm |beta * y_m| <= tol. This is that
iteration in numpy (classical Gram-Schmidt applied twice per step, Rayleigh-Ritz by
numpy.linalg.eigh, about 3/5 of the basis kept per restart), so a CPU solve costs what the
reference's CPU solve costs in matvecs and vector work. Pinned against LAPACK on the small
sectors by tests/test_oracle.py.
"""
from __future__ import annotations

import numpy as np


def eigsolve(matvec, D: int, nev: int, krylovdim: int = 30, tol: float = 1e-12,
             maxiter: int = 100, v0=None, seed: int = 0):
    """Lowest `nev` eigenpairs of the Hermitian map x -> matvec(x) on C^D.

    Returns (vals ascending, vecs [nev, D], info dict with iterations, matvecs, converged,
    max_residual)."""
    nev = min(nev, D)
    m = min(max(krylovdim, nev + 2), D)
    V = np.empty((m + 1, D), dtype=np.complex128)
    rng = np.random.default_rng(seed)
    v = (np.asarray(v0, dtype=np.complex128).copy() if v0 is not None
         else rng.random(D) + 1j * rng.random(D))
    V[0] = v / np.linalg.norm(v)
    T = np.zeros((m, m))
    k, matvecs, nconv, beta_last = 0, 0, 0, 0.0
    hnorm = 0.0
    theta = Y = res = None
    it = 0
    for it in range(1, maxiter + 1):
        for j in range(k, m):
            w = matvec(V[j])
            matvecs += 1
            Vj = V[: j + 1]
            h = Vj.conj() @ w
            w = w - Vj.T @ h
            h2 = Vj.conj() @ w
            w = w - Vj.T @ h2
            h = h + h2
            T[j, j] = h[j].real
            beta = float(np.linalg.norm(w))
            hnorm = max(hnorm, abs(h[j].real) + beta + (abs(T[j, j - 1]) if j else 0.0))
            if not beta > 1e-14 * max(hnorm, 1e-300):
                # invariant subspace: continue from a random direction orthogonal to it
                for _ in range(4):
                    w = rng.random(D) + 1j * rng.random(D)
                    for _p in range(2):
                        w = w - Vj.T @ (Vj.conj() @ w)
                    if np.linalg.norm(w) > 1e-8:
                        break
                V[j + 1] = w / np.linalg.norm(w)
                beta = 0.0
            else:
                V[j + 1] = w / beta
            if j + 1 < m:
                T[j, j + 1] = T[j + 1, j] = beta
            beta_last = beta
        theta, Y = np.linalg.eigh(T)
        res = np.abs(beta_last * Y[m - 1, :])
        nconv = 0
        while nconv < nev and res[nconv] <= tol:
            nconv += 1
        if nconv >= nev or it == maxiter or m == D:
            break
        keep = min(max((3 * m + 2 * nconv) // 5, nev), m - 1)
        V[:keep] = Y[:, :keep].T @ V[:m]
        V[keep] = V[m]
        T[:] = 0.0
        for c in range(keep):
            T[c, c] = theta[c]
            T[c, keep] = T[keep, c] = beta_last * Y[m - 1, c]
        k = keep
    vecs = Y[:, :nev].T @ V[:m]
    info = {"iterations": it, "matvecs": matvecs, "converged": nconv,
            "max_residual": float(res[:nev].max())}
    return theta[:nev].copy(), vecs, info
