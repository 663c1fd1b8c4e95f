"""Solvers of the assembled systems on the device.

gmres: restarted GMRES without preconditioner, the solver of NiHu's
Burton-Miller application (nihu/applications/helmholtz/helmholtz_3d_coll.cpp:99-108:
Eigen::GMRES with the identity preconditioner, tolerance 1e-8 on the relative
residual, restart 3000).  The matrix stays complex64 as assembled; every
product with it is the engine's FP64-accumulating matrix-vector kernel
(nb_gemv_c64: complex64 entries, complex128 vectors), and the Arnoldi basis,
the Gram-Schmidt sums (classical, applied twice) and the Givens rotations are
complex128, so the iteration converges to the solution of the stored system
to the requested tolerance instead of stalling at complex64 rounding.

Each iteration reads the matrix once (8 bytes per entry): for the
20480-unknown systems of the serving workload the tens of iterations the
Burton-Miller operator needs cost far less than a dense LU factorisation.
"""

from __future__ import annotations

import numpy as np

from . import nbem


class GmresResult:
    def __init__(self, x, iterations, residual, converged):
        self.x = x
        self.iterations = iterations
        self.residual = residual        # estimated relative residual |b - A x| / |b|
        self.converged = converged


def gmres(A, b, tol: float = 1e-8, restart: int = 3000, maxiter: int | None = None, stream=None,
          matvec=None) -> GmresResult:
    """Solve A x = b for a column-major complex64 GPU matrix A (n x n, a
    transposed view as nbem.colmajor makes) and a GPU vector b; x is
    complex128.  Stops when the relative residual estimate drops below tol,
    after maxiter iterations (default: n), or when the Krylov space is
    exhausted.  matvec(v) replaces the product with A (A then only gives the
    size and device)."""
    import torch
    n = A.shape[0]
    if A.shape != (n, n) or b.shape != (n,):
        raise ValueError("A must be n x n and b of length n")
    dev = A.device
    b = b.to(torch.complex128)
    bnorm = torch.linalg.vector_norm(b).item()
    x = torch.zeros(n, dtype=torch.complex128, device=dev)
    if bnorm == 0.0:
        return GmresResult(x, 0, 0.0, True)
    maxiter = n if maxiter is None else int(maxiter)
    m = max(1, min(int(restart), n, maxiter))
    if matvec is None:
        work = nbem.gemv_workspace(A)

        def matvec(v):
            return nbem.gemv(A, v, work=work, stream=stream)
    V = torch.empty((m + 1, n), dtype=torch.complex128, device=dev)
    H = np.zeros((m + 1, m), dtype=np.complex128)
    it = 0
    r = b.clone()
    beta = bnorm
    while True:
        V[0] = r / beta
        g = np.zeros(m + 1, dtype=np.complex128)
        g[0] = beta
        cs = np.zeros(m, dtype=np.complex128)
        sn = np.zeros(m, dtype=np.complex128)
        j_done = 0
        res = beta / bnorm
        for j in range(m):
            w = matvec(V[j])
            # classical Gram-Schmidt, twice
            h = V[:j + 1].conj() @ w
            w = w - V[:j + 1].T @ h
            h2 = V[:j + 1].conj() @ w
            w = w - V[:j + 1].T @ h2
            h = h + h2
            hn = torch.linalg.vector_norm(w)
            # one transfer (and host wait) per iteration: the new column of H
            col = torch.cat([h, hn.reshape(1).to(torch.complex128)]).cpu().numpy()
            hnv = col[j + 1].real
            if hnv != 0.0:
                V[j + 1] = w / hnv
            # the earlier rotations, then a new one that zeroes col[j + 1]
            for t in range(j):
                a0 = np.conj(cs[t]) * col[t] + np.conj(sn[t]) * col[t + 1]
                col[t + 1] = -sn[t] * col[t] + cs[t] * col[t + 1]
                col[t] = a0
            den = np.hypot(abs(col[j]), abs(col[j + 1]))
            if den == 0.0:
                cs[j], sn[j] = 1.0, 0.0
            else:
                cs[j], sn[j] = col[j] / den, col[j + 1] / den
            col[j] = den
            col[j + 1] = 0.0
            H[:j + 2, j] = col
            g[j + 1] = -sn[j] * g[j]
            g[j] = np.conj(cs[j]) * g[j]
            it += 1
            j_done = j + 1
            res = abs(g[j + 1]) / bnorm
            if res < tol or it >= maxiter or hnv == 0.0:
                break
        # x += V y with H y = g (upper triangular)
        y = np.linalg.solve(np.triu(H[:j_done, :j_done]), g[:j_done]) if j_done else np.zeros(0)
        x = x + V[:j_done].T @ torch.from_numpy(y).to(dev)
        if res < tol or it >= maxiter:
            break
        r = b - matvec(x)
        beta = torch.linalg.vector_norm(r).item()
        if beta / bnorm < tol:
            res = beta / bnorm
            break
    return GmresResult(x, it, res, res < tol)
