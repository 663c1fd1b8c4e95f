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
          matvec=None, x0=None, iterations_done: int = 0) -> GmresResult:
    """Solve A x = b for a column-major complex64 GPU matrix A (n x n, a
    transposed view as nbem.colmajor makes) and a GPU vector b; x is
    complex128.  Stops when the relative residual estimate drops below tol,
    after maxiter iterations (default: n; iterations_done of them already
    spent by the caller), or when the Krylov space is exhausted.  matvec(v)
    replaces the product with A (A then only gives the size and device); x0
    is a starting guess (default 0)."""
    import torch
    n = A.shape[0]
    if A.shape != (n, n) or b.shape != (n,):
        raise ValueError("A must be n x n and b of length n")
    dev = A.device
    b = b.to(torch.complex128)
    bnorm = torch.linalg.vector_norm(b).item()
    x = torch.zeros(n, dtype=torch.complex128, device=dev) if x0 is None else x0.to(torch.complex128).clone()
    if bnorm == 0.0:
        return GmresResult(torch.zeros_like(x), 0, 0.0, True)
    maxiter = n if maxiter is None else int(maxiter)
    if matvec is None:
        work = nbem.gemv_workspace(A)

        def matvec(v):
            return nbem.gemv(A, v, work=work, stream=stream)
    it = int(iterations_done)
    r = b.clone() if x0 is None else b - matvec(x)
    beta = torch.linalg.vector_norm(r).item()
    if beta / bnorm < tol or it >= maxiter:
        return GmresResult(x, it, beta / bnorm, beta / bnorm < tol)
    m = max(1, min(int(restart), n, maxiter - it))
    V = torch.empty((m + 1, n), dtype=torch.complex128, device=dev)
    H = np.zeros((m + 1, m), dtype=np.complex128)
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
            cs[j], sn[j] = _givens_column(col, cs, sn, j)
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


def _givens_column(col, cs, sn, j):
    """Apply the earlier rotations to the new Hessenberg column col (length
    j + 2), then a new one that zeroes col[j + 1]; returns the rotation."""
    for t in range(j):
        a0 = np.conj(cs[t]) * col[t] + np.conj(sn[t]) * col[t + 1]
        col[t + 1] = -sn[t] * col[t] + cs[t] * col[t + 1]
        col[t] = a0
    den = np.hypot(abs(col[j]), abs(col[j + 1]))
    c, s_ = (1.0, 0.0) if den == 0.0 else (col[j] / den, col[j + 1] / den)
    col[j] = den
    col[j + 1] = 0.0
    return c, s_


def gmres_batched(As, bs, tol: float = 1e-8, restart: int = 3000, maxiter: int | None = None, stream=None,
                  matvecs=None):
    """Solve A_g x_g = b_g for a batch of systems of one size (the requests
    of one mesh) in lockstep: each Arnoldi step launches every active
    system's product, then orthogonalises all of them with batched products
    and brings the new Hessenberg columns to the host in one transfer -- one
    host wait per step for the batch instead of one per system.  Every
    system's iteration is the one gmres() runs (same rotations, same
    stopping rule); systems stop as they converge.  A system still short of
    tol after a full restart cycle goes on alone in gmres() from its iterate.
    As: list of n x n matrices (column-major complex64 GPU views), bs: (B, n).
    matvecs: per system, a callable replacing the product (as in gmres)."""
    import torch
    B = len(As)
    n = As[0].shape[0]
    if bs.shape != (B, n) or any(A.shape != (n, n) for A in As):
        raise ValueError("every A must be n x n and bs of shape (len(As), n)")
    dev = As[0].device
    if matvecs is None:
        matvecs = []
        for A in As:
            work = nbem.gemv_workspace(A)
            matvecs.append(lambda v, A=A, work=work: nbem.gemv(A, v, work=work, stream=stream))
    b = bs.to(torch.complex128)
    bnorm = torch.linalg.vector_norm(b, dim=1).cpu().numpy()
    maxiter = n if maxiter is None else int(maxiter)
    m = max(1, min(int(restart), n, maxiter))
    V = torch.empty((B, m + 1, n), dtype=torch.complex128, device=dev)
    scale = torch.from_numpy(np.where(bnorm > 0, 1.0 / np.where(bnorm > 0, bnorm, 1.0), 0.0)).to(dev)
    V[:, 0] = b * scale[:, None]
    H = np.zeros((B, m + 1, m), dtype=np.complex128)
    g = np.zeros((B, m + 1), dtype=np.complex128)
    g[:, 0] = bnorm
    cs = np.zeros((B, m), dtype=np.complex128)
    sn = np.zeros((B, m), dtype=np.complex128)
    j_done = np.zeros(B, dtype=np.int64)
    res = np.where(bnorm > 0, 1.0, 0.0)
    active = bnorm > 0
    W = torch.empty((B, n), dtype=torch.complex128, device=dev)
    for j in range(m):
        idx = np.nonzero(active)[0]
        if len(idx) == 0:
            break
        for q in idx:
            W[q] = matvecs[q](V[q, j])
        ti = torch.from_numpy(idx).to(dev)
        Vj = V[ti, :j + 1]                         # (B', j + 1, n), gathered
        w = W[ti]
        # classical Gram-Schmidt, twice, as batched products
        h = torch.bmm(Vj.conj(), w.unsqueeze(2)).squeeze(2)
        w = w - torch.bmm(h.unsqueeze(1), Vj).squeeze(1)
        h2 = torch.bmm(Vj.conj(), w.unsqueeze(2)).squeeze(2)
        w = w - torch.bmm(h2.unsqueeze(1), Vj).squeeze(1)
        h = h + h2
        hn = torch.linalg.vector_norm(w, dim=1)
        cols = torch.cat([h, hn.unsqueeze(1).to(torch.complex128)], dim=1).cpu().numpy()
        hnv = cols[:, j + 1].real.copy()
        inv = torch.from_numpy(np.where(hnv != 0.0, 1.0 / np.where(hnv != 0.0, hnv, 1.0), 0.0)).to(dev)
        V[ti, j + 1] = w * inv[:, None]
        for r_, q in enumerate(idx):
            col = cols[r_]
            cs[q, j], sn[q, j] = _givens_column(col, cs[q], sn[q], j)
            H[q, :j + 2, j] = col
            g[q, j + 1] = -sn[q, j] * g[q, j]
            g[q, j] = np.conj(cs[q, j]) * g[q, j]
            j_done[q] = j + 1
            res[q] = abs(g[q, j + 1]) / bnorm[q]
            if res[q] < tol or j + 1 >= maxiter or hnv[r_] == 0.0:
                active[q] = False
    out = []
    for q in range(B):
        jd = int(j_done[q])
        x = torch.zeros(n, dtype=torch.complex128, device=dev)
        if jd:
            y = np.linalg.solve(np.triu(H[q, :jd, :jd]), g[q, :jd])
            x = torch.from_numpy(y).to(dev).unsqueeze(0).mm(V[q, :jd]).squeeze(0)
        if bnorm[q] == 0.0 or res[q] < tol or jd >= maxiter:
            out.append(GmresResult(x, jd, float(res[q]), bool(res[q] < tol)))
        else:
            # a full cycle without convergence: restart this system alone
            out.append(gmres(As[q], bs[q], tol=tol, restart=restart, maxiter=maxiter, stream=stream,
                             matvec=matvecs[q], x0=x, iterations_done=jd))
    return out
