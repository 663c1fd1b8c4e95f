"""Host logic of the GMRES solver (nihu_amd/solve.py) with a CPU matrix-vector
product standing in for the engine's kernel: it reaches the requested
relative residual, restarts correctly, and matches a direct solve; on the
Burton-Miller system of the oracle's matrices it converges in the number of
iterations of a reference GMRES (scipy) to the same tolerance."""

import numpy as np
import pytest
import torch

from nihu_amd.solve import gmres


def _run(A, b, **kw):
    At = torch.from_numpy(A)
    return gmres(At, torch.from_numpy(b), matvec=lambda v: At @ v, **kw)


def test_gmres_random_system():
    rng = np.random.default_rng(0)
    n = 200
    A = np.eye(n) * 4 + rng.normal(size=(n, n)) / np.sqrt(n) + 1j * rng.normal(size=(n, n)) / np.sqrt(n)
    b = rng.normal(size=n) + 1j * rng.normal(size=n)
    xe = np.linalg.solve(A, b)
    for restart in (200, 7):
        r = _run(A, b, tol=1e-10, restart=restart)
        x = r.x.numpy()
        print("restart", restart, "iterations", r.iterations, "residual", r.residual)
        assert r.converged
        assert np.linalg.norm(A @ x - b) / np.linalg.norm(b) < 1e-9
        assert np.linalg.norm(x - xe) / np.linalg.norm(xe) < 1e-8


def test_gmres_zero_rhs_and_iteration_cap():
    A = np.eye(5, dtype=np.complex128)
    r = _run(A, np.zeros(5, dtype=np.complex128))
    assert r.converged and r.iterations == 0 and not r.x.abs().any()
    rng = np.random.default_rng(1)
    n = 50
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    r = _run(A, rng.normal(size=n) + 0j, tol=1e-14, maxiter=5)
    assert r.iterations == 5 and not r.converged


def test_gmres_burton_miller_matches_scipy(orc):
    from scipy.sparse.linalg import gmres as sp_gmres
    from nihu_amd.mesh import jittered_icosphere
    from nihu_amd import serve
    nodes, tris = jittered_icosphere(3, seed=0)
    k = 10.0
    A, B = orc.assemble_bm(nodes, tris, k, pol=orc.policy(rule=1, fixed_degree=5, near_mode=1, near_limit=2.0))
    c, nu = serve._centres_normals(nodes, tris)
    d = np.array([0.0, 0.6, 0.8])
    b = B @ ((1j * k) * (nu @ d) * np.exp(-1j * k * (c @ d)))
    r = _run(A, b, tol=1e-8)
    its = [0]
    sp_gmres(A, b, rtol=1e-8, restart=3000, maxiter=1, callback=lambda _: its.__setitem__(0, its[0] + 1),
             callback_type="pr_norm")
    xe = np.linalg.solve(A, b)
    err = np.linalg.norm(r.x.numpy() - xe) / np.linalg.norm(xe)
    print("iterations %d (scipy %d), error %.2e" % (r.iterations, its[0], err))
    assert r.converged and abs(r.iterations - its[0]) <= 1 and err < 1e-7


def test_gmres_batched_matches_single():
    """The lockstep batch: every system's solution and iteration count equal
    gmres() alone on it (to rounding: the batch orthogonalises with batched
    products); a zero right-hand side, a system that needs a restart (short
    restart length) and one stopped by maxiter inside the same batch."""
    from nihu_amd.solve import gmres_batched
    rng = np.random.default_rng(3)
    n = 120
    mats, rhs = [], []
    for q in range(5):
        A = np.eye(n) * (3 + q) + (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))) / np.sqrt(n)
        mats.append(torch.from_numpy(A))
        rhs.append(np.zeros(n, dtype=np.complex128) if q == 2 else rng.normal(size=n) + 1j * rng.normal(size=n))
    bs = torch.from_numpy(np.stack(rhs))
    mv = [lambda v, A=A: A @ v for A in mats]
    for restart, maxiter in ((200, None), (6, None), (200, 4)):
        out = gmres_batched(mats, bs, tol=1e-11, restart=restart, maxiter=maxiter, matvecs=mv)
        for q in range(5):
            one = gmres(mats[q], bs[q], tol=1e-11, restart=restart, maxiter=maxiter, matvec=mv[q])
            assert out[q].converged == one.converged and out[q].iterations == one.iterations, (restart, q)
            d = (out[q].x - one.x).abs().max().item()
            assert d <= 1e-12 * max(1.0, one.x.abs().max().item()), (restart, q, d)
            if maxiter is None:
                A = mats[q].numpy()
                x = out[q].x.numpy()
                assert np.linalg.norm(A @ x - rhs[q]) <= 1e-10 * max(np.linalg.norm(rhs[q]), 1e-300) or q == 2
        assert not out[2].x.abs().any() and out[2].converged
