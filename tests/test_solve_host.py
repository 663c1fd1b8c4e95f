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
