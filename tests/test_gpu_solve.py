"""Solved surface field and the Laplace configuration, through the C ABI.

* The Burton-Miller system A p = B q of plane-wave scattering from a rigid
  jittered sphere, assembled by the engine (complex64, FP32 far pairs) and by
  the FP64 oracle, both solved in complex128: the solved fields agree to 1e-6.
* The Laplace configuration (1280 triangles, SLP and DLP as real float64
  matrices, 7-point rule for every regular pair, no near-singular pass,
  analytic diagonal): relative Frobenius error <= 1e-5, FP64 diagonal
  entries within 1e-12.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def nb():
    import torch
    assert torch.cuda.is_available(), "GPU test without a GPU"
    from nihu_amd import nbem
    return nbem


def test_surface_field_vs_oracle(nb, orc):
    import torch
    from nihu_amd.mesh import jittered_icosphere
    nodes, tris = jittered_icosphere(3, seed=2)
    n = len(tris)
    k = 5.0
    gp = nb.policy(**nb.CONFIG_POLICY)
    op = orc.policy(rule=1, fixed_degree=5, near_mode=1, near_limit=2.0)
    m = nb.Mesh(nodes, tris)
    A, B = nb.colmajor(n, n), nb.colmajor(n, n)
    nb.assemble_bm(m, k, A, B, pol=gp)
    torch.cuda.synchronize()
    Ag, Bg = A.cpu().numpy().astype(np.complex128), B.cpu().numpy().astype(np.complex128)
    Ar, Br = orc.assemble_bm(nodes, tris, k, pol=op)
    X = nodes[tris]
    c = X.mean(1)
    nv = np.cross(X[:, 1] - X[:, 0], X[:, 2] - X[:, 0])
    nu = nv / np.linalg.norm(nv, axis=1)[:, None]
    d = np.array([0.3, -0.4, 0.866])
    d /= np.linalg.norm(d)
    q = (1j * k) * (nu @ d) * np.exp(-1j * k * (c @ d))
    pg = np.linalg.solve(Ag, Bg @ q)
    pr = np.linalg.solve(Ar, Br @ q)
    rel = np.linalg.norm(pg - pr) / np.linalg.norm(pr)
    print("matrices rel fro A %.2e B %.2e; solved field rel %.2e max abs %.2e" % (
        np.linalg.norm(Ag - Ar) / np.linalg.norm(Ar), np.linalg.norm(Bg - Br) / np.linalg.norm(Br), rel,
        np.abs(pg - pr).max()), flush=True)
    assert rel <= 1e-6


@pytest.mark.parametrize("ny", [0, 1])
def test_laplace_configuration(nb, orc, ny):
    import torch
    from nihu_amd.mesh import jittered_icosphere
    nodes, tris = jittered_icosphere(3, seed=0)
    n = len(tris)
    m = nb.Mesh(nodes, tris)
    out = nb.colmajor(n, n, "float64")
    nb.assemble(m, m, out, nb.NB_LAPLACE, 0, ny, 0.0, pol=nb.policy(rule="fixed", degree=5, near="none"),
                accumulate=False)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    ref = orc.assemble(nodes, tris, 0, 0, ny, pol=orc.policy(rule=1, fixed_degree=5, near_mode=2)).real
    rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    dg = np.abs(np.diag(got) - np.diag(ref)).max() / max(np.abs(np.diag(ref)).max(), 1e-300)
    print("Laplace %s: rel fro %.2e, diagonal %.2e" % ("DLP" if ny else "SLP", rel, dg), flush=True)
    assert rel <= 1e-5
    if ny == 0:
        assert dg <= 1e-12
    else:
        assert np.all(np.diag(got) == 0.0)
