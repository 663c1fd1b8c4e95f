"""Complex wavenumbers k + i ki (NiHu's Helmholtz kernel with a complex
WaveNumber) on the device, against the FP64 oracle on the same inputs:
the four Helmholtz operators and the Burton-Miller matrices, under NiHu's
degrees and the fixed-rule configuration, in mixed precision (FP32 regular
pairs with the attenuation e^{ki r} on the weights: relative Frobenius
<= 1e-5) and under the FP64 precision policy (<= 1e-11); a zero imaginary
part gives the real-wavenumber matrices bit for bit; the Galerkin
formalism (constant and isoparametric linear fields) against the oracle's
Galerkin matrices (<= 1e-5); quadrilateral meshes refuse complex
wavenumbers."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def nb():
    import torch
    assert torch.cuda.is_available(), "GPU test without a GPU"
    from nihu_amd import nbem
    return nbem


def relfro(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("precision,tol", [("mixed", 1e-5), ("fp64", 1e-11)])
@pytest.mark.parametrize("nx,ny", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_operators(nb, orc, precision, tol, nx, ny):
    import torch
    from nihu_amd.mesh import jittered_icosphere
    nodes, tris = jittered_icosphere(2, seed=5)
    n = len(tris)
    m = nb.Mesh(nodes, tris)
    k = complex(4.0, -0.6)
    alpha = 0.4 + 0.3j
    out = nb.colmajor(n, n, "complex128")
    nb.assemble(m, m, out, nb.NB_HELMHOLTZ, nx, ny, k, alpha=alpha, pol=nb.policy(precision=precision),
                accumulate=False)
    torch.cuda.synchronize()
    ref = orc.assemble(nodes, tris, 1, nx, ny, k=k, alpha=alpha)
    e = relfro(out.cpu().numpy(), ref)
    print("(%d, %d) %s: rel fro %.2e" % (nx, ny, precision, e), flush=True)
    assert e <= tol


@pytest.mark.parametrize("precision,tol", [("mixed", 1e-5), ("fp64", 1e-11)])
@pytest.mark.parametrize("which", ["config", "nihu"])
def test_burton_miller(nb, orc, precision, tol, which):
    import torch
    from nihu_amd.mesh import jittered_icosphere
    nodes, tris = jittered_icosphere(3, seed=2)
    n = len(tris)
    m = nb.Mesh(nodes, tris)
    k = complex(10.0, -1.5)
    if which == "config":
        pol = nb.policy(**nb.CONFIG_POLICY, precision=precision)
        opol = orc.policy(rule=1, fixed_degree=5, near_mode=1, near_limit=2.0)
    else:
        pol, opol = nb.policy("nihu", precision=precision), orc.policy()
    dt = "complex64" if precision == "mixed" else "complex128"
    A, B = nb.colmajor(n, n, dt), nb.colmajor(n, n, dt)
    nb.assemble_bm(m, k, A, B, pol=pol)
    torch.cuda.synchronize()
    Ar, Br = orc.assemble_bm(nodes, tris, k, pol=opol)
    ea, eb = relfro(A.cpu().numpy(), Ar), relfro(B.cpu().numpy(), Br)
    print("Burton-Miller k = %s (%s, %s): A %.2e B %.2e" % (k, which, precision, ea, eb), flush=True)
    assert ea <= tol and eb <= tol


def test_zero_imaginary_part_is_the_real_path(nb):
    import torch
    from nihu_amd.mesh import jittered_icosphere
    nodes, tris = jittered_icosphere(3, seed=3)
    n = len(tris)
    m = nb.Mesh(nodes, tris)
    for pol in (nb.policy(**nb.CONFIG_POLICY), nb.policy("nihu")):
        A, B = nb.colmajor(n, n), nb.colmajor(n, n)
        A2, B2 = nb.colmajor(n, n), nb.colmajor(n, n)
        nb.assemble_bm(m, 7.5, A, B, pol=pol)
        nb.assemble_bm(m, complex(7.5, 0.0), A2, B2, alpha=complex(0.0, -1.0 / 7.5), pol=pol)
        torch.cuda.synchronize()
        assert torch.equal(torch.view_as_real(A), torch.view_as_real(A2))
        assert torch.equal(torch.view_as_real(B), torch.view_as_real(B2))


@pytest.mark.parametrize("formalism", ["galerkin", "galerkin_linear"])
@pytest.mark.parametrize("nx,ny", [(0, 0), (0, 1), (1, 0)])
def test_galerkin(nb, orc, formalism, nx, ny):
    import torch
    from nihu_amd.mesh import jittered_icosphere
    nodes, tris = jittered_icosphere(2, seed=8)
    m = nb.Mesh(nodes, tris)
    k = complex(3.0, -0.7)
    linear = formalism == "galerkin_linear"
    n = m.n_nodes if linear else m.n
    out = nb.colmajor(n, n, "complex128")
    nb.assemble(m, m, out, nb.NB_HELMHOLTZ, nx, ny, k, alpha=0.8 - 0.1j, accumulate=False, formalism=formalism)
    torch.cuda.synchronize()
    ref = orc.assemble_galerkin(nodes, tris, 1, nx, ny, k, alpha=0.8 - 0.1j, linear=linear)
    e = relfro(out.cpu().numpy(), ref)
    print("Galerkin %s (%d, %d): rel fro %.2e" % (formalism, nx, ny, e), flush=True)
    assert e <= 1e-5


def test_refusals(nb):
    from nihu_amd.mesh import jittered_icosphere, sphere_boundary
    nodes, tris = jittered_icosphere(1, seed=3)
    m = nb.Mesh(nodes, tris)
    k = complex(2.0, -0.5)
    qn, qe = sphere_boundary(1.0, 2)
    q = nb.Mesh(qn, qe)
    with pytest.raises(NotImplementedError):
        nb.assemble_bm(q, k, nb.colmajor(q.n, q.n), nb.colmajor(q.n, q.n), pol=nb.policy(rule="fixed", degree=5))
    with pytest.raises(ValueError):
        nb.assemble_bm(m, complex(-1.0, 0.5), nb.colmajor(m.n, m.n), nb.colmajor(m.n, m.n))
