"""The FP64 precision policy (every pair in FP64, NiHu's plain method) against
the FP64 oracle on the same inputs, through the C ABI:

* the Laplace configuration of 1280 jittered triangles: SLP and DLP as real
  float64 matrices, 7-point rule for every regular pair, no near-singular
  pass, analytic diagonal -- relative Frobenius error <= 1e-12 (the mixed
  precision path reaches 2e-7 here);
* the four Helmholtz kernels (k = 4), Burton-Miller matrices under the
  configuration policy (k = 10) and under NiHu's distance-driven degrees
  (k = 5), complex128 outputs: <= 1e-11;
* a frequency sweep with the FP64 policy equals its single-wavenumber
  assemblies bit for bit; the Galerkin formalism refuses the policy;
* quadrilateral meshes (Gauss square rules in FP64 from per-element tables)
  and meshes with both element types: the four kernels, real and complex k,
  NiHu's degrees and a fixed degree, and the Burton-Miller pair: <= 1e-11.

The bound is set by the rounding of the quadrature points: the engine forms
r = (c_j - c_i) + (y_q - c_j) from tables, the oracle y_q - x directly."""

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


@pytest.mark.parametrize("ny", [0, 1])
def test_laplace_configuration_fp64(nb, orc, ny):
    import torch
    from nihu_amd.mesh import jittered_icosphere
    nodes, tris = jittered_icosphere(3, seed=0)
    n = len(tris)
    m = nb.Mesh(nodes, tris)
    out = nb.colmajor(n, n, "float64")
    out.fill_(float("nan"))
    pol = nb.policy(rule="fixed", degree=5, near="none", precision="fp64")
    nb.assemble(m, m, out, nb.NB_LAPLACE, 0, ny, 0.0, pol=pol, accumulate=False)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    assert np.all(np.isfinite(got))
    ref = orc.assemble(nodes, tris, 0, 0, ny, pol=orc.policy(rule=1, fixed_degree=5, near_mode=2)).real
    rel = relfro(got, ref)
    off = ~np.eye(n, dtype=bool)
    mx = np.abs(got[off] - ref[off]).max() / np.abs(ref[off]).max()
    print("Laplace %s FP64: rel fro %.2e, max %.2e" % ("DLP" if ny else "SLP", rel, mx), flush=True)
    assert rel <= 1e-12 and mx <= 1e-12


@pytest.mark.parametrize("nx,ny", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_helmholtz_kernels_fp64(nb, orc, nx, ny):
    import torch
    from nihu_amd.mesh import jittered_icosphere
    nodes, tris = jittered_icosphere(2, seed=4)
    n = len(tris)
    m = nb.Mesh(nodes, tris)
    out = nb.colmajor(n, n, "complex128")
    alpha = 0.7 - 0.2j
    nb.assemble(m, m, out, nb.NB_HELMHOLTZ, nx, ny, 4.0, alpha=alpha, pol=nb.policy(precision="fp64"),
                accumulate=False)
    torch.cuda.synchronize()
    ref = orc.assemble(nodes, tris, 1, nx, ny, k=4.0, alpha=alpha)
    rel = relfro(out.cpu().numpy(), ref)
    print("Helmholtz (%d, %d) FP64: rel fro %.2e" % (nx, ny, rel), flush=True)
    assert rel <= 1e-11


@pytest.mark.parametrize("which", ["config", "nihu"])
def test_burton_miller_fp64(nb, orc, which):
    import torch
    from nihu_amd.mesh import jittered_icosphere
    nodes, tris = jittered_icosphere(3, seed=2)
    n = len(tris)
    m = nb.Mesh(nodes, tris)
    if which == "config":
        k, pol = 10.0, nb.policy(**nb.CONFIG_POLICY, precision="fp64")
        opol = orc.policy(rule=1, fixed_degree=5, near_mode=1, near_limit=2.0)
    else:
        k, pol, opol = 5.0, nb.policy("nihu", precision="fp64"), orc.policy()
    A, B = nb.colmajor(n, n, "complex128"), nb.colmajor(n, n, "complex128")
    nb.assemble_bm(m, k, A, B, pol=pol)
    torch.cuda.synchronize()
    Ar, Br = orc.assemble_bm(nodes, tris, k, pol=opol)
    ea, eb = relfro(A.cpu().numpy(), Ar), relfro(B.cpu().numpy(), Br)
    print("Burton-Miller FP64 (%s policy, k %g): A %.2e B %.2e" % (which, k, ea, eb), flush=True)
    assert ea <= 1e-11 and eb <= 1e-11


def test_sweep_and_refusals_fp64(nb):
    import torch
    from nihu_amd.mesh import jittered_icosphere
    nodes, tris = jittered_icosphere(2, seed=1)
    n = len(tris)
    m = nb.Mesh(nodes, tris)
    pol = nb.policy(precision="fp64")
    ks = np.array([1.5, 6.0, 11.0])
    A = torch.zeros((len(ks), n, n), dtype=torch.complex64, device="cuda").transpose(1, 2)
    B = torch.zeros_like(A)
    nb.assemble_sweep(m, ks, A, B, pol=pol)
    for f, k in enumerate(ks):
        A1, B1 = nb.colmajor(n, n), nb.colmajor(n, n)
        nb.assemble_bm(m, float(k), A1, B1, pol=pol)
        torch.cuda.synchronize()
        assert torch.equal(torch.view_as_real(A[f]), torch.view_as_real(A1))
        assert torch.equal(torch.view_as_real(B[f]), torch.view_as_real(B1))
    with pytest.raises(NotImplementedError):
        nb.assemble(m, m, nb.colmajor(m.n_nodes, m.n_nodes, "complex128"), nb.NB_HELMHOLTZ, 0, 0, 2.0, pol=pol,
                    formalism="galerkin_linear")


QNIHU = (dict(rule="nihu", near="mean_edge", limit=1.5, precision="fp64"),
         dict(rule=0, fixed_degree=0, near_mode=0, near_limit=1.5))
QFIXED = (dict(rule="fixed", degree=5, near="mean_edge", limit=1.5, precision="fp64"),
          dict(rule=1, fixed_degree=5, near_mode=0, near_limit=1.5))


@pytest.mark.parametrize("fam,k", [(0, 0.0), (1, 3.0), (1, complex(3.0, -0.5))])
@pytest.mark.parametrize("nx,ny", [(0, 0), (0, 1), (1, 0), (1, 1)])
@pytest.mark.parametrize("which", ["nihu", "fixed"])
def test_quadrilateral_kernels_fp64(nb, orc, fam, k, nx, ny, which):
    import torch
    from nihu_amd.mesh import jittered_cubed_sphere
    gp, op = QNIHU if which == "nihu" else QFIXED
    nodes, quads = jittered_cubed_sphere(5, seed=4)
    n = len(quads)
    m = nb.Mesh(nodes, quads)
    out = nb.colmajor(n, n, "complex128")
    out.fill_(complex("nan+nanj"))
    nb.assemble(m, m, out, fam, nx, ny, k, alpha=0.7 + 0.2j, pol=nb.policy(**gp), accumulate=False)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    assert np.isfinite(got).all()
    ref = orc.assemble(nodes, quads, fam, nx, ny, k=k, alpha=0.7 + 0.2j, pol=orc.policy(**op))
    e = relfro(got, ref)
    print("quadrilaterals FP64 fam %d (%d, %d) k = %s %s: rel fro %.2e" % (fam, nx, ny, k, which, e), flush=True)
    assert e <= 1e-11


def test_quadrilateral_and_mixed_burton_miller_fp64(nb, orc):
    import torch
    from nihu_amd.mesh import jittered_cubed_sphere, mixed_cubed_sphere
    from tests.test_gpu_mixed import oracle_mixed
    nodes, quads = jittered_cubed_sphere(5, seed=4)
    n = len(quads)
    m = nb.Mesh(nodes, quads)
    A, B = nb.colmajor(n, n, "complex128"), nb.colmajor(n, n, "complex128")
    nb.assemble_bm(m, 4.0, A, B, pol=nb.policy(**QNIHU[0]))
    torch.cuda.synchronize()
    Ar, Br = orc.assemble_bm(nodes, quads, 4.0, pol=orc.policy(**QNIHU[1]))
    ea, eb = relfro(A.cpu().numpy(), Ar), relfro(B.cpu().numpy(), Br)
    print("quadrilateral Burton-Miller FP64: A %.2e B %.2e" % (ea, eb), flush=True)
    assert ea <= 1e-11 and eb <= 1e-11
    # both element types in one mesh
    mn, table = mixed_cubed_sphere(5, seed=3)
    mm = nb.Mesh.from_elements(mn, table)
    nm = len(table)
    for nx, ny in ((0, 0), (1, 1)):
        out = nb.colmajor(nm, nm, "complex128")
        out.fill_(complex("nan+nanj"))
        nb.assemble(mm, mm, out, nb.NB_HELMHOLTZ, nx, ny, 3.0, pol=nb.policy(**QNIHU[0]), accumulate=False)
        torch.cuda.synchronize()
        got = out.cpu().numpy()
        ref = oracle_mixed(orc, mn, table, 1, nx, ny, 3.0, QNIHU[1])
        e = relfro(got, ref)
        print("mixed mesh FP64 (%d, %d): rel fro %.2e" % (nx, ny, e), flush=True)
        assert np.isfinite(got).all() and e <= 1e-11
