"""The quadrilateral gauss field (NiHu's quad_1_gauss_field: four dofs per
element at the 2 x 2 Gauss points, collocated there; the default build of
nihu/applications/helmholtz/helmholtz_3d_coll.cpp) through the C ABI, against
the FP64 oracle on the same inputs:

* every kernel, Laplace and Helmholtz (real and complex wavenumber), NiHu's
  policy and a fixed Gauss degree 5: rows to 2e-6 of their largest entry and
  1e-5 relative Frobenius; the entries of the FP64 passes (near-singular
  polar / Telles blocks and the diagonal blocks with Duffy and Guiggiani at
  the off-centre nodes) to 1e-9;
* the Burton-Miller pair and the identity over the dofs; field points (a
  constant test field on another mesh) against the gauss trial field;
* sampled rows of a mesh with many row tiles; the pulsating-sphere answer
  from the device matrices; the FP64 precision policy (<= 1e-11); host outputs
  with accumulation; a failed inverse
  mapping (forced from the host) as an error; the refusals.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROW_TOL = 2e-6
NIHU = (dict(rule="nihu", near="mean_edge", limit=1.5), dict(rule=0, fixed_degree=0, near_mode=0, near_limit=1.5))
FIXED = (dict(rule="fixed", degree=5, near="mean_edge", limit=1.5),
         dict(rule=1, fixed_degree=5, near_mode=0, near_limit=1.5))


@pytest.fixture(scope="module")
def nb():
    import torch
    assert torch.cuda.is_available(), "GPU test without a GPU"
    from nihu_amd import nbem
    return nbem


@pytest.fixture(scope="module")
def qmesh():
    from nihu_amd.mesh import jittered_cubed_sphere
    return jittered_cubed_sphere(5, seed=4)


def rowscaled(got, ref):
    scale = np.abs(ref).max(axis=1, keepdims=True)
    scale[scale == 0] = 1.0
    return (np.abs(got - ref) / scale).max()


def relfro(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def fp64_mask(nb, m, nx, ny, gp):
    """Entries of the FP64 passes: the element classes (near < 0, diagonal)
    expanded to the 4 x 4 dof blocks."""
    cls = nb.classify(m, m, nx, ny, nb.policy(**gp)).cpu().numpy()
    return np.kron(cls < 0, np.ones((4, 4), dtype=bool))


@pytest.mark.parametrize("fam,k", [(0, 0.0), (1, 3.0), (1, complex(2.5, -0.5))])
@pytest.mark.parametrize("nxy", [(0, 0), (0, 1), (1, 0), (1, 1)])
@pytest.mark.parametrize("pol", ["nihu", "fixed"])
def test_gauss_operators(nb, orc, qmesh, fam, k, nxy, pol):
    import torch
    gp, op = NIHU if pol == "nihu" else FIXED
    nodes, quads = qmesh
    nx, ny = nxy
    print("gauss operator", fam, k, nxy, pol, flush=True)
    m = nb.Mesh(nodes, quads)
    nd = 4 * m.n
    out = nb.colmajor(nd, nd, "complex128")
    out.fill_(complex("nan+nanj"))
    alpha = 0.6 - 0.3j
    nb.assemble(m, m, out, fam, nx, ny, k, alpha=alpha, pol=nb.policy(**gp), accumulate=False,
                formalism="collocation_gauss")
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    assert np.isfinite(got).all()
    ref = orc.assemble_gauss(nodes, quads, fam, nx, ny, k=k, alpha=alpha, pol=orc.policy(**op))
    mask = fp64_mask(nb, m, nx, ny, gp)
    e64 = np.abs(got[mask] - ref[mask]).max() / np.abs(ref[mask]).max()
    rs, fr = rowscaled(got, ref), relfro(got, ref)
    print("  row-scaled %.2e  rel fro %.2e  fp64 entries %.2e (%d)" % (rs, fr, e64, mask.sum()), flush=True)
    assert rs <= ROW_TOL and fr <= 1e-5
    assert e64 <= 1e-9


@pytest.mark.parametrize("k", [4.0, complex(4.0, -0.8)])
def test_gauss_burton_miller_and_identity(nb, orc, qmesh, k):
    import torch
    nodes, quads = qmesh
    m = nb.Mesh(nodes, quads)
    nd = 4 * m.n
    A, B = nb.colmajor(nd, nd), nb.colmajor(nd, nd)
    nb.assemble_bm_gauss(m, k, A, B, pol=nb.policy(**NIHU[0]))
    torch.cuda.synchronize()
    Ar, Br = orc.assemble_bm_gauss(nodes, quads, k, pol=orc.policy(**NIHU[1]))
    ea, eb = relfro(A.cpu().numpy(), Ar), relfro(B.cpu().numpy(), Br)
    print("gauss Burton-Miller k = %s: A %.2e B %.2e" % (k, ea, eb), flush=True)
    assert ea <= 1e-5 and eb <= 1e-5
    # the identity operator over the dofs
    I = nb.colmajor(nd, nd, "complex128")
    nb.assemble(m, m, I, nb.NB_IDENTITY, 0, 0, 0.0, alpha=0.5 + 0.25j, accumulate=False,
                formalism="collocation_gauss")
    torch.cuda.synchronize()
    assert np.array_equal(I.cpu().numpy(), (0.5 + 0.25j) * np.eye(nd))


def test_gauss_field_points(nb, orc, qmesh):
    """A constant test field on a triangle mesh outside the sphere against the
    gauss trial field (post-processing, pf = Mf p - Lf q)."""
    import torch
    from nihu_amd.mesh import icosphere
    nodes, quads = qmesh
    fn, ft = icosphere(2)
    fn = 1.3 * fn
    mq, mf = nb.Mesh(nodes, quads), nb.Mesh(fn, ft)
    for nx, ny in ((0, 0), (0, 1)):
        out = nb.colmajor(mf.n, 4 * mq.n, "complex128")
        out.fill_(complex("nan+nanj"))
        nb.assemble(mf, mq, out, nb.NB_HELMHOLTZ, nx, ny, 3.0, pol=nb.policy(**NIHU[0]), accumulate=False,
                    formalism="collocation_gauss_trial")
        torch.cuda.synchronize()
        got = out.cpu().numpy()
        ref = orc.assemble_gauss(fn, ft, 1, nx, ny, k=3.0, pol=orc.policy(**NIHU[1]), trial=(nodes, quads),
                                 test_gauss=False)
        rs, fr = rowscaled(got, ref), relfro(got, ref)
        print("field points (%d, %d): row-scaled %.2e rel fro %.2e" % (nx, ny, rs, fr), flush=True)
        assert np.isfinite(got).all() and rs <= ROW_TOL and fr <= 1e-5


def test_gauss_many_row_tiles(nb, orc):
    """3456 rows (14 row tiles of the regular pass): sampled rows of the
    Burton-Miller pair against the oracle."""
    import torch
    from nihu_amd.mesh import jittered_cubed_sphere
    nodes, quads = jittered_cubed_sphere(12, seed=7)
    m = nb.Mesh(nodes, quads)
    nd = 4 * m.n
    A, B = nb.colmajor(nd, nd), nb.colmajor(nd, nd)
    A.fill_(complex("nan+nanj"))
    B.fill_(complex("nan+nanj"))
    nb.assemble_bm_gauss(m, 6.0, A, B, pol=nb.policy(**NIHU[0]))
    torch.cuda.synchronize()
    rows = np.sort(np.random.default_rng(2).choice(nd, 48, replace=False)).astype(np.int32)
    Ar, Br = orc.assemble_bm_gauss(nodes, quads, 6.0, pol=orc.policy(**NIHU[1]), rows=rows)
    Ag, Bg = A.cpu().numpy(), B.cpu().numpy()
    assert np.isfinite(Ag).all() and np.isfinite(Bg).all()
    ea, eb = relfro(Ag[rows], Ar), relfro(Bg[rows], Br)
    print("3456 dofs, 48 rows: A %.2e B %.2e" % (ea, eb), flush=True)
    assert ea <= 1e-5 and eb <= 1e-5


def test_gauss_pulsating_sphere(nb):
    """p = -1/(1 + ik) for q = 1 from the device matrices (NiHu's blown-up
    cube sphere, 6 elements per cube edge, 864 dofs)."""
    import torch
    from nihu_amd.mesh import cubed_sphere
    nodes, quads = cubed_sphere(6)
    m = nb.Mesh(nodes, quads)
    nd = 4 * m.n
    k = 1.0
    L, M = nb.colmajor(nd, nd, "complex128"), nb.colmajor(nd, nd, "complex128")
    nb.assemble(m, m, L, nb.NB_HELMHOLTZ, 0, 0, k, accumulate=False, formalism="collocation_gauss")
    nb.assemble(m, m, M, nb.NB_HELMHOLTZ, 0, 1, k, accumulate=False, formalism="collocation_gauss")
    torch.cuda.synchronize()
    Ln, Mn = L.cpu().numpy(), M.cpu().numpy()
    p = np.linalg.solve(Mn - 0.5 * np.eye(nd), Ln @ np.ones(nd))
    panal = -1.0 / (1.0 + 1j * k)
    err = np.abs(p - panal).max() / abs(panal)
    print("gauss-field pulsating sphere on the device: %.2e" % err, flush=True)
    assert err < 0.015


@pytest.mark.parametrize("fam,k", [(0, 0.0), (1, 3.0), (1, complex(2.5, -0.5))])
@pytest.mark.parametrize("nxy", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_gauss_operators_fp64(nb, orc, qmesh, fam, k, nxy):
    """The FP64 precision policy: every block in FP64, against the oracle to 1e-11."""
    import torch
    nodes, quads = qmesh
    nx, ny = nxy
    m = nb.Mesh(nodes, quads)
    nd = 4 * m.n
    out = nb.colmajor(nd, nd, "complex128")
    out.fill_(complex("nan+nanj"))
    nb.assemble(m, m, out, fam, nx, ny, k, alpha=0.6 - 0.3j, pol=nb.policy(**NIHU[0], precision="fp64"),
                accumulate=False, formalism="collocation_gauss")
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    ref = orc.assemble_gauss(nodes, quads, fam, nx, ny, k=k, alpha=0.6 - 0.3j, pol=orc.policy(**NIHU[1]))
    e = relfro(got, ref)
    print("gauss field FP64 fam %d (%d, %d) k = %s: rel fro %.2e" % (fam, nx, ny, k, e), flush=True)
    assert np.isfinite(got).all() and e <= 1e-11


def test_gauss_host_output_accumulates(nb, orc, qmesh):
    """Host outputs (numpy, Fortran order) with accumulation: two calls give
    twice the oracle's matrix."""
    nodes, quads = qmesh
    m = nb.Mesh(nodes, quads)
    nd = 4 * m.n
    out = np.zeros((nd, nd), dtype=np.complex128, order="F")
    for _ in range(2):
        nb.assemble(m, m, out, nb.NB_HELMHOLTZ, 0, 1, 2.0, pol=nb.policy(**NIHU[0]), accumulate=True,
                    formalism="collocation_gauss")
    ref = orc.assemble_gauss(nodes, quads, 1, 0, 1, k=2.0, pol=orc.policy(**NIHU[1]))
    assert relfro(out, 2 * ref) <= 1e-5


def test_gauss_inverse_mapping_failure_is_an_error(nb, qmesh):
    """As for constant quadrilaterals: a near-singular block whose collocation
    point has no image makes the call fail (NiHu throws), forced from the host
    with NBEM_INVERSE_MAPPING_STEPS=1 so that every kernel still finishes."""
    import os
    import torch
    nodes, quads = qmesh
    m = nb.Mesh(nodes, quads)
    nd = 4 * m.n
    A, B = nb.colmajor(nd, nd), nb.colmajor(nd, nd)
    pol = nb.policy(**NIHU[0])
    os.environ["NBEM_INVERSE_MAPPING_STEPS"] = "1"
    try:
        with pytest.raises(nb.NbemError, match="inverse mapping"):
            nb.assemble_bm_gauss(m, 3.0, A, B, pol=pol)
    finally:
        del os.environ["NBEM_INVERSE_MAPPING_STEPS"]
    torch.cuda.synchronize()
    nb.assemble_bm_gauss(m, 3.0, A, B, pol=pol)
    torch.cuda.synchronize()
    assert torch.isfinite(torch.view_as_real(A)).all() and torch.isfinite(torch.view_as_real(B)).all()


def test_gauss_refusals(nb, qmesh):
    from nihu_amd.mesh import icosphere
    nodes, quads = qmesh
    m = nb.Mesh(nodes, quads)
    tn, tt = icosphere(1)
    t = nb.Mesh(tn, tt)
    nd = 4 * m.n
    with pytest.raises(NotImplementedError):
        nb.assemble(t, t, nb.colmajor(4 * t.n, 4 * t.n), 1, 0, 0, 1.0, formalism="collocation_gauss")
    with pytest.raises(NotImplementedError):
        nb.assemble(m, m, nb.colmajor(m.n, nd), 1, 0, 0, 1.0, formalism="collocation_gauss_trial")
    with pytest.raises(ValueError):
        nb.assemble(m, m, nb.colmajor(m.n, m.n), 1, 0, 0, 1.0, formalism="collocation_gauss")
    with pytest.raises(NotImplementedError):
        nb.assemble_bm_gauss(t, 2.0, nb.colmajor(4 * t.n, 4 * t.n), nb.colmajor(4 * t.n, 4 * t.n))
