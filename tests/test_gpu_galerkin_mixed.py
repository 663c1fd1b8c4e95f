"""Galerkin pairs with mixed field types through the C ABI, against the FP64
oracle: constant test x isoparametric trial ("galerkin_const_linear", rows =
elements, columns = nodes) and the reverse ("galerkin_linear_const"), on
triangle and quadrilateral meshes.  The engine runs the nodal kernels with the
constant side's dof map naming the element at every corner; the bar is that of
the other Galerkin matrices: rows within 2e-6 of their largest entry, 1e-5
relative Frobenius, the mass matrix to 1e-14.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROW_TOL = 2e-6
FRO_TOL = 1e-5
CODES = {"galerkin_const_linear": 2, "galerkin_linear_const": 3}
NIHU = (dict(rule="nihu", near="mean_edge", limit=1.5), dict(rule=0, fixed_degree=0, near_mode=0, near_limit=1.5))


@pytest.fixture(scope="module")
def nb():
    import torch
    assert torch.cuda.is_available(), "GPU test without a GPU"
    from nihu_amd import nbem
    return nbem


@pytest.fixture(scope="module", params=["tri", "quad"])
def mesh(request):
    from nihu_amd.mesh import jittered_icosphere, jittered_cubed_sphere
    if request.param == "tri":
        return request.param, jittered_icosphere(2, seed=4)
    return request.param, jittered_cubed_sphere(4, seed=4)


def rowscaled(got, ref):
    scale = np.abs(ref).max(axis=1, keepdims=True)
    scale[scale == 0] = 1.0
    return (np.abs(got - ref) / scale).max()


def relfro(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _assemble(nb, m, form, fam, nx, ny, k, alpha, gp, dtype="complex128"):
    import torch
    out = nb.colmajor(m.dofs(form, "test"), m.dofs(form, "trial"), dtype)
    nb.assemble(m, m, out, fam, nx, ny, k, alpha=alpha, pol=nb.policy(**gp), accumulate=False, formalism=form)
    torch.cuda.synchronize()
    return out.cpu().numpy()


@pytest.mark.parametrize("form", ["galerkin_const_linear", "galerkin_linear_const"])
@pytest.mark.parametrize("fam,k", [(0, 0.0), (1, 3.0), (1, complex(3.0, -0.5))])
@pytest.mark.parametrize("nxy", [(0, 0), (0, 1), (1, 0)])
def test_mixed_operator(nb, orc, mesh, form, fam, k, nxy):
    kind, (nodes, el) = mesh
    m = nb.Mesh(nodes, el)
    print("case", kind, form, fam, k, nxy, flush=True)
    got = _assemble(nb, m, form, fam, nxy[0], nxy[1], k, 0.8 + 0.25j, NIHU[0])
    ref = orc.assemble_galerkin(nodes, el, fam, nxy[0], nxy[1], k=k, alpha=0.8 + 0.25j, pol=orc.policy(**NIHU[1]),
                                linear=CODES[form])
    assert got.shape == ref.shape
    rs, rf = rowscaled(got, ref), relfro(got, ref)
    print("row-scaled %.2e  rel fro %.2e" % (rs, rf), flush=True)
    assert rs <= ROW_TOL and rf <= FRO_TOL


@pytest.mark.parametrize("form", ["galerkin_const_linear", "galerkin_linear_const"])
def test_mixed_mass_matrix(nb, orc, mesh, form):
    kind, (nodes, el) = mesh
    m = nb.Mesh(nodes, el)
    got = _assemble(nb, m, form, nb.NB_IDENTITY, 0, 0, 0.0, -0.5, NIHU[0])
    ref = orc.assemble_galerkin(nodes, el, 2, 0, 0, alpha=-0.5, linear=CODES[form])
    assert np.abs(got - ref).max() <= 1e-14 * np.abs(ref).max()


def test_mixed_then_plain_layouts(nb, orc, mesh):
    """The node-dof and element-dof layouts of one mesh live side by side:
    a mixed assembly, then the plain nodal one, then the mixed one again, give
    the oracle's matrices each time."""
    kind, (nodes, el) = mesh
    m = nb.Mesh(nodes, el)
    for form, code in (("galerkin_const_linear", 2), ("galerkin_linear", 1), ("galerkin_const_linear", 2)):
        got = _assemble(nb, m, form, 1, 0, 0, 2.0, 1.0, NIHU[0])
        ref = orc.assemble_galerkin(nodes, el, 1, 0, 0, k=2.0, pol=orc.policy(**NIHU[1]), linear=code)
        assert rowscaled(got, ref) <= ROW_TOL and relfro(got, ref) <= FRO_TOL


@pytest.mark.parametrize("form", ["galerkin_const_linear", "galerkin_linear_const"])
@pytest.mark.parametrize("nxy", [(0, 0), (1, 1)])
def test_mixed_other_mesh(nb, orc, mesh, form, nxy):
    """Different test and trial meshes of one element type: no singular
    check, so the hypersingular kernel assembles too; each mesh uses the
    layout of its own side's dof kind."""
    import torch
    from nihu_amd.mesh import icosphere, cubed_sphere
    kind, (nodes, el) = mesh
    fn, fe = icosphere(1) if kind == "tri" else cubed_sphere(2)
    fn = fn * 1.7 + np.array([0.2, -0.1, 0.3])
    mt, ms = nb.Mesh(nodes, el), nb.Mesh(fn, fe)
    out = nb.colmajor(mt.dofs(form, "test"), ms.dofs(form, "trial"), "complex128")
    nb.assemble(mt, ms, out, 1, nxy[0], nxy[1], 2.5, pol=nb.policy(**NIHU[0]), accumulate=False, formalism=form)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    ref = orc.assemble_galerkin(nodes, el, 1, nxy[0], nxy[1], k=2.5, pol=orc.policy(**NIHU[1]), trial=(fn, fe),
                                linear=CODES[form])
    assert got.shape == ref.shape
    assert rowscaled(got, ref) <= ROW_TOL and relfro(got, ref) <= FRO_TOL
