"""Edge cases of the assembly against the FP64 oracle, through the C ABI:
a single triangle, a flat open plate (every near pair coplanar), a very large
and a very small wavenumber, meshes with unused nodes and with inconsistently
oriented elements (NiHu's node-overlap scan then picks other start corners for
the Galerkin singular rules; the engine must make the same choices).

Bars as in the parity tests: 2e-6 of the row's largest entry and 1e-5
relative Frobenius, also at k = 150, where the FP32 per-point phase
correction k (r - d) reaches 15 rad (the pair phase k d is reduced in FP64).
"""

import numpy as np
import pytest

from tests import bemref

pytestmark = pytest.mark.gpu

ROW_TOL = 2e-6
NIHU = (dict(rule="nihu", near="mean_edge", limit=1.5), dict(rule=0, fixed_degree=0, near_mode=0, near_limit=1.5))
CONFIG = (dict(rule="fixed", degree=5, near="diameter", limit=2.0),
          dict(rule=1, fixed_degree=5, near_mode=1, near_limit=2.0))


@pytest.fixture(scope="module")
def nb():
    import torch
    assert torch.cuda.is_available(), "GPU test without a GPU"
    from nihu_amd import nbem
    return nbem


def rowscaled(got, ref):
    scale = np.abs(ref).max(axis=1, keepdims=True)
    scale[scale == 0] = 1.0
    return (np.abs(got - ref) / scale).max()


def relfro(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _coll(nb, nodes, tris, fam, nx, ny, k, gp, dtype="complex128"):
    import torch
    m = nb.Mesh(nodes, tris)
    out = nb.colmajor(len(tris), len(tris), dtype)
    nb.assemble(m, m, out, fam, nx, ny, k, pol=nb.policy(**gp), accumulate=False)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def _gal(nb, nodes, tris, fam, nx, ny, k, gp, form):
    import torch
    m = nb.Mesh(nodes, tris)
    out = nb.colmajor(m.dofs(form), m.dofs(form), "complex128")
    nb.assemble(m, m, out, fam, nx, ny, k, pol=nb.policy(**gp), accumulate=False, formalism=form)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def test_single_triangle(nb, orc):
    nodes = np.array([[0.0, 0.0, 0.0], [0.4, 0.05, 0.0], [0.1, 0.3, 0.02]])
    tris = np.array([[0, 1, 2]])
    for nx, ny in ((0, 0), (0, 1), (1, 0), (1, 1)):
        got = _coll(nb, nodes, tris, 1, nx, ny, 3.0, NIHU[0])
        ref = orc.assemble(nodes, tris, 1, nx, ny, k=3.0, pol=orc.policy(**NIHU[1]))
        assert np.abs(got - ref).max() <= 1e-12 * max(np.abs(ref).max(), 1.0), (nx, ny)
    for form, lin in (("galerkin", False), ("galerkin_linear", True)):
        got = _gal(nb, nodes, tris, 1, 0, 0, 3.0, NIHU[0], form)
        ref = orc.assemble_galerkin(nodes, tris, 1, 0, 0, k=3.0, pol=orc.policy(**NIHU[1]), linear=lin)
        assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max(), form


@pytest.mark.parametrize("nxy", [(0, 0), (1, 1)])
def test_flat_plate(nb, orc, nxy):
    """Coplanar open surface: near pairs in the element plane (z = 0)."""
    nodes, tris = bemref.square_mesh(8)
    got = _coll(nb, nodes, tris, 1, nxy[0], nxy[1], 4.0, NIHU[0])
    ref = orc.assemble(nodes, tris, 1, nxy[0], nxy[1], k=4.0, pol=orc.policy(**NIHU[1]))
    rs = rowscaled(got, ref)
    print("plate", nxy, "row-scaled %.2e" % rs, flush=True)
    assert rs <= ROW_TOL and relfro(got, ref) <= 1e-5
    for form, lin in (("galerkin", False), ("galerkin_linear", True)):
        g = _gal(nb, nodes, tris, 1, 0, 0, 4.0, NIHU[0], form)
        r = orc.assemble_galerkin(nodes, tris, 1, 0, 0, k=4.0, pol=orc.policy(**NIHU[1]), linear=lin)
        assert rowscaled(g, r) <= ROW_TOL, form


def test_large_wavenumber(nb, orc):
    from nihu_amd.mesh import jittered_icosphere
    nodes, tris = jittered_icosphere(3, seed=6)
    k = 150.0
    for nxy in ((0, 0), (1, 1)):
        got = _coll(nb, nodes, tris, 1, nxy[0], nxy[1], k, CONFIG[0])
        ref = orc.assemble(nodes, tris, 1, nxy[0], nxy[1], k=k, pol=orc.policy(**CONFIG[1]))
        rs = rowscaled(got, ref)
        print("k = 150", nxy, "row-scaled %.2e rel fro %.2e" % (rs, relfro(got, ref)), flush=True)
        assert rs <= ROW_TOL and relfro(got, ref) <= 1e-5


def test_small_wavenumber(nb, orc):
    from nihu_amd.mesh import jittered_icosphere
    nodes, tris = jittered_icosphere(2, seed=6)
    for nxy in ((0, 0), (0, 1), (1, 1)):
        got = _coll(nb, nodes, tris, 1, nxy[0], nxy[1], 1e-4, NIHU[0])
        ref = orc.assemble(nodes, tris, 1, nxy[0], nxy[1], k=1e-4, pol=orc.policy(**NIHU[1]))
        assert rowscaled(got, ref) <= ROW_TOL, nxy


def test_unused_nodes_and_flipped_elements(nb, orc):
    """Nodal Galerkin on a mesh with extra unreferenced nodes (their rows and
    columns stay zero) and a quarter of the elements reversed."""
    from nihu_amd.mesh import jittered_icosphere
    nodes, tris = jittered_icosphere(2, seed=8)
    rng = np.random.default_rng(3)
    flip = rng.random(len(tris)) < 0.25
    tris = tris.copy()
    tris[flip] = tris[flip][:, [0, 2, 1]]
    extra = rng.standard_normal((5, 3)) * 3.0
    nodes = np.vstack([nodes, extra])
    for nxy in ((0, 0), (0, 1)):
        got = _gal(nb, nodes, tris, 1, nxy[0], nxy[1], 2.0, NIHU[0], "galerkin_linear")
        ref = orc.assemble_galerkin(nodes, tris, 1, nxy[0], nxy[1], k=2.0, pol=orc.policy(**NIHU[1]), linear=True)
        assert np.all(got[-5:, :] == 0) and np.all(got[:, -5:] == 0)
        rs = rowscaled(got[:-5, :-5], ref[:-5, :-5])
        print("flipped", nxy, "row-scaled %.2e" % rs, flush=True)
        assert rs <= ROW_TOL
