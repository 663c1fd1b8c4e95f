"""The oracle's quadrilateral gauss field (nihu/library/quad_1_gauss_field.hpp:
four dofs per element at the 2 x 2 Gauss points, collocated there), checked
against what does not restate its formulas:

* the shape functions against the Lagrange polynomials written from their
  definition (nodes of lib_shape.cpp:128-134), with finite-difference
  gradients;
* partition of unity: the four trial dofs of every element integral sum to
  the constant-field integral of the same method at the same point (the
  constant-field quadrilateral path is pinned in test_oracle_quad.py) -- this
  covers the regular, polar, Telles, Duffy and Guiggiani integrals, the last
  one off the centre;
* each dof on its own: brute-force subdivided quadrature with the shape
  function as weight (regular and near-singular pairs), and for the
  hypersingular diagonal the finite part by exclusion of a ball with the
  shape function as weight (Richardson in the radius);
* whole matrices on cubed spheres: the pulsating sphere, the transparent
  point-source known answer at the collocation nodes, Gauss's identity for
  the Laplace double layer and the hypersingular operator annihilating
  constants; the Burton-Miller pair equals its operators plus the identity.
"""

import numpy as np
import pytest

from tests import bemref as B
from tests.test_oracle_quad import WARPED, _linearised, _point_cases

PAIRS = [(0, 0), (0, 1), (1, 0), (1, 1)]
G = np.sqrt(3.0) / 3.0     # the nodes as lib_shape.cpp:128-134 writes them
NODES = [(-G, -G), (G, -G), (-G, G), (G, G)]


def test_shape_functions(orc):
    rng = np.random.default_rng(3)
    for xi, eta in rng.uniform(-1, 1, size=(20, 2)):
        N, dx, dy, nodes = orc.gauss_shape(xi, eta)
        ref = np.array([B.gauss_field_shape(b)(xi, eta) for b in range(4)])
        assert np.allclose(N, ref, rtol=0, atol=1e-15)
        assert abs(N.sum() - 1.0) < 1e-15 and abs(dx.sum()) < 1e-15 and abs(dy.sum()) < 1e-15
        h = 1e-6
        fdx = (orc.gauss_shape(xi + h, eta)[0] - orc.gauss_shape(xi - h, eta)[0]) / (2 * h)
        fdy = (orc.gauss_shape(xi, eta + h)[0] - orc.gauss_shape(xi, eta - h)[0]) / (2 * h)
        assert np.allclose(dx, fdx, atol=1e-9) and np.allclose(dy, fdy, atol=1e-9)
    assert np.allclose(nodes, NODES, rtol=0, atol=3e-16)
    for a, (xi, eta) in enumerate(NODES):
        assert np.allclose(orc.gauss_shape(xi, eta)[0], np.eye(4)[a], rtol=0, atol=1e-15)


@pytest.mark.parametrize("fam,k", [(0, 0.0), (1, 3.0), (1, complex(2.5, -0.5))])
@pytest.mark.parametrize("nxy", PAIRS)
def test_partition_of_unity(orc, fam, k, nxy):
    """Sum over the four trial dofs = the constant-field integral."""
    nx, ny = nxy
    x0, nx0 = _point_cases()["above 0.05"]
    far = dict(x0=np.array([3.0, -1.0, 2.0]), nx0=np.array([0.0, 0.6, 0.8]))
    cases = [
        (orc.GAUSS_REGULAR, orc.QUAD_REGULAR, dict(far, o1=9)),
        (orc.GAUSS_POLAR, orc.QUAD_POLAR, dict(x0=x0, nx0=nx0, o1=15, o2=15)),
        (orc.GAUSS_TELLES, orc.QUAD_TELLES, dict(x0=x0, nx0=nx0, o1=30)),
        (orc.GAUSS_NEAR, orc.QUAD_NEAR, dict(x0=x0, nx0=nx0)),
        (orc.GAUSS_DUFFY, orc.QUAD_DUFFY, dict(xi0=(0.4, -0.25), o1=7)),
    ]
    if nx == ny == 1:
        cases.append((orc.GAUSS_GUIGGIANI, orc.QUAD_GUIGGIANI, dict(xi0=(0.4, -0.25), o1=5, o2=9)))
    for wg, wc, kw in cases:
        g = orc.gauss_elem(wg, fam, nx, ny, k, WARPED, **kw)
        c = orc.quad_elem(wc, fam, nx, ny, k, WARPED, **kw)
        assert abs(g.sum() - c) <= 1e-12 * max(abs(c), np.abs(g).max()), (wg, g.sum(), c)
    # the singular integral at node a: Duffy degree 7 or Guiggiani (5, 9) there
    for a, xi0 in enumerate(NODES):
        g = orc.gauss_elem(orc.GAUSS_SINGULAR, fam, nx, ny, k, WARPED, o1=a)
        if nx == ny == 1:
            c = orc.quad_elem(orc.QUAD_GUIGGIANI, fam, 1, 1, k, WARPED, xi0=xi0, o1=5, o2=9)
        else:
            c = orc.quad_elem(orc.QUAD_DUFFY, fam, nx, ny, k, WARPED, xi0=xi0, o1=7)
        assert abs(g.sum() - c) <= 1e-12 * max(abs(c), np.abs(g).max()), (a, g.sum(), c)


@pytest.mark.parametrize("fam,k", [(0, 0.0), (1, 3.0)])
def test_regular_brute_force(orc, fam, k):
    x0 = np.array([3.0, -1.0, 2.0])
    nx0 = np.array([0.0, 0.6, 0.8])
    for nx, ny in PAIRS:
        g = orc.gauss_elem(orc.GAUSS_REGULAR, fam, nx, ny, k, WARPED, x0=x0, nx0=nx0, o1=13)
        ref = np.array([B.quad_brute(fam, nx, ny, k, WARPED, x0, nx0, levels=2, n_gauss=12,
                                     weight=B.gauss_field_shape(b)) for b in range(4)])
        assert np.abs(g - ref).max() <= 1e-9 * np.abs(ref).max(), (nx, ny)


@pytest.mark.parametrize("fam,k", [(0, 0.0), (1, 3.0)])
def test_near_brute_force(orc, fam, k):
    """Telles per dof against brute force (its error is the constant-field
    one, up to 1.7e-3 of the block for the hypersingular kernel 0.05 above
    the element); the polar method's remaining difference per dof is exactly
    N_b(xi0) times that of its planar part."""
    for name, (x0, nx0) in _point_cases().items():
        xi = orc.quad_elem(orc.QUAD_INVERSE, 0, 0, 0, 0.0, WARPED, x0=x0)
        lin = _linearised(WARPED, xi[0], xi[1])
        N0 = orc.gauss_shape(xi[0], xi[1])[0]
        for nx, ny in PAIRS:
            ref = np.array([B.quad_brute(fam, nx, ny, k, WARPED, x0, nx0, levels=6, n_gauss=10,
                                         weight=B.gauss_field_shape(b)) for b in range(4)])
            scale = np.abs(ref).max()
            tel = orc.gauss_elem(orc.GAUSS_TELLES, fam, nx, ny, k, WARPED, x0=x0, nx0=nx0, o1=30)
            assert np.abs(tel - ref).max() <= 3e-3 * scale, (name, nx, ny)
            pol_hi = orc.gauss_elem(orc.GAUSS_POLAR, fam, nx, ny, k, WARPED, x0=x0, nx0=nx0, o1=101, o2=101)
            planar = orc.quad_elem(orc.QUAD_PLANAR, fam, nx, ny, k, lin, x0=x0, nx0=nx0)
            planar_ref = B.quad_brute(fam, nx, ny, k, lin, x0, nx0, levels=6, n_gauss=10)
            dev = (pol_hi - ref) - N0 * (planar - planar_ref)
            assert np.abs(dev).max() <= 2e-6 * scale, (name, nx, ny, np.abs(dev).max() / scale)


@pytest.mark.parametrize("fam,k", [(0, 0.0), (1, 3.0), (1, complex(2.0, -0.5))])
@pytest.mark.parametrize("a", [0, 3])
def test_guiggiani_finite_part_per_dof(orc, fam, k, a):
    """The hypersingular diagonal at collocation node a, each dof against the
    finite part by exclusion of a ball with the shape function as weight
    (the shape-function terms of the Laurent coefficients matter here: their
    sum over the dofs is zero)."""
    xi0 = NODES[a]
    hi = orc.gauss_elem(orc.GAUSS_GUIGGIANI, fam, 1, 1, k, WARPED, xi0=xi0, o1=21, o2=41)
    lo = orc.gauss_elem(orc.GAUSS_SINGULAR, fam, 1, 1, k, WARPED, o1=a)
    ref = np.array([B.hsp_finite_part(fam, k, WARPED, xi0, weight=B.gauss_field_shape(b)) for b in range(4)])
    scale = np.abs(ref).max()
    print(a, "hi %.1e lo %.1e" % (np.abs(hi - ref).max() / scale, np.abs(lo - ref).max() / scale), flush=True)
    assert np.abs(hi - ref).max() <= 1e-7 * scale
    assert np.abs(lo - ref).max() <= 5e-3 * scale


def test_duffy_converges(orc):
    """The weakly singular diagonal per dof is the degree-7 blind Duffy rule
    around the node (NiHu's singular quadrature order); the construction
    converges with the degree (41 against 61: 1e-6 of the block).  Degree 7
    is coarse on this strongly warped element for the double layers (up to
    30 % at the nodes, as for the constant field at the same points); on the
    SLP it is within 1e-2."""
    for a, xi0 in enumerate(NODES):
        for nx, ny in ((0, 0), (0, 1), (1, 0)):
            lo = orc.gauss_elem(orc.GAUSS_SINGULAR, 1, nx, ny, 2.0, WARPED, o1=a)
            d7 = orc.gauss_elem(orc.GAUSS_DUFFY, 1, nx, ny, 2.0, WARPED, xi0=xi0, o1=7)
            assert np.array_equal(lo, d7)
            hi = orc.gauss_elem(orc.GAUSS_DUFFY, 1, nx, ny, 2.0, WARPED, xi0=xi0, o1=41)
            top = orc.gauss_elem(orc.GAUSS_DUFFY, 1, nx, ny, 2.0, WARPED, xi0=xi0, o1=61)
            scale = np.abs(top).max()
            assert np.abs(hi - top).max() <= 1e-6 * scale, (a, nx, ny)
            if nx == ny == 0:
                assert np.abs(lo - top).max() <= 1e-2 * scale, (a, nx, ny)


def _node_geometry(nodes, quads):
    X, N = [], []
    for q in quads:
        for xi, eta in NODES:
            x, nj = B.quad_geom(nodes[q], xi, eta)
            X.append(x)
            N.append(nj / np.linalg.norm(nj))
    return np.array(X), np.array(N)


@pytest.mark.parametrize("k", [1.0, complex(1.0, -0.3)])
def test_pulsating_sphere(orc, k):
    """p = -1/(1 + ik) for q = 1 on NiHu's blown-up cube sphere: the error
    falls with the mesh (4 and 6 elements per cube edge) and is below the
    constant field's on the same mesh."""
    from nihu_amd.mesh import cubed_sphere
    panal = -1.0 / (1.0 + 1j * k)
    errs = []
    for n in (4, 6):
        nodes, quads = cubed_sphere(n)
        L = orc.assemble_gauss(nodes, quads, 1, 0, 0, k)
        M = orc.assemble_gauss(nodes, quads, 1, 0, 1, k)
        p = np.linalg.solve(M - 0.5 * np.eye(len(M)), L @ np.ones(len(M)))
        errs.append(np.abs(p - panal).max() / abs(panal))
    Lc = orc.assemble(nodes, quads, 1, 0, 0, k)
    Mc = orc.assemble(nodes, quads, 1, 0, 1, k)
    pc = np.linalg.solve(Mc - 0.5 * np.eye(len(Mc)), Lc @ np.ones(len(Mc)))
    ec = np.abs(pc - panal).max() / abs(panal)
    print("gauss-field pulsating sphere, k = %s: %.2e, %.2e (constant field %.2e)" % (k, errs[0], errs[1], ec))
    assert errs[1] < 0.015 and errs[1] < 0.6 * errs[0] and errs[1] < ec


def test_transparent_point_source(orc):
    from nihu_amd.mesh import jittered_cubed_sphere
    nodes, quads = jittered_cubed_sphere(5, seed=2)
    k = 2.0
    xs = np.array([0.2, -0.1, 0.15])
    X, N = _node_geometry(nodes, quads)
    rv = X - xs
    r = np.linalg.norm(rv, axis=1)
    p = np.exp(-1j * k * r) / (4 * np.pi * r)
    q = p * (-1j * k - 1 / r) * np.sum(rv * N, axis=1) / r
    G = orc.assemble_gauss(nodes, quads, 1, 0, 0, k)
    H = orc.assemble_gauss(nodes, quads, 1, 0, 1, k)
    rel = np.linalg.norm((H - 0.5 * np.eye(len(p))) @ p - G @ q) / np.linalg.norm(G @ q)
    print("gauss-field transparent residual: %.2e" % rel)
    assert rel < 5e-3


def test_laplace_identities_and_bm(orc):
    from nihu_amd.mesh import jittered_cubed_sphere
    nodes, quads = jittered_cubed_sphere(4, seed=1)
    D = orc.assemble_gauss(nodes, quads, 0, 0, 1).real
    rs = D.sum(axis=1)
    print("gauss-field Laplace DLP row sums: max |row sum + 1/2| %.2e" % np.abs(rs + 0.5).max())
    assert np.abs(rs + 0.5).max() < 1e-2
    Hs = orc.assemble_gauss(nodes, quads, 0, 1, 1).real
    assert np.abs(Hs.sum(axis=1)).max() < 5e-2 * np.abs(np.diag(Hs)).max()
    k = 3.0
    A, Bm = orc.assemble_bm_gauss(nodes, quads, k)
    al = -1j / k
    ops = {(nx, ny): orc.assemble_gauss(nodes, quads, 1, nx, ny, k) for nx, ny in PAIRS}
    I = np.eye(A.shape[0])
    assert np.allclose(A, al * ops[(1, 1)] - 0.5 * I + ops[(0, 1)], rtol=0, atol=1e-13 * np.abs(A).max())
    assert np.allclose(Bm, ops[(0, 0)] + al * ops[(1, 0)] + al / 2 * I, rtol=0, atol=1e-13 * np.abs(Bm).max())


def test_field_points_against_constant_sum(orc):
    """Field points (a constant test field on another mesh): summing the four
    dofs of each trial element gives the constant-field matrix where both use
    the same rule -- a fixed Gauss degree here."""
    from nihu_amd.mesh import jittered_cubed_sphere, icosphere
    nodes, quads = jittered_cubed_sphere(3, seed=6)
    fn, ft = icosphere(1)
    fn = 2.2 * fn
    pol = orc.policy(rule=1, fixed_degree=5, near_mode=0, near_limit=1.5)
    for nx, ny in ((0, 0), (0, 1)):
        g = orc.assemble_gauss(fn, ft, 1, nx, ny, 2.0, pol=pol, trial=(nodes, quads), test_gauss=False)
        c = orc.assemble(fn, ft, 1, nx, ny, k=2.0, pol=pol, trial=(nodes, quads))
        gs = g.reshape(g.shape[0], -1, 4).sum(axis=2)
        assert np.abs(gs - c).max() <= 1e-13 * np.abs(c).max()
