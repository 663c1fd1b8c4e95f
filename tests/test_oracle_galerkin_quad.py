"""Pins of the oracle's Galerkin formalism on quadrilaterals (CPU only):
constant and isoparametric (bilinear, nodal) fields.

NiHu's tests hold no Galerkin values for quadrilaterals, so the checks stand
apart from the oracle's own formulas:
* the canonical quad x quad singular rules against the closed-form double
  integral of 1/r over the unit square, (4/3)(1 - sqrt 2) + 4 ln(1 + sqrt 2),
  and against a semi-analytic computation for squares sharing an edge or a
  corner (the inner integral over the trial square in closed form at points
  of its plane, the outer one by Gauss rules graded toward the edges);
* element pairs in every relabelling of their nodes (the rotations of the
  rules to the shared corners) against those values, and nodal blocks
  against the degree-31 rule weighted by each element's own shape functions;
* regular pairs of warped quadrilaterals against brute-force tensor Gauss
  rules; the degree of NiHu's policy from the distance table;
* the mass matrix of a square in closed form; Gauss's identity for the
  double layer on a closed surface; NiHu's pulsating-sphere answer
  (test/stallone_unit/helmholtz_3d_pulsating_sphere_test.cpp:28-73) in its
  Galerkin form; the real limit and analyticity of a complex wavenumber.
"""

import numpy as np
import pytest

from tests import bemref
from tests.test_oracle_quad import WARPED

FACE = 4.0 / 3.0 * (1.0 - np.sqrt(2.0)) + 4.0 * np.log(1.0 + np.sqrt(2.0))   # int int 1/r, unit square
UNIT = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], float)


def _shapes(xi, eta):
    xi, eta = np.asarray(xi, float), np.asarray(eta, float)
    return np.stack([(1 - xi) * (1 - eta), (1 + xi) * (1 - eta), (1 + xi) * (1 + eta), (1 - xi) * (1 + eta)], -1) / 4


def _grid(n=3):
    """flat n x n grid of unit squares in z = 0, counter-clockwise nodes"""
    nodes = np.array([[i, j, 0.0] for j in range(n + 1) for i in range(n + 1)])
    idx = lambda i, j: j * (n + 1) + i
    quads = np.array([[idx(i, j), idx(i + 1, j), idx(i + 1, j + 1), idx(i, j + 1)] for j in range(n) for i in range(n)])
    return nodes, quads


def _shift(ids, s):
    return [ids[(s + q) % 4] for q in range(4)]


def _double_wrap(a, b):
    """both elements hold their shared edge at local positions (3, 0): NiHu's
    node-overlap scan (nihu/core/element.hpp:375-413) meets the trial's
    position 3 first and then starts the trial edge at position 0, so the
    edge rule is rotated onto the wrong trial edge and misses the singular
    line -- the same as for triangles at positions (2, 0)"""
    shared = set(a) & set(b)
    return (len(shared) == 2 and sorted(a.index(v) for v in shared) == [0, 3]
            and sorted(b.index(v) for v in shared) == [0, 3])


# ------------------------------------------------------ canonical rules

@pytest.mark.parametrize("match,count", [(2, 2048), (1, 1536), (0, 1280)])
def test_rule_size_weights_domain(orc, match, count):
    r = orc.quad_singular_rule(match)
    assert len(r) == count
    assert abs(r[:, 4].sum() - 16.0) < 1e-13
    assert np.all(np.abs(r[:, :4]) <= 1.0)


def test_face_rule_closed_form(orc):
    """Test and trial points of the face rule on the unit square: the order-7
    rule of NiHu's singular_quadrature_order and the same construction at
    degree 31 against the closed form."""
    for deg, tol in ((7, 2e-6), (31, 1e-14)):
        r = orc.quad_singular_rule(2, deg)
        x, y = (r[:, 0:2] + 1) / 2, (r[:, 2:4] + 1) / 2
        v = np.sum(r[:, 4] / 16 / np.linalg.norm(x - y, axis=1))
        assert abs(v / FACE - 1) < tol, (deg, v)


def test_semi_analytic_helper_closed_form():
    v = bemref.coplanar_squares_slp(((0, 1), (0, 1)), ((0, 1), (0, 1)))
    assert abs(v / FACE - 1) < 1e-12


@pytest.mark.parametrize("match", [1, 0])
def test_edge_corner_rules_semi_analytic(orc, match):
    """Canonical edge (trial square reflected across eta_2 = -1) and corner
    (reflected through (-1, -1)) configurations of unit squares."""
    if match == 1:
        ref = bemref.coplanar_squares_slp(((0, 1), (0, 1)), ((0, 1), (-1, 0)))
    else:
        ref = bemref.coplanar_squares_slp(((0, 1), (0, 1)), ((-1, 0), (-1, 0)))
    for deg, tol in ((7, 1.5e-6), (31, 1e-12)):
        r = orc.quad_singular_rule(match, deg)
        x, y = (r[:, 0:2] + 1) / 2, (r[:, 2:4] + 1) / 2
        if match == 1:
            y[:, 1] = -y[:, 1]
        else:
            y = -y
        v = np.sum(r[:, 4] / 16 / np.linalg.norm(x - y, axis=1))
        assert abs(v / ref - 1) < tol, (deg, v, ref)


# ------------------------------------------- element pairs of a flat grid

def test_overlap_names_the_shared_corners(orc):
    """For every pair of a grid sharing nodes, the start corners of the scan
    put the shared edge at test corners (i1, i1 + 1) = trial corners
    (i2 + 1, i2), the shared corner at (i1, i2), in every relabelling but
    the double wrap, where the scan gives i2 = 0 as NiHu's does."""
    nodes, quads = _grid(3)
    seen = {1: 0, 2: 0, "wrap": 0}
    for e in quads:
        for f in quads:
            for s in range(4):
                for t in range(4):
                    a, b = _shift([int(v) for v in e], s), _shift([int(v) for v in f], t)
                    n, i1, i2 = orc.overlap_quads(a, b)
                    shared = set(a) & set(b)
                    assert n == len(shared)
                    if n == 2 and _double_wrap(a, b):
                        assert (i1, i2) == (3, 0)
                        assert not (a[i1] == b[(i2 + 1) % 4] and a[(i1 + 1) % 4] == b[i2])
                        seen["wrap"] += 1
                    elif n == 2:
                        assert a[i1] == b[(i2 + 1) % 4] and a[(i1 + 1) % 4] == b[i2]
                    elif n == 1:
                        assert a[i1] == b[i2]
                    if n in seen:
                        seen[n] += 1
    assert seen[1] > 0 and seen[2] > 0 and seen["wrap"] > 0


@pytest.mark.parametrize("name,fi,ref_args", [
    ("face", 4, (((1, 2), (1, 2)), ((1, 2), (1, 2)))),
    ("edge", 5, (((1, 2), (1, 2)), ((2, 3), (1, 2)))),
    ("corner", 8, (((1, 2), (1, 2)), ((2, 3), (2, 3))))])
def test_slp_blocks_every_relabelling(orc, name, fi, ref_args):
    """Laplace single layer, constant fields, centre square of a 3 x 3 grid
    against itself, a side neighbour and a diagonal neighbour: every cyclic
    relabelling of both elements gives the semi-analytic value to the order-7
    rules' error, except the double wrap of an edge pair (off by > 1e-3, as
    in NiHu)."""
    nodes, quads = _grid(3)
    ref = bemref.coplanar_squares_slp(*ref_args) / (4 * np.pi)
    e = [int(v) for v in quads[4]]
    f = [int(v) for v in quads[fi]]
    n_wrap = 0
    for s in range(4):
        for t in range(4):
            a, b = _shift(e, s), _shift(f, t)
            if name == "face":
                b = a
            blk = orc.galerkin_quad_block(0, 0, 0, 0.0, nodes[a], a, nodes[b], b, same_elem=(name == "face"),
                                          linear=False)
            assert blk[0, 0].imag == 0.0
            if name == "edge" and _double_wrap(a, b):
                n_wrap += 1
                assert abs(blk[0, 0].real / ref - 1) > 1e-3, (s, t)
            else:
                assert abs(blk[0, 0].real / ref - 1) < 2e-6, (name, s, t, blk[0, 0], ref)
    assert n_wrap == (1 if name == "edge" else 0)


@pytest.mark.parametrize("fi", [5, 8, 7, 6])
def test_nodal_blocks_follow_the_nodes(orc, fi):
    """Isoparametric blocks of an edge or corner pair: relabelling the
    elements permutes the block with the nodes and nothing else (the rules
    are rotated to the same shared corners)."""
    nodes, quads = _grid(3)
    e, f = [int(v) for v in quads[4]], [int(v) for v in quads[fi]]
    base = None
    for s in range(4):
        for t in range(4):
            a, b = _shift(e, s), _shift(f, t)
            if _double_wrap(a, b):
                continue
            blk = orc.galerkin_quad_block(1, 0, 0, 2.0, nodes[a], a, nodes[b], b, linear=True)
            by_node = {(a[i], b[j]): blk[i, j] for i in range(4) for j in range(4)}
            if base is None:
                base = by_node
                scale = max(abs(v) for v in by_node.values())
            for key, v in by_node.items():
                assert abs(v - base[key]) < 1e-13 * scale


@pytest.mark.parametrize("match", [1, 0])
def test_nodal_block_against_degree_31(orc, match):
    """Laplace single layer, nodal fields: the oracle's order-7 block against
    the degree-31 canonical rule weighted by each element's own shape
    functions at the physical points (no rotation machinery)."""
    # test: unit square, nodes 0..3; trial: shares test corners (0, 1) as its
    # corners (1, 0) (edge) or test corner 0 as its corner 0 (corner)
    if match == 1:
        Y = np.array([[1, 0, 0], [0, 0, 0], [0, -1, 0], [1, -1, 0]], float)
        sn = [1, 0, 4, 5]
        # reference coordinates of the trial element at a physical point
        to_ref = lambda P: (1 - 2 * P[:, 0], -2 * P[:, 1] - 1)
    else:
        Y = np.array([[0, 0, 0], [-1, 0, 0], [-1, -1, 0], [0, -1, 0]], float)
        sn = [0, 4, 5, 6]
        to_ref = lambda P: (-2 * P[:, 0] - 1, -2 * P[:, 1] - 1)
    blk = orc.galerkin_quad_block(0, 0, 0, 0.0, UNIT, [0, 1, 2, 3], Y, sn, linear=True)
    r = orc.quad_singular_rule(match, 31)
    x = (r[:, 0:2] + 1) / 2
    y = (r[:, 2:4] + 1) / 2
    if match == 1:
        y[:, 1] = -y[:, 1]
    else:
        y = -y
    G = r[:, 4] / 16 / np.linalg.norm(x - y, axis=1) / (4 * np.pi)
    Nt = _shapes(r[:, 0], r[:, 1])
    Ns = _shapes(*to_ref(y))
    ref = (Nt * G[:, None]).T @ Ns
    assert np.abs(blk.real - ref).max() < 1e-5 * np.abs(ref).max()
    # partition of unity: the block sums to the constant-field value
    con = orc.galerkin_quad_block(0, 0, 0, 0.0, UNIT, [0, 1, 2, 3], Y, sn, linear=False)[0, 0]
    assert abs(blk.sum() - con) < 1e-14 * abs(con)


def test_hypersingular_face_edge_throw_corner_finite(orc):
    nodes, quads = _grid(3)
    e = list(quads[4])
    with pytest.raises(RuntimeError):
        orc.galerkin_quad_block(1, 1, 1, 2.0, nodes[e], e, nodes[e], e, same_elem=True)
    f = list(quads[5])
    with pytest.raises(RuntimeError):
        orc.galerkin_quad_block(1, 1, 1, 2.0, nodes[e], e, nodes[f], f)
    f = list(quads[8])
    v = orc.galerkin_quad_block(1, 1, 1, 2.0, nodes[e], e, nodes[f], f)
    assert np.all(np.isfinite(v))


# ------------------------------------------------------- regular pairs

def _brute_block(family, nx, ny, k, tc, sc, n=12, linear=True):
    g, w = np.polynomial.legendre.leggauss(n)
    XI, ETA = np.meshgrid(g, g, indexing="ij")
    W = np.outer(w, w).ravel()
    xi, eta = XI.ravel(), ETA.ravel()
    Xt, NJt = bemref.quad_geom(tc, xi, eta)
    Ys, NJs = bemref.quad_geom(sc, xi, eta)
    Jt, Js = np.linalg.norm(NJt, axis=1), np.linalg.norm(NJs, axis=1)
    N = _shapes(xi, eta) if linear else np.ones((len(xi), 1))
    out = np.zeros((N.shape[1], N.shape[1]), complex)
    for x, nxv, jt, wt, na in zip(Xt, NJt / Jt[:, None], Jt, W, N):
        K = bemref.np_kernel_pts(family, nx, ny, k, x, Ys, nxv, NJs / Js[:, None])
        out += wt * jt * np.outer(na, (K * W * Js) @ N)
    return out


@pytest.mark.parametrize("nxy", [(0, 0), (0, 1), (1, 0), (1, 1)])
@pytest.mark.parametrize("linear", [False, True])
def test_regular_block_warped_vs_brute(orc, nxy, linear):
    """A regular pair of warped quadrilaterals at a fixed degree of 13 (7 x 7
    points on each) against 12 x 12-point tensor Gauss rules."""
    nx, ny = nxy
    sc = WARPED[[1, 2, 3, 0]] * 0.9 + np.array([0.4, -0.3, 2.2])
    pol = orc.policy(rule=1, fixed_degree=13)
    for k in (3.0, complex(3.0, -0.5)):
        blk = orc.galerkin_quad_block(1, nx, ny, k, WARPED, [0, 1, 2, 3], sc, [4, 5, 6, 7], pol=pol,
                                      linear=linear)
        ref = _brute_block(1, nx, ny, k, WARPED, sc, linear=linear)
        n = 4 if linear else 1
        assert np.abs(blk[:n, :n] - ref).max() < 1e-11 * np.abs(ref).max(), (k, nxy)


@pytest.mark.parametrize("linear", [False, True])
def test_policy_degree_from_distance_table(orc, linear):
    """NiHu's policy: degree = field complexity (1 constant, 2 nodal on
    quad_1) + the 1/r^2 interval of d / R_trial; the block equals the block
    at that fixed degree."""
    tab = [(1.1, 8), (1.5, 6), (2.8, 4), (15.8, 2), (np.inf, 0)]
    R = np.mean([np.linalg.norm(WARPED[(q + 1) % 4] - WARPED[q]) for q in range(4)])
    c = bemref.quad_geom(WARPED, 0.0, 0.0)[0]
    for shift in (1.3, 1.8, 3.5, 20.0):
        tc = WARPED + np.array([0.0, 0.0, shift])
        d = np.linalg.norm(bemref.quad_geom(tc, 0.0, 0.0)[0] - c)
        deg = (2 if linear else 1) + next(dg for lim, dg in tab if d / R < lim)
        blk = orc.galerkin_quad_block(1, 0, 1, 2.0, tc, [4, 5, 6, 7], WARPED, [0, 1, 2, 3], linear=linear)
        fix = orc.galerkin_quad_block(1, 0, 1, 2.0, tc, [4, 5, 6, 7], WARPED, [0, 1, 2, 3],
                                      pol=orc.policy(rule=1, fixed_degree=deg), linear=linear)
        assert np.array_equal(blk, fix), (shift, deg)


# ----------------------------------------------------- whole-mesh answers

@pytest.mark.parametrize("linear", [False, True])
def test_mass_matrix(orc, linear):
    nodes, quads = _grid(2)
    M = orc.assemble_galerkin(nodes, quads, 2, 0, 0, linear=linear).real
    if linear:
        loc = np.array([[4, 2, 1, 2], [2, 4, 2, 1], [1, 2, 4, 2], [2, 1, 2, 4]]) / 36.0
        ref = np.zeros((len(nodes), len(nodes)))
        for q in quads:
            ref[np.ix_(q, q)] += loc
    else:
        ref = np.eye(len(quads))
    assert np.abs(M - ref).max() < 1e-15
    # warped: the entries sum to the Gauss rule of the area (degree 3, 2 x 2
    # points, nodal; degree 1, the centre, constant), close to the area
    Mw = orc.assemble_galerkin(WARPED, np.array([[0, 1, 2, 3]]), 2, 0, 0, linear=linear).real

    def area(n):
        g, w = np.polynomial.legendre.leggauss(n)
        XI, ETA = np.meshgrid(g, g, indexing="ij")
        J = np.linalg.norm(bemref.quad_geom(WARPED, XI.ravel(), ETA.ravel())[1], axis=1)
        return np.sum(J * np.outer(w, w).ravel())

    assert abs(Mw.sum() / area(2 if linear else 1) - 1) < 1e-14
    assert abs(Mw.sum() / area(20) - 1) < 3e-2


@pytest.mark.parametrize("linear", [False, True])
def test_gauss_identity_closed_surface(orc, linear):
    """int_Gamma dG/dn_y dS_y = -1/2 on a closed surface of bilinear patches:
    row sums of the Galerkin double layer are -1/2 times the mass-matrix row
    sums."""
    from nihu_amd.mesh import cubed_sphere
    nodes, quads = cubed_sphere(4)
    Dm = orc.assemble_galerkin(nodes, quads, 0, 0, 1, linear=linear)
    Ms = orc.assemble_galerkin(nodes, quads, 2, 0, 0, linear=linear)
    rs, ms = Dm.sum(1).real, Ms.sum(1).real
    assert np.abs(rs + 0.5 * ms).max() < 2e-3 * ms.max()


@pytest.mark.parametrize("linear", [False, True])
def test_pulsating_sphere(orc, linear):
    """(M - I/2) p = L q with Galerkin matrices on NiHu's blown-up cube sphere
    of quadrilaterals: p = -q / (1 + i k) for the pulsating unit sphere, at a
    real and a complex wavenumber."""
    from nihu_amd.mesh import cubed_sphere
    nodes, quads = cubed_sphere(5)
    n = len(nodes) if linear else len(quads)
    I = orc.assemble_galerkin(nodes, quads, 2, 0, 0, linear=linear)
    for k in (1.0, complex(1.0, -0.3)):
        L = orc.assemble_galerkin(nodes, quads, 1, 0, 0, k=k, linear=linear)
        M = orc.assemble_galerkin(nodes, quads, 1, 0, 1, k=k, linear=linear)
        p = np.linalg.solve(M - 0.5 * I, L @ np.ones(n))
        panal = -1.0 / (1.0 + 1j * k)
        err = np.abs(p - panal).max() / abs(panal)
        print("Galerkin quadrilaterals, linear=%s, k=%s: %.2e" % (linear, k, err))
        assert err < 0.02


def test_complex_k_real_limit_and_analytic(orc):
    nodes, quads = _grid(3)
    e, f = list(quads[4]), list(quads[5])
    for nx, ny in ((0, 0), (0, 1), (1, 0)):
        F = lambda k: orc.galerkin_quad_block(1, nx, ny, k, nodes[e], e, nodes[f], f)
        assert np.abs(F(complex(2.5, 0.0)) - F(2.5)).max() <= 1e-15 * np.abs(F(2.5)).max()
        kc, h = complex(2.5, -0.4), 1e-5
        d_re = (F(kc + h) - F(kc - h)) / (2 * h)
        d_im = (F(kc + 1j * h) - F(kc - 1j * h)) / (2 * h)
        assert np.abs(d_im - 1j * d_re).max() <= 1e-6 * np.abs(d_re).max()


def test_other_mesh_is_regular_only(orc):
    """Different meshes: no singular check, a hypersingular operator between
    disjoint quadrilateral meshes assembles."""
    nodes, quads = _grid(2)
    far = (nodes + np.array([0.0, 0.0, 2.0]), quads)
    A = orc.assemble_galerkin(nodes, quads, 1, 1, 1, k=1.0, trial=far, linear=True)
    assert np.all(np.isfinite(A))
    with pytest.raises(Exception):
        orc.assemble_galerkin(nodes, quads, 1, 1, 1, k=1.0, linear=True)
