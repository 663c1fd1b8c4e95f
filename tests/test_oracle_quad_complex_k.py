"""Pins of the oracle's quadrilateral path at a complex wavenumber (CPU only).
NiHu's Helmholtz kernel takes a complex WaveNumber (helmholtz_kernel.hpp:103)
and its tests hold no complex-wavenumber values, so every check is against
something other than the oracle's own formula: the real-wavenumber
quadrilateral path (pinned in test_oracle_quad.py) at Im k = 0 and continued
analytically (Cauchy-Riemann in k), brute-force subdivided quadrature of near
pairs, the Hadamard finite part by exclusion of a ball for the hypersingular
diagonal, a high-order Duffy rule for the weakly singular diagonal, and the
exact answers of the pulsating-sphere and transparent point-source problems,
which hold for complex k as for real k."""

import numpy as np
import pytest

from tests import bemref as B
from tests.test_oracle_quad import WARPED, _linearised, _point_cases

KC = complex(3.0, -0.4)
PAIRS = [(0, 0), (0, 1), (1, 0), (1, 1)]


def _cases(orc):
    """(what, kwargs) of every element integral of the quadrilateral path."""
    x0, nx0 = _point_cases()["above 0.05"]
    far = dict(x0=np.array([3.0, -1.0, 2.0]), nx0=np.array([0.0, 0.6, 0.8]))
    return [
        (orc.QUAD_REGULAR, dict(far, o1=9)),
        (orc.QUAD_SINGULAR, {}),
        (orc.QUAD_POLAR, dict(x0=x0, nx0=nx0, o1=15, o2=15)),
        (orc.QUAD_TELLES, dict(x0=x0, nx0=nx0, o1=30)),
        (orc.QUAD_NEAR, dict(x0=x0, nx0=nx0)),
        (orc.QUAD_DUFFY, dict(xi0=(0.3, -0.2), o1=7)),
    ]


@pytest.mark.parametrize("nxy", PAIRS)
def test_real_limit(orc, nxy):
    """k + 0i through the complex path gives the real-wavenumber integrals."""
    a, b = nxy
    for what, kw in _cases(orc):
        real = orc.quad_elem(what, 1, a, b, 3.0, WARPED, **kw)
        cplx = orc.quad_elem(what, 1, a, b, complex(3.0, 0.0), WARPED, **kw)
        assert abs(cplx - real) <= 1e-14 * abs(real), (what, a, b, cplx, real)
    if a == b == 1:
        real = orc.quad_elem(orc.QUAD_GUIGGIANI, 1, 1, 1, 3.0, WARPED, xi0=(0.5, -0.3), o1=5, o2=9)
        cplx = orc.quad_elem(orc.QUAD_GUIGGIANI, 1, 1, 1, complex(3.0, 0.0), WARPED, xi0=(0.5, -0.3), o1=5, o2=9)
        assert abs(cplx - real) <= 1e-14 * abs(real)


@pytest.mark.parametrize("nxy", PAIRS)
def test_analytic_in_k(orc, nxy):
    """Each integral is analytic in k: its derivative along Im k is i times its
    derivative along Re k (central differences)."""
    a, b = nxy
    h = 1e-5
    for what, kw in _cases(orc):
        F = lambda k: orc.quad_elem(what, 1, a, b, k, WARPED, **kw)
        d_re = (F(KC + h) - F(KC - h)) / (2 * h)
        d_im = (F(KC + 1j * h) - F(KC - 1j * h)) / (2 * h)
        assert abs(d_im - 1j * d_re) <= 1e-6 * abs(d_re) + 1e-12, (what, a, b, d_im, d_re)


@pytest.mark.parametrize("k", [complex(3.0, -0.6), complex(2.0, 0.5)])
def test_near_methods_vs_brute(orc, k):
    """Telles against brute force: its error (that of the 1/r^3 core, up to
    1.7e-3 for the hypersingular kernel 0.05 above the element) is the one it
    makes at Re k, to 1e-6; the polar method's remaining difference is exactly
    that of its planar part, as for real k."""
    for name, (x0, nx0) in _point_cases().items():
        xi = orc.quad_elem(orc.QUAD_INVERSE, 0, 0, 0, 0.0, WARPED, x0=x0)
        lin = _linearised(WARPED, xi[0], xi[1])
        for nx, ny in PAIRS:
            ref = B.quad_brute(1, nx, ny, k, WARPED, x0, nx0, levels=6, n_gauss=10)
            ref_real = B.quad_brute(1, nx, ny, k.real, WARPED, x0, nx0, levels=6, n_gauss=10)
            tel = orc.quad_elem(orc.QUAD_TELLES, 1, nx, ny, k, WARPED, x0=x0, nx0=nx0, o1=30)
            tel_real = orc.quad_elem(orc.QUAD_TELLES, 1, nx, ny, k.real, WARPED, x0=x0, nx0=nx0, o1=30)
            assert abs(tel - ref) <= 3e-3 * abs(ref), (name, nx, ny)
            assert abs((tel - ref) - (tel_real - ref_real)) <= 1e-6 * abs(ref), (name, nx, ny)
            assert abs(ref - ref_real) > 1e-2 * abs(ref), (name, nx, ny)
            pol_hi = orc.quad_elem(orc.QUAD_POLAR, 1, nx, ny, k, WARPED, x0=x0, nx0=nx0, o1=101, o2=101)
            planar = orc.quad_elem(orc.QUAD_PLANAR, 1, nx, ny, k, lin, x0=x0, nx0=nx0)
            planar_ref = B.quad_brute(1, nx, ny, k, lin, x0, nx0, levels=6, n_gauss=10)
            assert abs((pol_hi - ref) - (planar - planar_ref)) <= 2e-6 * abs(ref), (name, nx, ny)


@pytest.mark.parametrize("k", [complex(3.0, -0.5), complex(1.0, 0.8)])
@pytest.mark.parametrize("xi0", [(0.0, 0.0), (0.5, -0.3)])
def test_guiggiani_warped_finite_part(orc, k, xi0):
    """Hypersingular diagonal: Guiggiani's finite part (the Laurent terms are
    those of the Laplace core for any k) against the finite part by exclusion
    of a ball, computed independently."""
    ref = B.hsp_finite_part(1, k, WARPED, xi0)
    hi = orc.quad_elem(orc.QUAD_GUIGGIANI, 1, 1, 1, k, WARPED, xi0=xi0, o1=21, o2=41)
    lo = orc.quad_elem(orc.QUAD_GUIGGIANI, 1, 1, 1, k, WARPED, xi0=xi0, o1=5, o2=9)
    ref_real = B.hsp_finite_part(1, k.real, WARPED, xi0)
    assert abs(hi - ref) / abs(ref) < 1e-7, abs(hi - ref) / abs(ref)
    assert abs(lo - ref) / abs(ref) < 5e-3
    assert abs(hi - ref) < 1e-3 * abs(ref - ref_real)


def test_singular_blind_against_high_duffy(orc):
    """Weakly singular diagonal (SLP, DLP, DLPt): the order-7 blind rule
    against the same construction at degree 41.  Its error (3 % for DLP and
    DLPt on the warped element, NiHu's rule) comes from the static singular
    core: at complex k it is the error at Re k to within 5 % of itself."""
    for nx, ny in ((0, 0), (0, 1), (1, 0)):
        err = []
        for k in (KC, KC.real):
            lo = orc.quad_elem(orc.QUAD_SINGULAR, 1, nx, ny, k, WARPED)
            hi = orc.quad_elem(orc.QUAD_DUFFY, 1, nx, ny, k, WARPED, xi0=(0.0, 0.0), o1=41)
            assert abs(lo - hi) <= 5e-2 * abs(hi), (nx, ny, k, lo, hi)
            err.append(lo - hi)
        assert abs(err[0] - err[1]) <= 5e-2 * abs(err[1]), (nx, ny, err)


def test_pulsating_sphere(orc):
    """Analytic p = -1/(1 + ik) q for the pulsating unit sphere
    (nihu/test/stallone_unit/helmholtz_3d_pulsating_sphere_test.cpp:28-73) on
    NiHu's blown-up cube sphere of quadrilaterals, at a complex wavenumber."""
    from nihu_amd.mesh import cubed_sphere
    nodes, quads = cubed_sphere(6)
    k = complex(1.0, -0.3)
    L = orc.assemble(nodes, quads, 1, 0, 0, k)
    M = orc.assemble(nodes, quads, 1, 0, 1, k)
    p = np.linalg.solve(M - 0.5 * np.eye(len(quads)), L @ np.ones(len(quads)))
    panal = -1.0 / (1.0 + 1j * k)
    err = np.abs(p - panal).max() / abs(panal)
    print("quadrilateral pulsating sphere, complex k: %.2e" % err)
    assert err < 0.02


def test_transparent_point_source(orc):
    """Interior point source: p = G(x, x_s), q = dG/dn_y on the surface satisfy
    (DLP - I/2) p = SLP q to discretisation error, at a complex wavenumber."""
    from nihu_amd.mesh import jittered_cubed_sphere
    nodes, quads = jittered_cubed_sphere(6, seed=2)
    k = complex(2.0, -0.4)
    xs = np.array([0.2, -0.1, 0.15])
    cents, norms = [], []
    for e in nodes[quads]:
        x, n = B.quad_geom(e, 0.0, 0.0)
        cents.append(x)
        norms.append(n / np.linalg.norm(n))
    cents, norms = np.array(cents), np.array(norms)
    rv = cents - xs
    r = np.linalg.norm(rv, axis=1)
    p = np.exp(-1j * k * r) / (4 * np.pi * r)
    q = p * (-1j * k - 1 / r) * np.sum(rv * norms, axis=1) / r
    G = orc.assemble(nodes, quads, 1, 0, 0, k=k)
    H = orc.assemble(nodes, quads, 1, 0, 1, k=k)
    rel = np.linalg.norm((H - 0.5 * np.eye(len(quads))) @ p - G @ q) / np.linalg.norm(G @ q)
    print("quadrilateral transparent residual, complex k: %.2e" % rel)
    assert rel < 2e-2


def test_bm_is_sum_of_operators(orc):
    from nihu_amd.mesh import jittered_cubed_sphere
    nodes, quads = jittered_cubed_sphere(3, seed=4)
    k = complex(3.0, -0.6)
    al = -1j / k
    A, Bm = orc.assemble_bm(nodes, quads, k)
    G = orc.assemble(nodes, quads, 1, 0, 0, k)
    H = orc.assemble(nodes, quads, 1, 0, 1, k)
    Ht = orc.assemble(nodes, quads, 1, 1, 0, k)
    D = orc.assemble(nodes, quads, 1, 1, 1, k)
    I = np.eye(len(quads))
    assert np.allclose(A, al * D - 0.5 * I + H, rtol=0, atol=1e-13 * np.abs(A).max())
    assert np.allclose(Bm, G + al * Ht + al / 2 * I, rtol=0, atol=1e-13 * np.abs(Bm).max())


def test_whole_matrix_real_limit(orc):
    """Whole quadrilateral and mixed-type blocks: k + 0i equals the real path."""
    from nihu_amd.mesh import jittered_cubed_sphere, icosphere
    nodes, quads = jittered_cubed_sphere(3, seed=5)
    tn, tt = icosphere(1)
    for nx, ny in PAIRS:
        a = orc.assemble(nodes, quads, 1, nx, ny, 2.5)
        b = orc.assemble(nodes, quads, 1, nx, ny, complex(2.5, 0.0))
        assert np.abs(a - b).max() <= 1e-14 * np.abs(a).max()
        # triangles against quadrilaterals and the reverse
        a = orc.assemble(nodes, quads, 1, nx, ny, 2.5, trial=(0.5 * tn, tt))
        b = orc.assemble(nodes, quads, 1, nx, ny, complex(2.5, 0.0), trial=(0.5 * tn, tt))
        assert np.abs(a - b).max() <= 1e-14 * np.abs(a).max()
        a = orc.assemble(0.5 * tn, tt, 1, nx, ny, 2.5, trial=(nodes, quads))
        b = orc.assemble(0.5 * tn, tt, 1, nx, ny, complex(2.5, 0.0), trial=(nodes, quads))
        assert np.abs(a - b).max() <= 1e-14 * np.abs(a).max()
