"""Pins of the oracle's complex-wavenumber path (CPU only; NiHu's Helmholtz
kernel takes a complex WaveNumber, helmholtz_kernel.hpp:103, and its tests
hold no complex-wavenumber values).  Every check is against something other
than the oracle's own formula: the real-wavenumber path (pinned by NiHu's
golden values in test_oracle.py) continued analytically (Cauchy-Riemann in k),
finite differences of the kernel, brute-force subdivided quadrature of near
pairs, closed-form polar integrals of the diagonal terms, and the exact
answers of the transparent point-source and pulsating-sphere problems, which
hold for complex k as for real k."""

import numpy as np
import pytest

from tests import bemref

KC = complex(3.0, -0.4)
PAIRS = [(0, 0), (0, 1), (1, 0), (1, 1)]


def _geom(seed=7):
    rng = np.random.default_rng(seed)
    x, y = rng.normal(size=3), rng.normal(size=3)
    nx, ny = [v / np.linalg.norm(v) for v in rng.normal(size=(2, 3))]
    return x, y, nx, ny


@pytest.mark.parametrize("nxy", PAIRS)
def test_real_limit_and_analytic_in_k(orc, nxy):
    """k + 0i gives the real-wavenumber kernel; at a complex k the derivative
    along Im k equals i times the derivative along Re k (the kernel is the
    analytic continuation of the real one, not another function of k)."""
    x, y, nx, ny = _geom()
    a, b = nxy
    K = lambda k: orc.kernel(1, a, b, k, x, y, nx, ny)
    real = K(3.0)
    assert abs(K(complex(3.0, 0.0)) - real) <= 1e-15 * abs(real)
    h = 1e-5
    d_re = (K(KC + h) - K(KC - h)) / (2 * h)
    d_im = (K(KC + 1j * h) - K(KC - 1j * h)) / (2 * h)
    assert abs(d_im - 1j * d_re) <= 1e-7 * abs(d_re), (d_im, d_re)


def test_kernel_finite_differences(orc):
    """DLP, DLPt and HSP as normal derivatives of the SLP kernel at a complex k
    (central differences)."""
    x, y, nx, ny = _geom(11)
    G = lambda p, q: orc.kernel(1, 0, 0, KC, p, q, nx, ny)
    e = 1e-5
    gy = (G(x, y + e * ny) - G(x, y - e * ny)) / (2 * e)
    gx = (G(x + e * nx, y) - G(x - e * nx, y)) / (2 * e)
    e2 = 1e-4
    gxy = (G(x + e2 * nx, y + e2 * ny) - G(x + e2 * nx, y - e2 * ny) - G(x - e2 * nx, y + e2 * ny)
           + G(x - e2 * nx, y - e2 * ny)) / (4 * e2 * e2)
    for num, (a, b) in ((gy, (0, 1)), (gx, (1, 0)), (gxy, (1, 1))):
        val = orc.kernel(1, a, b, KC, x, y, nx, ny)
        assert abs(num - val) <= 1e-6 * abs(val), (a, b, num, val)


@pytest.mark.parametrize("nxy", PAIRS)
@pytest.mark.parametrize("height", [0.15, -0.08])
def test_near_integral_brute_force(orc, nxy, height):
    """As the real-wavenumber test: the static part exact, the imaginary part
    of the smooth dynamic integrand to 1e-8, the real part bounded by
    2e-3 |k L|^2 (NiHu's 42-point dynamic rule on the cusp of K_helm - K_lap)."""
    coords = np.array([[0.1, -0.2, 0.05], [0.9, 0.1, -0.1], [0.2, 0.8, 0.2]])
    n = np.cross(coords[1] - coords[0], coords[2] - coords[0])
    n /= np.linalg.norm(n)
    x0 = coords.mean(axis=0) + np.array([0.05, -0.03, 0.0]) + height * n
    nxv = n + np.array([0.3, -0.2, 0.1])
    nxv /= np.linalg.norm(nxv)
    a, b = nxy
    k = complex(2.0, -0.5)
    xi, eta, w = orc.dunavant(14)
    Y, W = bemref.subdivided_rule(coords, xi, eta, w, 5)
    bf = np.sum(bemref.np_kernel(1, a, b, k, x0, Y, nxv, n) * W)
    val = orc.elem_integral(1, 1, a, b, k, coords, x0=x0, nx0=nxv, degree=120)
    assert abs(val - bf) <= 2e-3 * abs(k) ** 2 * abs(bf), (val, bf)
    # the error is small against what the imaginary part of k changes
    bf_real = np.sum(bemref.np_kernel(1, a, b, k.real, x0, Y, nxv, n) * W)
    assert abs(val - bf) < 0.1 * abs(bf - bf_real), (val, bf, bf_real)


DIAG = np.array([[0.0, 0.0, 0.0], [0.09, 0.01, 0.0], [0.02, 0.08, 0.0]])


@pytest.mark.parametrize("k", [complex(1.0, -0.5), complex(10.0, -2.0), complex(25.0, 3.0)])
def test_diagonal_closed_form(orc, k):
    """Polar closed forms at the centroid, continued to complex k:
    SLP = 1/4pi sum int (1 - e^{-ikP})/(ik) dtheta,
    HSP = 1/4pi [ -int e^{-ikP}/P dtheta - 2 pi i k ]; bounds as for real k
    with |k| R."""
    q, _ = np.linalg.qr(np.random.default_rng(1).normal(size=(3, 3)))
    coords = DIAG @ q.T + np.array([0.3, -0.1, 0.7])
    c = coords.mean(axis=0)
    slp_exact, _ = bemref.polar_integral(coords, c, lambda P, th: (1 - np.exp(-1j * k * P)) / (1j * k), 96)
    hsp_int, _ = bemref.polar_integral(coords, c, lambda P, th: -np.exp(-1j * k * P) / P, 96)
    slp_exact /= 4 * np.pi
    hsp_exact = (hsp_int - 2j * np.pi * k) / (4 * np.pi)
    kR = abs(k) * 0.09
    for a, ref in ((0, slp_exact), (1, hsp_exact)):
        val = orc.elem_integral(2, 1, a, a, k, coords)
        assert abs(val - ref) <= max(3e-3 * kR ** 2, 1e-12) * abs(ref), (a, val, ref)


def _transparent_error(orc, level, k):
    """Sources inside the surface give exact p0 = sum G(x0, y), q0 = dp0/dn;
    solve (M - I/2) p = L q0 (nihu/tutorial/helmholtz_3d_transparent_standalone.cpp:17-72,
    here with a complex wavenumber)."""
    from nihu_amd.mesh import jittered_icosphere
    nodes, tris = jittered_icosphere(level, seed=1)
    g = orc.geometry(nodes, tris)
    y, ny = g["center"], g["normal"]
    x0 = np.array([[0.2, 0.3, 0.1], [0.3, 0.2, 0.2]])
    p0 = np.zeros(len(tris), complex)
    q0 = np.zeros(len(tris), complex)
    for s, amp in enumerate([1.0, -1.0]):
        rv = y - x0[s]
        r = np.linalg.norm(rv, axis=1)
        ikr = 1j * k * r
        rdn = np.sum(rv * ny, axis=1) / r
        p0 += amp * np.exp(-ikr) / r / (4 * np.pi)
        q0 += amp * -(np.exp(-ikr) / r / r) * rdn / (4 * np.pi) * (1 + ikr)
    L = orc.assemble(nodes, tris, 1, 0, 0, k)
    M = orc.assemble(nodes, tris, 1, 0, 1, k)
    p = np.linalg.solve(M - 0.5 * np.eye(len(tris)), L @ q0)
    return np.linalg.norm(p - p0) / np.linalg.norm(p0)


def test_transparent_known_answer_converges(orc):
    k = complex(1.5, -0.3)
    e2 = _transparent_error(orc, 2, k)
    e3 = _transparent_error(orc, 3, k)
    assert e3 < 0.05 and e3 < 0.6 * e2, (e2, e3)


def test_pulsating_sphere(orc):
    """Analytic p = -1/(1 + ik) q for the pulsating unit sphere
    (nihu/test/stallone_unit/helmholtz_3d_pulsating_sphere_test.cpp:28-73),
    with a complex wavenumber."""
    from nihu_amd.mesh import icosphere
    nodes, tris = icosphere(3)
    k = complex(1.0, -0.3)
    L = orc.assemble(nodes, tris, 1, 0, 0, k)
    M = orc.assemble(nodes, tris, 1, 0, 1, k)
    p = np.linalg.solve(M - 0.5 * np.eye(len(tris)), L @ np.ones(len(tris)))
    panal = -1.0 / (1.0 + 1j * k)
    assert np.abs(p - panal).max() / abs(panal) < 0.02


def test_bm_is_sum_of_operators(orc):
    from nihu_amd.mesh import jittered_icosphere
    nodes, tris = jittered_icosphere(2, seed=4)
    k = complex(3.0, -0.6)
    al = -1j / k
    A, B = orc.assemble_bm(nodes, tris, k)
    G = orc.assemble(nodes, tris, 1, 0, 0, k)
    H = orc.assemble(nodes, tris, 1, 0, 1, k)
    Ht = orc.assemble(nodes, tris, 1, 1, 0, k)
    D = orc.assemble(nodes, tris, 1, 1, 1, k)
    I = np.eye(len(tris))
    assert np.allclose(A, al * D - 0.5 * I + H, rtol=0, atol=1e-13 * np.abs(A).max())
    assert np.allclose(B, G + al * Ht + al / 2 * I, rtol=0, atol=1e-13 * np.abs(B).max())


@pytest.mark.parametrize("linear", [False, True])
def test_galerkin_pulsating_sphere(orc, linear):
    """The pulsating-sphere answer in Galerkin form (constant and
    isoparametric linear fields) at a complex wavenumber:
    (M - M_I/2) p = L q with the mass matrix M_I, p = -1/(1 + ik) for q = 1."""
    from nihu_amd.mesh import icosphere
    nodes, tris = icosphere(2)
    k = complex(1.0, -0.3)
    L = orc.assemble_galerkin(nodes, tris, 1, 0, 0, k, linear=linear)
    M = orc.assemble_galerkin(nodes, tris, 1, 0, 1, k, linear=linear)
    I = orc.assemble_galerkin(nodes, tris, 2, 0, 0, 0.0, linear=linear)
    p = np.linalg.solve(M - 0.5 * I, L @ np.ones(L.shape[1]))
    panal = -1.0 / (1.0 + 1j * k)
    err = np.abs(p - panal).max() / abs(panal)
    print("Galerkin pulsating sphere (%s), complex k: %.2e" % ("linear" if linear else "constant", err))
    assert err < 0.03


def test_galerkin_real_limit(orc):
    """k + 0i through the complex path equals the real-wavenumber Galerkin
    matrices (the same quadrature, the real kernel)."""
    from nihu_amd.mesh import jittered_icosphere
    nodes, tris = jittered_icosphere(1, seed=2)
    for nx, ny in ((0, 0), (0, 1), (1, 0)):
        a = orc.assemble_galerkin(nodes, tris, 1, nx, ny, 2.5)
        b = orc.assemble_galerkin(nodes, tris, 1, nx, ny, complex(2.5, 0.0))
        assert np.abs(a - b).max() <= 1e-14 * np.abs(a).max()
