"""Independent numpy computations used by the tests to pin the oracle.

These are brute-force and closed-form evaluations (subdivided quadrature,
polar integration around the singular point, Legendre roots) -- different
algorithms from the semi-analytic ones the oracle restates, so a dropped term,
a wrong sign or a transposed operand in either shows up as a mismatch.
"""

from __future__ import annotations

import numpy as np

FOUR_PI = 4.0 * np.pi


def np_kernel(family, nx, ny, k, x, Y, nxv, nyv):
    """Vectorised 3D Laplace / Helmholtz kernel at trial points Y (n x 3), r = y - x."""
    rv = Y - x[None, :]
    r = np.linalg.norm(rv, axis=1)
    kk = 0.0 if family == 0 else k
    e = np.exp(-1j * kk * r)
    if nx == 0 and ny == 0:
        return e / (FOUR_PI * r)
    rdny = rv @ nyv / r
    rdnx = -(rv @ nxv) / r
    if nx + ny == 1:
        g1 = e * (-1.0 - 1j * kk * r) / (FOUR_PI * r * r)
        return g1 * (rdny if ny == 1 else rdnx)
    ikr = 1j * kk * r
    g = e / (FOUR_PI * r ** 3)
    return g * (3.0 + ikr * (3.0 + ikr)) * rdnx * rdny + g * (1.0 + ikr) * np.dot(nxv, nyv)


def subdivided_rule(coords, base_xi, base_eta, base_w, levels):
    """Physical points and weights (incl. Jacobian) of a base triangle rule
    applied on 4**levels congruent sub-triangles of the flat triangle."""
    tris = [np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]])]
    for _ in range(levels):
        nxt = []
        for t in tris:
            a, b, c = t
            ab, bc, ca = (a + b) / 2, (b + c) / 2, (c + a) / 2
            nxt += [np.array([a, ab, ca]), np.array([ab, b, bc]), np.array([ca, bc, c]), np.array([ab, bc, ca])]
        tris = nxt
    X0, X1, X2 = coords
    J = np.linalg.norm(np.cross(X1 - X0, X2 - X0))
    pts, wts = [], []
    for t in tris:
        a, b, c = t
        area_ratio = abs((b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0]))  # vs the reference triangle
        xi = a[0] + (b[0] - a[0]) * base_xi + (c[0] - a[0]) * base_eta
        eta = a[1] + (b[1] - a[1]) * base_xi + (c[1] - a[1]) * base_eta
        pts.append(np.outer(1 - xi - eta, X0) + np.outer(xi, X1) + np.outer(eta, X2))
        wts.append(base_w * area_ratio * J)
    return np.vstack(pts), np.concatenate(wts)


def polar_sides(coords, x0):
    """In-plane polar description of a flat triangle seen from the projection of
    x0: per side (theta1, theta2, distance to side line, foot angle), and the
    height z of x0 above the plane (normal (X1-X0)x(X2-X0))."""
    X0, X1, X2 = coords
    n = np.cross(X1 - X0, X2 - X0)
    n = n / np.linalg.norm(n)
    t1 = (X1 - X0) / np.linalg.norm(X1 - X0)
    t2 = np.cross(n, t1)
    z = np.dot(x0 - X0, n)
    p = x0 - z * n
    loc = [np.array([np.dot(Xc - p, t1), np.dot(Xc - p, t2)]) for Xc in coords]
    sides = []
    for s in range(3):
        a, b = loc[s], loc[(s + 1) % 3]
        th1, th2 = np.arctan2(a[1], a[0]), np.arctan2(b[1], b[0])
        dth = np.angle(np.exp(1j * (th2 - th1)))  # signed turn in (-pi, pi]
        l = (b - a) / np.linalg.norm(b - a)
        foot = a - l * np.dot(a, l)
        sides.append((th1, th1 + dth, np.linalg.norm(foot), np.arctan2(foot[1], foot[0])))
    return sides, z, n


def polar_integral(coords, x0, radial, n_gauss=64):
    """sum over sides of int dtheta radial(P(theta)) with P the in-plane
    distance from the projected x0 to the side along the ray theta."""
    sides, z, _ = polar_sides(coords, x0)
    xg, wg = np.polynomial.legendre.leggauss(n_gauss)
    tot = 0.0
    for th1, th2, d, th0 in sides:
        th = th1 + (th2 - th1) * (xg + 1) / 2
        P = d / np.cos(th - th0)
        tot = tot + np.sum(radial(P, th) * wg) * (th2 - th1) / 2
    return tot, z


def polar_area_rule(coords, x0, n_theta=48, n_rho=48):
    """Points and weights of a polar rule centred at x0 (x0 in the plane of the
    flat triangle): weights carry rho, so integrands O(1/rho) are integrated
    to high accuracy."""
    sides, z, nrm = polar_sides(coords, x0)
    X0, X1, _ = coords
    t1 = (X1 - X0) / np.linalg.norm(X1 - X0)
    t2 = np.cross(nrm, t1)
    xg, wg = np.polynomial.legendre.leggauss(n_theta)
    xr, wr = np.polynomial.legendre.leggauss(n_rho)
    pts, wts = [], []
    for th1, th2, d, th0 in sides:
        th = th1 + (th2 - th1) * (xg + 1) / 2
        wth = wg * (th2 - th1) / 2
        P = d / np.cos(th - th0)
        for t, w, p in zip(th, wth, P):
            rho = p * (xr + 1) / 2
            ww = w * wr * p / 2 * rho
            pts.append(x0[None, :] + np.outer(rho * np.cos(t), t1) + np.outer(rho * np.sin(t), t2))
            wts.append(ww)
    return np.vstack(pts), np.concatenate(wts)


def dunavant_monomial_exact(a, b):
    """int over the reference triangle of xi^a eta^b = a! b! / (a+b+2)!"""
    from math import factorial
    return factorial(a) * factorial(b) / factorial(a + b + 2)


def shape_values(coords, Y):
    """Linear shape functions (1 - xi - eta, xi, eta) of a flat triangle at
    points Y lying in its plane (least-squares inverse map)."""
    X0, X1, X2 = coords
    M = np.array([X1 - X0, X2 - X0]).T
    sol, *_ = np.linalg.lstsq(M, (Y - X0).T, rcond=None)
    xi, eta = sol
    return np.stack([1 - xi - eta, xi, eta], axis=1)


def polar_rule_projected(coords, x, n_theta=24, n_rho=24):
    """polar_area_rule centred at the orthogonal projection of x onto the
    triangle's plane (x may lie off the plane)."""
    X0, X1, X2 = coords
    nrm = np.cross(X1 - X0, X2 - X0)
    nrm = nrm / np.linalg.norm(nrm)
    p = x - np.dot(x - X0, nrm) * nrm
    return polar_area_rule(coords, p, n_theta, n_rho)


def galerkin_pair_brute(family, nx, ny, k, tc, sc, outer_points, outer_weights, n_theta=24, n_rho=24):
    """3x3 block int_T N_a(x) int_T' K(x, y) N_b(y) dy dx: outer rule given,
    inner polar rule around the projection of each outer point."""
    nxv = np.cross(tc[1] - tc[0], tc[2] - tc[0])
    nxv = nxv / np.linalg.norm(nxv)
    nyv = np.cross(sc[1] - sc[0], sc[2] - sc[0])
    nyv = nyv / np.linalg.norm(nyv)
    Nt = shape_values(tc, outer_points)
    blk = np.zeros((3, 3), complex)
    for x, wx, nt in zip(outer_points, outer_weights, Nt):
        Y, Wy = polar_rule_projected(sc, x, n_theta, n_rho)
        K = np_kernel(family, nx, ny, k, x, Y, nxv, nyv)
        blk += wx * np.outer(nt, (K * Wy) @ shape_values(sc, Y))
    return blk


def solid_angle(coords, X):
    """Signed solid angle of a flat triangle seen from points X (Van Oosterom and
    Strackee, 1983): int (y - x).n / |y - x|^3 dS_y, n = (X1-X0)x(X2-X0)/|.|."""
    R = [c[None, :] - X for c in coords]
    r = [np.linalg.norm(v, axis=1) for v in R]
    num = np.einsum("ij,ij->i", R[0], np.cross(R[1], R[2]))
    den = (r[0] * r[1] * r[2] + np.einsum("ij,ij->i", R[0], R[1]) * r[2]
           + np.einsum("ij,ij->i", R[0], R[2]) * r[1] + np.einsum("ij,ij->i", R[1], R[2]) * r[0])
    return 2.0 * np.arctan2(num, den)


def square_mesh(n):
    """[-1, 1]^2 in the z = 0 plane, n x n cells, two triangles each."""
    g = np.linspace(-1.0, 1.0, n + 1)
    X, Y = np.meshgrid(g, g, indexing="ij")
    nodes = np.stack([X.ravel(), Y.ravel(), np.zeros(X.size)], 1)
    tris = []
    for i in range(n):
        for j in range(n):
            a = i * (n + 1) + j
            b, d = a + n + 1, a + 1
            tris += [[a, b, b + 1], [a, b + 1, d]]
    return nodes, np.array(tris, dtype=np.int32)
