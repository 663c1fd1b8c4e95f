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


# ---------------------------------------------------------------------------
# Quadrilaterals (bilinear, corners (-1,-1), (1,-1), (1,1), (-1,1))
# ---------------------------------------------------------------------------
QC = np.array([[-1.0, -1.0], [1.0, -1.0], [1.0, 1.0], [-1.0, 1.0]])


def quad_geom(coords, xi, eta):
    """Points and Jacobian vectors dx/dxi x dx/deta of a bilinear quadrilateral
    at arrays of reference coordinates."""
    xi = np.asarray(xi, dtype=float)
    eta = np.asarray(eta, dtype=float)
    N = np.stack([(1 - xi) * (1 - eta), (1 + xi) * (1 - eta), (1 + xi) * (1 + eta), (1 - xi) * (1 + eta)], -1) / 4
    Dxi = np.stack([-(1 - eta), (1 - eta), (1 + eta), -(1 + eta)], -1) / 4
    Deta = np.stack([-(1 - xi), -(1 + xi), (1 + xi), (1 - xi)], -1) / 4
    C = np.asarray(coords, dtype=float)
    return N @ C, np.cross(Dxi @ C, Deta @ C)


def np_kernel_pts(family, nx, ny, k, x, Y, nxv, NY):
    """np_kernel with one trial normal per point (rows of NY)."""
    rv = Y - x[None, :]
    r = np.linalg.norm(rv, axis=1)
    kk = 0.0 if family == 0 else k
    e = np.exp(-1j * kk * r)
    if nx == 0 and ny == 0:
        return e / (FOUR_PI * r)
    rdny = np.einsum("ij,ij->i", rv, NY) / r
    rdnx = -(rv @ nxv) / r
    if nx + ny == 1:
        g1 = e * (-1.0 - 1j * kk * r) / (FOUR_PI * r * r)
        return g1 * (rdny if ny == 1 else rdnx)
    ikr = 1j * kk * r
    g = e / (FOUR_PI * r ** 3)
    return g * (3.0 + ikr * (3.0 + ikr)) * rdnx * rdny + g * (1.0 + ikr) * (NY @ nxv)


def quad_brute(family, nx, ny, k, coords, x0, nx0, levels=4, n_gauss=8, weight=None):
    """Integral of K(x0, y) (times weight(xi, eta) if given) over a
    quadrilateral by Gauss rules on a uniform 2^levels x 2^levels subdivision
    of the reference square."""
    g, w = np.polynomial.legendre.leggauss(n_gauss)
    m = 2 ** levels
    h = 2.0 / m
    a = -1.0 + h * (np.arange(m) + 0.5)
    cx = (a[:, None] + g[None, :] * h / 2).ravel()
    cw = np.tile(w * h / 2, m)
    XI, ETA = np.meshgrid(cx, cx, indexing="ij")
    W = np.outer(cw, cw).ravel()
    Y, NJ = quad_geom(coords, XI.ravel(), ETA.ravel())
    J = np.linalg.norm(NJ, axis=1)
    K = np_kernel_pts(family, nx, ny, k, np.asarray(x0, float), Y, np.asarray(nx0, float), NJ / J[:, None])
    if weight is not None:
        K = K * weight(XI.ravel(), ETA.ravel())
    return np.sum(K * J * W)


def polygon_exact(corners, p, kind):
    """Closed forms over a planar polygon (corners: n x 2, counter-clockwise)
    at an interior point p of its plane: kind 'slp' = int 1/(4 pi r) dA,
    kind 'hsp' = finite part of int 1/(4 pi r^3) dA.  Per side, with h the
    distance from p to the side's line and phi the angle from the foot of the
    perpendicular: int rho dphi = h ln|sec + tan|, int dphi / rho = sin / h."""
    C = np.asarray(corners, float)
    p = np.asarray(p, float)
    tot = 0.0
    for i in range(len(C)):
        a, b = C[i] - p, C[(i + 1) % len(C)] - p
        t = (b - a) / np.linalg.norm(b - a)
        foot = a - t * np.dot(a, t)
        h = np.linalg.norm(foot)
        s1, s2 = np.dot(a, t) / h, np.dot(b, t) / h     # tan(phi) at the side's ends
        if kind == "slp":
            tot += h * (np.arcsinh(s2) - np.arcsinh(s1))
        else:
            tot += (s2 / np.sqrt(1 + s2 * s2) - s1 / np.sqrt(1 + s1 * s1)) / h
    return tot / FOUR_PI if kind == "slp" else -tot / FOUR_PI


def hsp_finite_part(family, k, coords, xi0, eps=(4e-3, 2e-3, 1e-3), n_theta=48, n_rho=24, weight=None):
    """Hadamard finite part of the HSP kernel over a (warped) quadrilateral at
    x0 = x(xi0), n_x = n(xi0): lim [int over the surface outside the ball
    |y - x0| < eps of K dS - 1/(2 eps)], from polar integration in the
    reference square (the ball's boundary found by Newton per direction) and
    Richardson extrapolation over the eps values (error O(eps))."""
    xi0 = np.asarray(xi0, float)
    x0, n0 = quad_geom(coords, xi0[0], xi0[1])
    nx0 = n0 / np.linalg.norm(n0)
    gt, wt = np.polynomial.legendre.leggauss(n_theta)
    gr, wr = np.polynomial.legendre.leggauss(n_rho)
    vals = []
    eps_list = tuple(eps)
    for eps in eps_list:
        tot = 0.0
        for s in range(4):
            c1, c2 = QC[s] - xi0, QC[(s + 1) % 4] - xi0
            t1, t2 = np.arctan2(c1[1], c1[0]), np.arctan2(c2[1], c2[0])
            if t2 < t1:
                t2 += 2 * np.pi
            th = (t1 + t2) / 2 + gt * (t2 - t1) / 2
            wth = wt * (t2 - t1) / 2
            d = np.stack([np.cos(th), np.sin(th)], -1)
            # distance to the side along each direction
            nrm = np.array([c2[1] - c1[1], -(c2[0] - c1[0])])
            rho_hat = np.dot(c1, nrm) / (d @ nrm)
            # ball boundary: |x(xi0 + rho d) - x0| = eps, Newton from the linear guess
            rho_e = np.full(n_theta, eps / np.linalg.norm(quad_geom(coords, xi0[0] + 1e-7, xi0[1])[0] - x0) * 1e-7)
            for it in range(30):
                P = xi0[None, :] + rho_e[:, None] * d
                Y, _ = quad_geom(coords, P[:, 0], P[:, 1])
                f = np.linalg.norm(Y - x0, axis=1) - eps
                h = 1e-9
                P2 = xi0[None, :] + (rho_e + h)[:, None] * d
                Y2, _ = quad_geom(coords, P2[:, 0], P2[:, 1])
                df = (np.linalg.norm(Y2 - x0, axis=1) - eps - f) / h
                rho_e = rho_e - f / df
            # radial integral on geometric subintervals from rho_e to rho_hat
            for i in range(n_theta):
                edges = np.geomspace(rho_e[i], rho_hat[i], 12)
                R = ((edges[:-1, None] + edges[1:, None]) / 2 + gr[None, :] * (edges[1:, None] - edges[:-1, None]) / 2).ravel()
                WR = (wr[None, :] * (edges[1:, None] - edges[:-1, None]) / 2).ravel()
                P = xi0[None, :] + R[:, None] * d[i][None, :]
                Y, NJ = quad_geom(coords, P[:, 0], P[:, 1])
                J = np.linalg.norm(NJ, axis=1)
                K = np_kernel_pts(family, 1, 1, k, x0, Y, nx0, NJ / J[:, None])
                if weight is not None:
                    K = K * weight(P[:, 0], P[:, 1])
                tot += wth[i] * np.sum(K * J * R * WR)
        w0 = 1.0 if weight is None else float(weight(np.array([xi0[0]]), np.array([xi0[1]]))[0])
        vals.append(tot - w0 / (2.0 * eps))
    # Richardson on eps, 2 eps ratios (errors O(eps) + O(eps^2))
    e = np.array(eps_list)
    A = np.stack([np.ones_like(e), e, e * e], -1)
    return np.linalg.solve(A, np.array(vals, dtype=complex))[0]


def gauss_field_shape(b):
    """The shape function of dof b of the quadrilateral gauss field, written
    from its definition (the bilinear Lagrange polynomial equal to 1 at node
    b and 0 at the other three nodes, nodes (-,-), (+,-), (-,+), (+,+) at
    +-1/sqrt(3)), as a function of (xi, eta)."""
    g = 1.0 / np.sqrt(3.0)
    nodes = np.array([[-g, -g], [g, -g], [-g, g], [g, g]])
    xb, yb = nodes[b]
    xo = [x for x in (-g, g) if x != xb][0]
    yo = [y for y in (-g, g) if y != yb][0]
    return lambda xi, eta: (np.asarray(xi) - xo) / (xb - xo) * (np.asarray(eta) - yo) / (yb - yo)


# ---------------------------------------------------------------------------
# Galerkin pairs of coplanar squares: the inner integral of 1/r over a
# rectangle at a point of its plane in closed form, the outer integral by
# Gauss rules graded toward the ends of each axis.
# ---------------------------------------------------------------------------
def _corner_antiderivative(p, q):
    """F(p, q) with d^2 F / dp dq = 1 / sqrt(p^2 + q^2):
    p ln(q + r) + q ln(p + r), each term 0 where its factor is 0."""
    p = np.asarray(p, float)
    q = np.asarray(q, float)
    r = np.hypot(p, q)

    def log_sum(a, b):
        # ln(a + r) without cancellation: a + r = b^2 / (r - a) for a < 0
        with np.errstate(divide="ignore", invalid="ignore"):
            return np.where(a >= 0.0, np.log(a + r), np.log(b * b) - np.log(r - a))

    with np.errstate(divide="ignore", invalid="ignore"):
        t1 = np.where(p == 0.0, 0.0, p * log_sum(q, p))
        t2 = np.where(q == 0.0, 0.0, q * log_sum(p, q))
    return t1 + t2


def rect_potential(x, y, u0, u1, v0, v1):
    """int_{u0}^{u1} int_{v0}^{v1} du dv / |(u, v) - (x, y)| at in-plane points."""
    F = _corner_antiderivative
    return (F(u1 - x, v1 - y) - F(u0 - x, v1 - y) - F(u1 - x, v0 - y) + F(u0 - x, v0 - y))


def graded_rule(a, b, ratio=0.2, levels=24, n=12):
    """Composite Gauss rule on [a, b] with subintervals shrinking
    geometrically toward both ends (for integrands like t ln t there)."""
    g, w = np.polynomial.legendre.leggauss(n)
    m = 0.5 * (a + b)
    cuts = [0.0] + [ratio ** j for j in range(levels, 0, -1)] + [1.0]
    X, W = [], []
    for lo, hi in zip(cuts[:-1], cuts[1:]):
        for s in (-1.0, 1.0):
            # [lo, hi] of the half length, measured from the end
            t0, t1 = (a + lo * (m - a), a + hi * (m - a)) if s < 0 else (b - hi * (b - m), b - lo * (b - m))
            X.append(0.5 * (t0 + t1) + 0.5 * (t1 - t0) * g)
            W.append(0.5 * (t1 - t0) * w)
    return np.concatenate(X), np.concatenate(W)


def coplanar_squares_slp(test, trial):
    """int_test int_trial 1 / |x - y| for axis-aligned squares
    ((x0, x1), (y0, y1)) of one plane."""
    (a0, a1), (b0, b1) = test
    X, WX = graded_rule(a0, a1)
    Y, WY = graded_rule(b0, b1)
    XX, YY = np.meshgrid(X, Y, indexing="ij")
    (u0, u1), (v0, v1) = trial
    P = rect_potential(XX, YY, u0, u1, v0, v1)
    return float(np.sum(P * np.outer(WX, WY)))
