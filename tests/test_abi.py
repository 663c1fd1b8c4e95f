"""CPU checks of the engine library: it loads, exports every symbol of
include/nbem.h, and its host-side tables and geometry agree with the oracle
(bit for bit where the classification depends on it).  No GPU calls."""

import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def nb():
    from nihu_amd import nbem
    return nbem


@pytest.fixture(scope="module")
def raw(nb):
    l = nb.lib()
    D = C.POINTER(C.c_double)
    l.nbg_dunavant.argtypes = [C.c_int, D, D, D, C.c_int]
    l.nbg_gauss_legendre.argtypes = [C.c_int, D, D]
    l.nbg_geometry.argtypes = [D, C.c_int, C.POINTER(C.c_int32), C.c_int, D, D, D, D, D, D]
    l.nbg_threshold.argtypes = [C.c_double, C.c_double]
    l.nbg_threshold.restype = C.c_double
    return l


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def test_header_symbols_exported(nb):
    hdr = open(os.path.join(ROOT, "include", "nbem.h")).read()
    declared = sorted(set(re.findall(r"\b(nb_[a-z0-9_]+)\s*\(", hdr)))
    assert set(declared) == set(nb.EXPORTS), declared
    l = nb.lib()
    for name in declared:
        assert hasattr(l, name), name


def test_version(nb):
    assert nb.lib().nb_version() >= 10000


@pytest.mark.parametrize("degree", range(15))
def test_engine_dunavant_matches_oracle(orc, raw, degree):
    xi, eta, w = (np.zeros(64) for _ in range(3))
    n = raw.nbg_dunavant(degree, _p(xi), _p(eta), _p(w), 64)
    oxi, oeta, ow = orc.dunavant(degree)
    assert n == len(ow)
    a = sorted(zip(np.round(xi[:n], 15), np.round(eta[:n], 15), np.round(w[:n], 15)))
    b = sorted(zip(np.round(oxi, 15), np.round(oeta, 15), np.round(ow, 15)))
    assert np.allclose(np.array(a), np.array(b), rtol=0, atol=2e-15)


@pytest.mark.parametrize("n", [6, 21])
def test_engine_gauss_legendre(raw, n):
    x, w = np.zeros(n), np.zeros(n)
    raw.nbg_gauss_legendre(n, _p(x), _p(w))
    xr, wr = np.polynomial.legendre.leggauss(n)
    assert np.abs(x - xr).max() < 1e-15 and np.abs(w - wr).max() < 1e-14


def test_engine_geometry_bitwise_equals_oracle(orc, raw):
    from nihu_amd.mesh import jittered_icosphere
    nodes, tris = jittered_icosphere(3, seed=7)
    n = len(tris)
    c, nu, J, R, diam, X = (np.zeros(s) for s in (3 * n, 3 * n, n, n, n, 9 * n))
    t32 = np.ascontiguousarray(tris, dtype=np.int32)
    st = raw.nbg_geometry(_p(nodes), len(nodes), t32.ctypes.data_as(C.POINTER(C.c_int32)), n,
                          _p(c), _p(nu), _p(J), _p(R), _p(diam), _p(X))
    assert st == 0
    g = orc.geometry(nodes, tris)
    assert np.array_equal(c.reshape(3, n).T, g["center"])
    assert np.array_equal(R, g["R"]) and np.array_equal(diam, g["diam"]) and np.array_equal(J, g["jac"])
    assert np.allclose(nu.reshape(3, n).T, g["normal"], rtol=0, atol=1e-15)


def test_engine_quad_geometry_bitwise_equals_oracle(orc, raw):
    """Centres, mean edges and diameters of quadrilaterals round exactly as the
    oracle's (the pair classification compares against thresholds built on them)."""
    from nihu_amd.mesh import jittered_cubed_sphere
    nodes, quads = jittered_cubed_sphere(7, seed=3)
    n = len(quads)
    c, nu, J, R, diam, X = (np.zeros(s) for s in (3 * n, 3 * n, n, n, n, 12 * n))
    q32 = np.ascontiguousarray(quads, dtype=np.int32)
    raw.nbg_geometry_quad.argtypes = [C.POINTER(C.c_double), C.c_int, C.POINTER(C.c_int32), C.c_int] + \
        [C.POINTER(C.c_double)] * 6
    st = raw.nbg_geometry_quad(_p(nodes), len(nodes), q32.ctypes.data_as(C.POINTER(C.c_int32)), n,
                               _p(c), _p(nu), _p(J), _p(R), _p(diam), _p(X))
    assert st == 0
    g = orc.geometry_quad(nodes, quads)
    assert np.array_equal(c.reshape(3, n).T, g["center"])
    assert np.array_equal(R, g["R"]) and np.array_equal(diam, g["diam"])
    assert np.allclose(nu.reshape(3, n).T, g["normal"], rtol=0, atol=1e-15)


def test_threshold_reproduces_sqrt_division_test(raw):
    """d2 < threshold(R, x)  <=>  sqrt(d2) / R < x, checked at and around the
    threshold (where a rounding slip would show)."""
    rng = np.random.default_rng(0)
    for _ in range(300):
        R = rng.uniform(0.01, 0.2)
        x = rng.choice([1.1, 1.5, 2.0, 9.1, 22.3, 11. / 10., 158. / 10.])
        t = raw.nbg_threshold(R, x)
        for d2 in [t, np.nextafter(t, 0), np.nextafter(t, 1), np.nextafter(np.nextafter(t, 0), 0), t * (1 + 1e-12)]:
            assert (d2 < t) == (np.sqrt(d2) / R < x)


def test_bad_mesh_rejected_without_gpu(nb):
    with pytest.raises(ValueError):
        nb.Mesh(np.zeros((3, 3)), np.array([[0, 1, 5]]))


def test_empty_and_degenerate_meshes_rejected_without_gpu(nb):
    """Empty element lists, a zero-area triangle, a degenerate quadrilateral and
    a node index past the end are refused by nb_mesh_create before any device
    work (nihu's mesh would accept them and fail later, or divide by zero)."""
    nodes = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [1.0, 1.0, 0.0]])
    with pytest.raises(ValueError, match="empty mesh"):
        nb.Mesh(nodes, np.zeros((0, 3), dtype=np.int32))
    with pytest.raises(ValueError, match="empty mesh"):
        nb.Mesh(nodes, np.zeros((0, 4), dtype=np.int32))
    with pytest.raises(ValueError):
        nb.Mesh(nodes, np.array([[0, 1, 1]]))            # zero area
    with pytest.raises(ValueError):
        nb.Mesh(nodes, np.array([[0, 1, 1, 0]]))         # collapsed quadrilateral
    with pytest.raises(ValueError):
        nb.Mesh(nodes, np.array([[0, 1, -1]]))           # negative node index
    with pytest.raises(ValueError):
        nb.Mesh(nodes, np.ones((2, 5), dtype=np.int32))   # neither triangles nor quadrilaterals
    # NiHu element tables: unknown type ids, and a degenerate element in a table of one type
    with pytest.raises(ValueError, match="type id"):
        nb.Mesh.from_elements(nodes, np.array([[32303, 0, 1, 2, 0], [32505, 0, 1, 2, 3]]))
    with pytest.raises(ValueError, match="empty"):
        nb.Mesh.from_elements(nodes, np.zeros((0, 5), dtype=np.int32))
    with pytest.raises(ValueError):
        nb.Mesh.from_elements(nodes, np.array([[32303, 0, 1, 1]]))
    raw = np.array([[32303, 0, 1, 2, 0], [32505, 0, 1, 2, 3]], dtype=np.int32)
    h = __import__("ctypes").c_void_p()
    assert nb.lib().nb_mesh_create_elements(nodes.ctypes.data, 4, raw.ctypes.data, 2,
                                            __import__("ctypes").byref(h)) == nb.NB_ERR_ARG
    assert "type id" in nb.lib().nb_last_error().decode()


def test_mixed_table_helpers():
    """The mixed cubed sphere is closed and consistently oriented (every
    directed edge once, its reverse once) and round-trips through OFF."""
    import tempfile
    from nihu_amd.mesh import mixed_cubed_sphere, table_faces, write_off, read_off_elements
    nodes, table = mixed_cubed_sphere(4, seed=1)
    tri = table[:, 0] == 32303
    assert tri.sum() == 2 * ((6 * 16 + 2) // 3) and (~tri).sum() == 6 * 16 - (6 * 16 + 2) // 3
    edges = {}
    for f in table_faces(table):
        for a in range(len(f)):
            e = (f[a], f[(a + 1) % len(f)])
            edges[e] = edges.get(e, 0) + 1
    assert all(v == 1 and edges.get((b, a)) == 1 for (a, b), v in edges.items())
    with tempfile.TemporaryDirectory() as d:
        p = d + "/m.off"
        write_off(p, nodes, table_faces(table))
        n2, t2 = read_off_elements(p)
    assert np.array_equal(n2, nodes) and np.array_equal(t2, table)
