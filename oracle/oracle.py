"""ctypes front end of the FP64 oracle (liboracle.so built from nbem_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg, as the checker or the timed CPU baseline.  The
engine in nihu_amd/ never imports it.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_LIB_V4 = os.path.join(_HERE, "liboracle_v4.so")
_lib = None
_loaded = {}
# x86-64-v4: the AVX-512 subsets the tuned build may use
_V4_FLAGS = {"avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl"}


def tuned_supported():
    """True if the tuned build exists and this host's CPU reports the
    x86-64-v4 features it was compiled for (read from /proc/cpuinfo)."""
    if not os.path.exists(_LIB_V4):
        return False
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("flags"):
                return _V4_FLAGS <= set(line.split(":", 1)[1].split())
    except OSError:
        pass
    return False


def use_build(name="exact"):
    """Switch the functions of this module between the checker build
    ("exact": liboracle.so, no FMA contraction) and the tuned build
    ("tuned": liboracle_v4.so, AVX-512 and FMA contraction; the timed CPU
    baseline).  Returns the previous name."""
    global _lib
    prev = "tuned" if _lib is not None and _lib is _loaded.get(_LIB_V4) else "exact"
    path = _LIB_V4 if name == "tuned" else _LIB
    if name == "tuned" and not tuned_supported():
        raise RuntimeError("the tuned oracle build is missing or this CPU lacks x86-64-v4")
    if path not in _loaded:
        _loaded[path] = _declare(C.CDLL(path))
    _lib = _loaded[path]
    return prev


class Policy(C.Structure):
    _fields_ = [("rule", C.c_int), ("fixed_degree", C.c_int), ("near_mode", C.c_int), ("near_limit", C.c_double)]


NIHU = dict(rule=0, fixed_degree=0, near_mode=0, near_limit=1.5)


def policy(rule=0, fixed_degree=0, near_mode=0, near_limit=1.5):
    return Policy(rule, fixed_degree, near_mode, near_limit)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        _lib = _loaded.setdefault(_LIB, _declare(C.CDLL(_LIB)))
    return _lib


def _declare(l):
    if True:
        D = C.POINTER(C.c_double)
        I = C.POINTER(C.c_int)
        P = C.POINTER(Policy)
        l.orc_last_error.restype = C.c_char_p
        l.orc_dunavant.argtypes = [C.c_int, D, D, D, C.c_int, I]
        l.orc_gauss_legendre.argtypes = [C.c_int, D, D]
        l.orc_geometry.argtypes = [D, C.c_int, I, C.c_int, D, D, D, D, D]
        l.orc_kernel.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, D, D, D, D, D]
        l.orc_elem_integral.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, D, D, D, C.c_int, D]
        l.orc_kernel_ck.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, D, D, D, D, D]
        l.orc_elem_integral_ck.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, D, D, D,
                                           C.c_int, D]
        l.orc_assemble_ck.argtypes = [D, C.c_int, I, C.c_int, D, C.c_int, I, C.c_int, C.c_int,
                                      C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                      P, I, C.c_int, D, C.c_int64, C.c_int, C.c_int]
        l.orc_assemble_bm_ck.argtypes = [D, C.c_int, I, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                         P, I, C.c_int, D, D, C.c_int64, C.c_int]
        l.orc_classify.argtypes = [D, C.c_int, I, C.c_int, D, C.c_int, I, C.c_int, C.c_int, C.c_int, C.c_int,
                                   P, I, C.c_int, C.POINTER(C.c_int8)]
        l.orc_assemble.argtypes = [D, C.c_int, I, C.c_int, D, C.c_int, I, C.c_int, C.c_int,
                                   C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                   P, I, C.c_int, D, C.c_int64, C.c_int, C.c_int]
        l.orc_identity.argtypes = [C.c_int, C.c_double, C.c_double, I, C.c_int, D, C.c_int64]
        l.orc_assemble_bm.argtypes = [D, C.c_int, I, C.c_int, C.c_double, C.c_double, C.c_double,
                                      P, I, C.c_int, D, D, C.c_int64, C.c_int]
        l.orc_sweep_bm.argtypes = [D, C.c_int, I, C.c_int, D, C.c_int, P, I, C.c_int, D, D, C.c_int]
        l.orc_assemble_galerkin.argtypes = [D, C.c_int, I, C.c_int, D, C.c_int, I, C.c_int, C.c_int, C.c_int,
                                            C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                            P, I, C.c_int, D, C.c_int64, C.c_int, C.c_int]
        l.orc_assemble_galerkin_ck.argtypes = [D, C.c_int, I, C.c_int, D, C.c_int, I, C.c_int, C.c_int, C.c_int,
                                               C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                               C.c_double, P, I, C.c_int, D, C.c_int64, C.c_int, C.c_int]
        l.orc_overlap.argtypes = [I, I, I, I]
        l.orc_assemble_galerkin_quad.argtypes = [D, C.c_int, I, C.c_int, D, C.c_int, I, C.c_int, C.c_int, C.c_int,
                                                 C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                                 C.c_double, P, I, C.c_int, D, C.c_int64, C.c_int, C.c_int]
        l.orc_overlap_quads.argtypes = [I, I, I, I]
        l.orc_quad_singular_rule.argtypes = [C.c_int, C.c_int, D, C.c_int, I]
        l.orc_galerkin_quad_block.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, P, D, I, D, I,
                                              C.c_int, C.c_int, C.c_int, D]
        l.orc_singular_rule.argtypes = [C.c_int, D, C.c_int, I]
        l.orc_galerkin_block.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, P, D, I, D, I, C.c_int, C.c_int,
                                         C.c_int, D]
        l.orc_assemble_elems.argtypes = [D, C.c_int, I, C.c_int, C.c_int, D, C.c_int, I, C.c_int, C.c_int, C.c_int,
                                         C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                         P, I, C.c_int, D, C.c_int64, C.c_int, C.c_int]
        l.orc_assemble_bm_elems.argtypes = [D, C.c_int, I, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                            P, I, C.c_int, D, D, C.c_int64, C.c_int]
        l.orc_quad_elem.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, D, D, D, D, C.c_int, C.c_int, D]
        l.orc_assemble_elems_ck.argtypes = [D, C.c_int, I, C.c_int, C.c_int, D, C.c_int, I, C.c_int, C.c_int,
                                            C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                            C.c_double, P, I, C.c_int, D, C.c_int64, C.c_int, C.c_int]
        l.orc_assemble_bm_elems_ck.argtypes = [D, C.c_int, I, C.c_int, C.c_int, C.c_double, C.c_double,
                                               C.c_double, C.c_double, P, I, C.c_int, D, D, C.c_int64, C.c_int]
        l.orc_quad_elem_ck.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, D, D, D, D,
                                       C.c_int, C.c_int, D]
        l.orc_assemble_gauss.argtypes = [D, C.c_int, I, C.c_int, C.c_int, D, C.c_int, I, C.c_int, C.c_int,
                                         C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                         C.c_double, P, I, C.c_int, D, C.c_int64, C.c_int, C.c_int]
        l.orc_assemble_bm_gauss.argtypes = [D, C.c_int, I, C.c_int, C.c_double, C.c_double, C.c_double,
                                            C.c_double, P, I, C.c_int, D, D, C.c_int64, C.c_int]
        l.orc_gauss_elem.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, D, D, D, D,
                                     C.c_int, C.c_int, D]
        l.orc_gauss_shape.argtypes = [C.c_double, C.c_double, D, D]
        l.orc_quad_duffy_rule.argtypes = [C.c_int, C.c_double, C.c_double, D, C.c_int, I]
        l.orc_telles_rule.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, D, C.c_int, I]
        l.orc_geometry_quad.argtypes = [D, C.c_int, I, C.c_int, D, D, D, D]
    return l


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _i(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_int))


def _check(status):
    if status != 0:
        msg = lib().orc_last_error().decode()
        if status == 2:
            raise IndexError(msg)
        raise RuntimeError(msg)


def _mesh(nodes, tris):
    return (np.ascontiguousarray(nodes, dtype=np.float64), np.ascontiguousarray(tris, dtype=np.int32))


def dunavant(degree):
    xi, eta, w = (np.zeros(64) for _ in range(3))
    n = C.c_int(0)
    _check(lib().orc_dunavant(degree, _d(xi), _d(eta), _d(w), 64, C.byref(n)))
    return xi[:n.value].copy(), eta[:n.value].copy(), w[:n.value].copy()


def gauss_legendre(n):
    x, w = np.zeros(n), np.zeros(n)
    lib().orc_gauss_legendre(n, _d(x), _d(w))
    return x, w


def geometry(nodes, tris):
    nodes, tris = _mesh(nodes, tris)
    n = len(tris)
    c, nu, J, R, diam = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(n), np.zeros(n), np.zeros(n)
    _check(lib().orc_geometry(_d(nodes), len(nodes), _i(tris), n, _d(c), _d(nu), _d(J), _d(R), _d(diam)))
    return dict(center=c, normal=nu, jac=J, R=R, diam=diam)


def _is_complex(k):
    return isinstance(k, (complex, np.complexfloating))


def kernel(family, nx, ny, k, x, y, nxv, nyv):
    """k real or complex (a complex wavenumber)."""
    out = np.zeros(2)
    arr = [np.ascontiguousarray(v, dtype=np.float64) for v in (x, y, nxv, nyv)]
    if _is_complex(k):
        _check(lib().orc_kernel_ck(family, nx, ny, k.real, k.imag, *[_d(a) for a in arr], _d(out)))
    else:
        _check(lib().orc_kernel(family, nx, ny, k, *[_d(a) for a in arr], _d(out)))
    return complex(out[0], out[1])


def elem_integral(kind, family, nx, ny, k, coords, x0=(0, 0, 0), nx0=(0, 0, 1), degree=0):
    """kind 0 regular (Dunavant degree), 1 near-singular, 2 singular (x0 = centre).
    coords: 3 x 3, one vertex per row."""
    out = np.zeros(2)
    cs = np.ascontiguousarray(coords, dtype=np.float64).reshape(9)
    x0 = np.ascontiguousarray(x0, dtype=np.float64)
    nx0 = np.ascontiguousarray(nx0, dtype=np.float64)
    if _is_complex(k):
        _check(lib().orc_elem_integral_ck(kind, family, nx, ny, k.real, k.imag, _d(cs), _d(x0), _d(nx0), degree,
                                          _d(out)))
    else:
        _check(lib().orc_elem_integral(kind, family, nx, ny, k, _d(cs), _d(x0), _d(nx0), degree, _d(out)))
    return complex(out[0], out[1])


def classify(nodes, tris, nx, ny, pol=None, rows=None, trial=None):
    """Pair classes (-1 diagonal, -2 near, else regular degree), [n_rows, n_trial] int8."""
    pol = pol or policy()
    t_nodes, t_tris = _mesh(nodes, tris)
    same = trial is None
    s_nodes, s_tris = (t_nodes, t_tris) if same else _mesh(*trial)
    rows_a = None if rows is None else np.ascontiguousarray(rows, dtype=np.int32)
    nr = len(t_tris) if rows_a is None else len(rows_a)
    cls = np.zeros((nr, len(s_tris)), dtype=np.int8)
    _check(lib().orc_classify(_d(t_nodes), len(t_nodes), _i(t_tris), len(t_tris),
                              _d(s_nodes), len(s_nodes), _i(s_tris), len(s_tris), int(same), nx, ny,
                              C.byref(pol), _i(rows_a), nr, cls.ctypes.data_as(C.POINTER(C.c_int8))))
    return cls


def assemble(nodes, tris, family, nx, ny, k=0.0, alpha=1.0, pol=None, rows=None, trial=None,
             out=None, accumulate=False, threads=0):
    """alpha * operator block [rows, all trial columns], complex128 (Fortran order)."""
    pol = pol or policy()
    t_nodes, t_tris = _mesh(nodes, tris)
    same = trial is None
    s_nodes, s_tris = (t_nodes, t_tris) if same else _mesh(*trial)
    rows_a = None if rows is None else np.ascontiguousarray(rows, dtype=np.int32)
    nr = len(t_tris) if rows_a is None else len(rows_a)
    if out is None:
        out = np.zeros((nr, len(s_tris)), dtype=np.complex128, order="F")
        accumulate = False
    alpha = complex(alpha)
    if _is_complex(k) and (t_tris.shape[1] != 3 or s_tris.shape[1] != 3):
        # complex wavenumber, quadrilaterals on either side
        _check(lib().orc_assemble_elems_ck(_d(t_nodes), len(t_nodes), _i(t_tris), len(t_tris), t_tris.shape[1],
                                           _d(s_nodes), len(s_nodes), _i(s_tris), len(s_tris), s_tris.shape[1],
                                           int(same), family, nx, ny, k.real, k.imag, alpha.real, alpha.imag,
                                           C.byref(pol), _i(rows_a), nr, out.ctypes.data_as(C.POINTER(C.c_double)),
                                           out.shape[0], int(accumulate), threads))
    elif _is_complex(k):
        _check(lib().orc_assemble_ck(_d(t_nodes), len(t_nodes), _i(t_tris), len(t_tris),
                                     _d(s_nodes), len(s_nodes), _i(s_tris), len(s_tris), int(same),
                                     family, nx, ny, k.real, k.imag, alpha.real, alpha.imag, C.byref(pol), _i(rows_a),
                                     nr, out.ctypes.data_as(C.POINTER(C.c_double)), out.shape[0], int(accumulate),
                                     threads))
    elif t_tris.shape[1] == 3 and s_tris.shape[1] == 3:
        _check(lib().orc_assemble(_d(t_nodes), len(t_nodes), _i(t_tris), len(t_tris),
                                  _d(s_nodes), len(s_nodes), _i(s_tris), len(s_tris), int(same),
                                  family, nx, ny, k, alpha.real, alpha.imag, C.byref(pol), _i(rows_a), nr,
                                  out.ctypes.data_as(C.POINTER(C.c_double)), out.shape[0], int(accumulate), threads))
    else:
        # quadrilaterals (4 nodes per element) on either side
        _check(lib().orc_assemble_elems(_d(t_nodes), len(t_nodes), _i(t_tris), len(t_tris), t_tris.shape[1],
                                        _d(s_nodes), len(s_nodes), _i(s_tris), len(s_tris), s_tris.shape[1],
                                        int(same), family, nx, ny, k, alpha.real, alpha.imag, C.byref(pol),
                                        _i(rows_a), nr, out.ctypes.data_as(C.POINTER(C.c_double)), out.shape[0],
                                        int(accumulate), threads))
    return out


def assemble_bm(nodes, tris, k, alpha=None, pol=None, rows=None, threads=0):
    """Burton-Miller (A, B) in NiHu's composition order; alpha defaults to -i/k."""
    pol = pol or policy()
    nodes, tris = _mesh(nodes, tris)
    if alpha is None:
        alpha = -1j / complex(k) if _is_complex(k) else complex(0.0, -1.0 / k)
    alpha = complex(alpha)
    rows_a = None if rows is None else np.ascontiguousarray(rows, dtype=np.int32)
    nr = len(tris) if rows_a is None else len(rows_a)
    A = np.zeros((nr, len(tris)), dtype=np.complex128, order="F")
    B = np.zeros((nr, len(tris)), dtype=np.complex128, order="F")
    if _is_complex(k) and tris.shape[1] != 3:
        _check(lib().orc_assemble_bm_elems_ck(_d(nodes), len(nodes), _i(tris), len(tris), tris.shape[1], k.real,
                                              k.imag, alpha.real, alpha.imag, C.byref(pol), _i(rows_a), nr,
                                              A.ctypes.data_as(C.POINTER(C.c_double)),
                                              B.ctypes.data_as(C.POINTER(C.c_double)), nr, threads))
    elif _is_complex(k):
        _check(lib().orc_assemble_bm_ck(_d(nodes), len(nodes), _i(tris), len(tris), k.real, k.imag, alpha.real,
                                        alpha.imag, C.byref(pol), _i(rows_a), nr,
                                        A.ctypes.data_as(C.POINTER(C.c_double)),
                                        B.ctypes.data_as(C.POINTER(C.c_double)), nr, threads))
    elif tris.shape[1] == 3:
        _check(lib().orc_assemble_bm(_d(nodes), len(nodes), _i(tris), len(tris), k, alpha.real, alpha.imag,
                                     C.byref(pol), _i(rows_a), nr, A.ctypes.data_as(C.POINTER(C.c_double)),
                                     B.ctypes.data_as(C.POINTER(C.c_double)), nr, threads))
    else:
        _check(lib().orc_assemble_bm_elems(_d(nodes), len(nodes), _i(tris), len(tris), tris.shape[1], k, alpha.real,
                                           alpha.imag, C.byref(pol), _i(rows_a), nr,
                                           A.ctypes.data_as(C.POINTER(C.c_double)),
                                           B.ctypes.data_as(C.POINTER(C.c_double)), nr, threads))
    return A, B


def sweep_bm(nodes, tris, ks, pol=None, rows=None, threads=0):
    """NiHu's OpenMP-over-frequencies pattern; returns A, B of shape [n_k, n_rows, n_tri]."""
    pol = pol or policy()
    nodes, tris = _mesh(nodes, tris)
    ks = np.ascontiguousarray(ks, dtype=np.float64)
    rows_a = None if rows is None else np.ascontiguousarray(rows, dtype=np.int32)
    nr = len(tris) if rows_a is None else len(rows_a)
    A = np.zeros((len(ks), len(tris), nr), dtype=np.complex128)
    B = np.zeros_like(A)
    _check(lib().orc_sweep_bm(_d(nodes), len(nodes), _i(tris), len(tris), _d(ks), len(ks), C.byref(pol),
                              _i(rows_a), nr, A.ctypes.data_as(C.POINTER(C.c_double)),
                              B.ctypes.data_as(C.POINTER(C.c_double)), threads))
    return A.transpose(0, 2, 1), B.transpose(0, 2, 1)


def assemble_galerkin(nodes, tris, family, nx, ny, k=0.0, alpha=1.0, pol=None, trial=None, linear=True,
                      out=None, accumulate=False, threads=0, test_elems=None):
    """alpha * Galerkin matrix, complex128 (Fortran order): linear=True gives
    isoparametric linear triangles ([test nodes, trial nodes]), linear=False
    constant triangles ([test elements, trial elements]); family 2 = identity.
    test_elems: only these test elements contribute (rows whose elements are
    all listed are complete)."""
    pol = pol or policy()
    t_nodes, t_tris = _mesh(nodes, tris)
    same = trial is None
    s_nodes, s_tris = (t_nodes, t_tris) if same else _mesh(*trial)
    # linear: True / 1 isoparametric x isoparametric, False / 0 constant x
    # constant, 2 constant test x isoparametric trial, 3 isoparametric test x
    # constant trial
    linear = int(linear)
    lt, ls = linear in (1, 3), linear in (1, 2)
    if out is None:
        shape = (len(t_nodes) if lt else len(t_tris), len(s_nodes) if ls else len(s_tris))
        out = np.zeros(shape, dtype=np.complex128, order="F")
        accumulate = False
    alpha = complex(alpha)
    te = None if test_elems is None else np.ascontiguousarray(test_elems, dtype=np.int32)
    if t_tris.shape[1] == 4 or s_tris.shape[1] == 4:
        if t_tris.shape[1] != 4 or s_tris.shape[1] != 4:
            raise ValueError("Galerkin pairs of a triangle and a quadrilateral are not built")
        kc = complex(k)
        _check(lib().orc_assemble_galerkin_quad(_d(t_nodes), len(t_nodes), _i(t_tris), len(t_tris),
                                                _d(s_nodes), len(s_nodes), _i(s_tris), len(s_tris), int(same),
                                                int(linear), family, nx, ny, kc.real, kc.imag, alpha.real,
                                                alpha.imag, C.byref(pol), _i(te), 0 if te is None else len(te),
                                                out.ctypes.data_as(C.POINTER(C.c_double)), out.shape[0],
                                                int(accumulate), threads))
        return out
    if _is_complex(k):
        _check(lib().orc_assemble_galerkin_ck(_d(t_nodes), len(t_nodes), _i(t_tris), len(t_tris),
                                              _d(s_nodes), len(s_nodes), _i(s_tris), len(s_tris), int(same),
                                              int(linear), family, nx, ny, k.real, k.imag, alpha.real, alpha.imag,
                                              C.byref(pol), _i(te), 0 if te is None else len(te),
                                              out.ctypes.data_as(C.POINTER(C.c_double)), out.shape[0],
                                              int(accumulate), threads))
        return out
    _check(lib().orc_assemble_galerkin(_d(t_nodes), len(t_nodes), _i(t_tris), len(t_tris),
                                       _d(s_nodes), len(s_nodes), _i(s_tris), len(s_tris), int(same), int(linear),
                                       family, nx, ny, k, alpha.real, alpha.imag, C.byref(pol), _i(te),
                                       0 if te is None else len(te), out.ctypes.data_as(C.POINTER(C.c_double)),
                                       out.shape[0],
                                       int(accumulate), threads))
    return out


def galerkin_quad_block(family, nx, ny, k, test_coords, test_nodes, trial_coords, trial_nodes,
                        same_mesh=True, same_elem=False, pol=None, linear=True):
    """4x4 Galerkin block (test shape a, trial shape b) of one quadrilateral
    pair, corners (4, 3) each (linear=False: the constant-field value in [0, 0])."""
    pol = pol or policy()
    kc = complex(k)
    tc = np.ascontiguousarray(test_coords, dtype=np.float64).reshape(12)
    sc = np.ascontiguousarray(trial_coords, dtype=np.float64).reshape(12)
    tn = np.ascontiguousarray(test_nodes, dtype=np.int32)
    sn = np.ascontiguousarray(trial_nodes, dtype=np.int32)
    out = np.zeros(32)
    _check(lib().orc_galerkin_quad_block(family, nx, ny, kc.real, kc.imag, C.byref(pol), _d(tc), _i(tn), _d(sc),
                                         _i(sn), int(same_mesh), int(same_elem), int(linear), _d(out)))
    return (out[0::2] + 1j * out[1::2]).reshape(4, 4)


def quad_singular_rule(match, degree=7):
    """Canonical quad x quad singular Galerkin rule of a match dimension (2
    face, 1 edge, 0 corner) built on the Gauss rule of a degree: columns
    (test xi, test eta, trial xi, trial eta, weight)."""
    n1 = degree // 2 + 1
    cap = 6 * n1 ** 4 * 2
    pts = np.zeros((cap, 5))
    n = C.c_int(0)
    _check(lib().orc_quad_singular_rule(match, degree, _d(pts), cap, C.byref(n)))
    return pts[:n.value].copy()


def overlap_quads(a, b):
    """(count, ind1, ind2) of NiHu's node-overlap scan of two quadrilaterals"""
    a = np.ascontiguousarray(a, dtype=np.int32)
    b = np.ascontiguousarray(b, dtype=np.int32)
    i1, i2 = C.c_int(0), C.c_int(0)
    n = lib().orc_overlap_quads(_i(a), _i(b), C.byref(i1), C.byref(i2))
    return n, i1.value, i2.value


def singular_rule(match):
    """Canonical singular Galerkin rule of a match dimension (2 face, 1 edge, 0 vertex):
    columns (test xi, test eta, trial xi, trial eta, weight)."""
    cap = 4 ** 4 * 6 * 2
    pts = np.zeros((cap, 5))
    n = C.c_int(0)
    _check(lib().orc_singular_rule(match, _d(pts), cap, C.byref(n)))
    return pts[:n.value].copy()


def galerkin_block(family, nx, ny, k, test_coords, test_nodes, trial_coords, trial_nodes,
                   same_mesh=True, same_elem=False, pol=None, linear=True):
    """3x3 Galerkin block (test shape a, trial shape b) of one element pair
    (linear=False: the constant-field value in [0, 0])."""
    pol = pol or policy()
    tc = np.ascontiguousarray(test_coords, dtype=np.float64).reshape(9)
    sc = np.ascontiguousarray(trial_coords, dtype=np.float64).reshape(9)
    tn = np.ascontiguousarray(test_nodes, dtype=np.int32)
    sn = np.ascontiguousarray(trial_nodes, dtype=np.int32)
    out = np.zeros(18)
    _check(lib().orc_galerkin_block(family, nx, ny, k, C.byref(pol), _d(tc), _i(tn), _d(sc), _i(sn),
                                    int(same_mesh), int(same_elem), int(linear), _d(out)))
    return (out[0::2] + 1j * out[1::2]).reshape(3, 3)


def overlap(a, b):
    """(count, ind1, ind2) of NiHu's node-overlap scan of two elements"""
    a = np.ascontiguousarray(a, dtype=np.int32)
    b = np.ascontiguousarray(b, dtype=np.int32)
    i1, i2 = C.c_int(0), C.c_int(0)
    n = lib().orc_overlap(_i(a), _i(b), C.byref(i1), C.byref(i2))
    return n, i1.value, i2.value


def num_threads():
    return lib().orc_num_threads()


QUAD_REGULAR, QUAD_SINGULAR, QUAD_POLAR, QUAD_TELLES, QUAD_GUIGGIANI, QUAD_DUFFY, QUAD_INVERSE, QUAD_NEAR, \
    QUAD_PLANAR = range(9)


def quad_elem(what, family, nx, ny, k, coords, x0=None, nx0=None, xi0=None, o1=0, o2=0):
    """Element integral over a quadrilateral (coords: 4 x 3, one corner per
    row); `what` is one of the QUAD_* constants (see orc_quad_elem).
    QUAD_INVERSE returns (xi, eta, zeta, converged) of x0 (o1: -log10 tol)."""
    cs = np.ascontiguousarray(coords, dtype=np.float64).reshape(12)
    arrs = [None if v is None else np.ascontiguousarray(v, dtype=np.float64) for v in (x0, nx0, xi0)]
    out = np.zeros(4)
    ptrs = [None if a is None else _d(a) for a in arrs]
    if _is_complex(k):
        _check(lib().orc_quad_elem_ck(what, family, nx, ny, k.real, k.imag, _d(cs), *ptrs, o1, o2, _d(out)))
    else:
        _check(lib().orc_quad_elem(what, family, nx, ny, k, _d(cs), *ptrs, o1, o2, _d(out)))
    if what == QUAD_INVERSE:
        return out.copy()
    return complex(out[0], out[1])


def quad_duffy_rule(degree, xi0=0.0, eta0=0.0):
    """Blind Duffy rule around (xi0, eta0) on [-1, 1]^2: rows (xi, eta, w)."""
    cap = 4 * (degree // 2 + 1) ** 2
    pts = np.zeros((cap, 3))
    n = C.c_int(0)
    _check(lib().orc_quad_duffy_rule(degree, xi0, eta0, _d(pts), cap, C.byref(n)))
    return pts[:n.value].copy()


def telles_rule(degree, xi0, eta0, d):
    """Telles-transformed Gauss square rule: rows (xi, eta, w)."""
    cap = (degree // 2 + 1) ** 2
    pts = np.zeros((cap, 3))
    n = C.c_int(0)
    _check(lib().orc_telles_rule(degree, xi0, eta0, d, _d(pts), cap, C.byref(n)))
    return pts[:n.value].copy()


def geometry_quad(nodes, quads):
    """Quadrilateral centres x(0, 0), unit normals there, mean edges, diameters."""
    nodes, quads = _mesh(nodes, quads)
    n = len(quads)
    c, nu, R, diam = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(n), np.zeros(n)
    _check(lib().orc_geometry_quad(_d(nodes), len(nodes), _i(quads), n, _d(c), _d(nu), _d(R), _d(diam)))
    return dict(center=c, normal=nu, R=R, diam=diam)


# ---------------------------------------------------------------- gauss field
GAUSS_REGULAR, GAUSS_SINGULAR, GAUSS_POLAR, GAUSS_TELLES, GAUSS_GUIGGIANI, GAUSS_DUFFY = range(6)
GAUSS_NEAR = 7


def gauss_shape(xi, eta):
    """(N[4], dN/dxi[4], dN/deta[4]) of the quadrilateral gauss field at
    (xi, eta), and its nodes (4 x 2)."""
    out, nodes = np.zeros(12), np.zeros(8)
    lib().orc_gauss_shape(float(xi), float(eta), _d(out), _d(nodes))
    return out[:4].copy(), out[4:8].copy(), out[8:].copy(), nodes.reshape(4, 2)


def gauss_elem(what, family, nx, ny, k, coords, x0=None, nx0=None, xi0=None, o1=0, o2=0):
    """Element integral over a quadrilateral with the gauss field: 4 complex
    values, one per trial dof (what: a GAUSS_* constant, see orc_gauss_elem)."""
    cs = np.ascontiguousarray(coords, dtype=np.float64).reshape(12)
    arrs = [None if v is None else np.ascontiguousarray(v, dtype=np.float64) for v in (x0, nx0, xi0)]
    out = np.zeros(8)
    kc = complex(k)
    _check(lib().orc_gauss_elem(what, family, nx, ny, kc.real, kc.imag, _d(cs),
                                *[None if a is None else _d(a) for a in arrs], o1, o2, _d(out)))
    return out[0::2] + 1j * out[1::2]


def assemble_gauss(nodes, quads, family, nx, ny, k=0.0, alpha=1.0, pol=None, rows=None, trial=None,
                   test_gauss=True, out=None, accumulate=False, threads=0):
    """alpha * operator with the quadrilateral gauss field (4 dofs per
    element, column 4 j + b) on the trial mesh (trial=None: the test mesh);
    rows: the nodes 4 i + a of the test mesh's gauss field (test_gauss) or
    its elements (field points, any element type), complex128 Fortran order."""
    pol = pol or policy()
    t_nodes, t_el = _mesh(nodes, quads)
    same = trial is None
    s_nodes, s_el = (t_nodes, t_el) if same else _mesh(*trial)
    if s_el.shape[1] != 4:
        raise ValueError("the gauss field needs a quadrilateral trial mesh")
    rows_a = None if rows is None else np.ascontiguousarray(rows, dtype=np.int32)
    n_test_rows = 4 * len(t_el) if test_gauss else len(t_el)
    nr = n_test_rows if rows_a is None else len(rows_a)
    if out is None:
        out = np.zeros((nr, 4 * len(s_el)), dtype=np.complex128, order="F")
        accumulate = False
    alpha = complex(alpha)
    kc = complex(k)
    _check(lib().orc_assemble_gauss(_d(t_nodes), len(t_nodes), _i(t_el), len(t_el), t_el.shape[1],
                                    _d(s_nodes), len(s_nodes), _i(s_el), len(s_el), int(same), int(test_gauss),
                                    family, nx, ny, kc.real, kc.imag, alpha.real, alpha.imag, C.byref(pol),
                                    _i(rows_a), nr, out.ctypes.data_as(C.POINTER(C.c_double)), out.shape[0],
                                    int(accumulate), threads))
    return out


def assemble_bm_gauss(nodes, quads, k, alpha=None, pol=None, rows=None, threads=0):
    """Burton-Miller (A, B) with the quadrilateral gauss field; alpha
    defaults to -i/k."""
    pol = pol or policy()
    nodes, quads = _mesh(nodes, quads)
    kc = complex(k)
    alpha = -1j / kc if alpha is None else complex(alpha)
    rows_a = None if rows is None else np.ascontiguousarray(rows, dtype=np.int32)
    nd = 4 * len(quads)
    nr = nd if rows_a is None else len(rows_a)
    A = np.zeros((nr, nd), dtype=np.complex128, order="F")
    B = np.zeros((nr, nd), dtype=np.complex128, order="F")
    _check(lib().orc_assemble_bm_gauss(_d(nodes), len(nodes), _i(quads), len(quads), kc.real, kc.imag,
                                       alpha.real, alpha.imag, C.byref(pol), _i(rows_a), nr,
                                       A.ctypes.data_as(C.POINTER(C.c_double)),
                                       B.ctypes.data_as(C.POINTER(C.c_double)), nr, threads))
    return A, B
