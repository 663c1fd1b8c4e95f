"""ctypes binding of libnbem.so (include/nbem.h).

Device buffers are torch tensors on the GPU (torch supplies memory and
streams; all arithmetic happens in the library's HIP kernels).  Host buffers
are numpy arrays.  There is no fallback: if the library is missing, loading
raises.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnbem.so")

NB_OK, NB_ERR_ARG, NB_ERR_UNSUPPORTED, NB_ERR_RANGE, NB_ERR_OOM, NB_ERR_DEVICE, NB_ERR_RUNTIME = range(7)
NB_LAPLACE, NB_HELMHOLTZ, NB_IDENTITY = 0, 1, 2
NB_RULE_NIHU, NB_RULE_FIXED = 0, 1
NB_NEAR_MEAN_EDGE, NB_NEAR_DIAMETER, NB_NEAR_NONE = 0, 1, 2
NB_PREC_MIXED, NB_PREC_FP64 = 0, 1
NB_C64, NB_C128, NB_F32, NB_F64 = 0, 1, 2, 3
NB_HOST, NB_DEVICE = 0, 1
NB_COLLOCATION, NB_GALERKIN, NB_GALERKIN_LINEAR = 0, 1, 2
NB_GALERKIN_CONST_LINEAR, NB_GALERKIN_LINEAR_CONST = 5, 6
# collocation with the quadrilateral gauss field (4 dofs per element at the
# 2 x 2 Gauss points, NiHu's quad_1_gauss_field): gauss test field, or the
# constant test field of another mesh (field points)
NB_COLLOCATION_GAUSS, NB_COLLOCATION_GAUSS_TRIAL = 3, 4
NB_TRIA_1_ID, NB_QUAD_1_ID = 32303, 32404   # element type ids of NiHu's tria_1 and quad_1
FORMALISMS = {"collocation": NB_COLLOCATION, "galerkin": NB_GALERKIN, "galerkin_linear": NB_GALERKIN_LINEAR,
              "collocation_gauss": NB_COLLOCATION_GAUSS, "collocation_gauss_trial": NB_COLLOCATION_GAUSS_TRIAL,
              "galerkin_const_linear": NB_GALERKIN_CONST_LINEAR, "galerkin_linear_const": NB_GALERKIN_LINEAR_CONST}

# symbols declared in include/nbem.h
EXPORTS = ["nb_last_error", "nb_version", "nb_device_info", "nb_mesh_create", "nb_mesh_create_quads",
           "nb_mesh_create_elements", "nb_mesh_destroy", "nb_mesh_info",
           "nb_assemble", "nb_assemble_bm", "nb_assemble_sweep", "nb_classify", "nb_autotune", "nb_set_launch",
           "nb_synchronize", "nb_profile", "nb_profile_read", "nb_eval_math",
           "nb_assemble_sweep_noalloc", "nb_gemv_workspace", "nb_gemv_c64", "nb_assemble_bm_ck",
           "nb_assemble_bm_gauss"]


class NbKernel(C.Structure):
    _fields_ = [("family", C.c_int32), ("nx", C.c_int32), ("ny", C.c_int32), ("k", C.c_double),
                ("k_im", C.c_double)]


class NbScalar(C.Structure):
    _fields_ = [("re", C.c_double), ("im", C.c_double)]


class NbPolicy(C.Structure):
    _fields_ = [("rule", C.c_int32), ("fixed_degree", C.c_int32), ("near_mode", C.c_int32), ("near_limit", C.c_double),
                ("precision", C.c_int32)]


class NbOut(C.Structure):
    _fields_ = [("ptr", C.c_void_p), ("rows", C.c_int64), ("cols", C.c_int64), ("ld", C.c_int64),
                ("dtype", C.c_int32), ("location", C.c_int32), ("accumulate", C.c_int32)]


class NbLaunch(C.Structure):
    _fields_ = [("tile_cols", C.c_int32), ("k_batch", C.c_int32), ("persistent", C.c_int32)]


class NbemError(RuntimeError):
    pass


_lib = None


def lib():
    """Load libnbem.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libnbem.so not built: run `python -m nihu_amd.build` (HIP engine missing)")
        l = C.CDLL(LIB_PATH)
        P = C.c_void_p
        l.nb_last_error.restype = C.c_char_p
        l.nb_device_info.argtypes = [C.c_char_p, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int64)]
        l.nb_mesh_create.argtypes = [P, C.c_int32, P, C.c_int32, C.POINTER(P)]
        l.nb_mesh_create_quads.argtypes = [P, C.c_int32, P, C.c_int32, C.POINTER(P)]
        l.nb_mesh_create_elements.argtypes = [P, C.c_int32, P, C.c_int32, C.POINTER(P)]
        l.nb_mesh_destroy.argtypes = [P]
        l.nb_mesh_info.argtypes = [P, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int64)]
        l.nb_assemble.argtypes = [P, P, C.c_int32, C.c_int32, NbKernel, NbScalar, C.POINTER(NbPolicy),
                                  C.POINTER(NbOut), P]
        l.nb_assemble_bm.argtypes = [P, C.c_double, NbScalar, C.POINTER(NbPolicy), C.POINTER(NbOut),
                                     C.POINTER(NbOut), P]
        l.nb_assemble_bm_ck.argtypes = [P, NbScalar, NbScalar, C.POINTER(NbPolicy), C.POINTER(NbOut),
                                        C.POINTER(NbOut), P]
        l.nb_assemble_bm_gauss.argtypes = [P, NbScalar, NbScalar, C.POINTER(NbPolicy), C.POINTER(NbOut),
                                           C.POINTER(NbOut), P]
        l.nb_assemble_sweep.argtypes = [P, P, C.c_int32, C.POINTER(NbPolicy), P, P, C.c_int64, C.c_int64,
                                        C.POINTER(NbLaunch), P]
        l.nb_assemble_sweep_noalloc.argtypes = [P, P, C.c_int32, C.POINTER(NbPolicy), P, P, C.c_int64, C.c_int64,
                                                C.POINTER(NbLaunch), P]
        l.nb_classify.argtypes = [P, P, C.c_int32, C.c_int32, C.c_int32, C.POINTER(NbPolicy), P, P]
        l.nb_autotune.argtypes = [P, C.c_int32, C.POINTER(NbPolicy), C.POINTER(NbLaunch)]
        l.nb_set_launch.argtypes = [C.POINTER(NbLaunch)]
        l.nb_synchronize.argtypes = [P]
        l.nb_eval_math.argtypes = [C.c_int32, P, P, C.c_int64, P]
        l.nb_gemv_workspace.argtypes = [C.c_int64, C.c_int64, C.POINTER(C.c_int64)]
        l.nb_gemv_c64.argtypes = [P, C.c_int64, C.c_int64, C.c_int64, P, P, P, C.c_int64, P]
        l.nb_profile.argtypes = [C.c_int32]
        l.nb_profile_read.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_int64), C.c_int32]
        _lib = l
    return _lib


def check(status: int):
    if status == NB_OK:
        return
    msg = lib().nb_last_error().decode()
    if status == NB_ERR_ARG:
        raise ValueError(msg)
    if status == NB_ERR_UNSUPPORTED:
        raise NotImplementedError(msg)
    if status == NB_ERR_RANGE:
        raise IndexError(msg)
    if status == NB_ERR_OOM:
        raise MemoryError(msg)
    raise NbemError(msg)


def policy(rule: str | int = "nihu", degree: int = 5, near: str | int = "mean_edge", limit: float = 1.5,
           precision: str | int = "mixed") -> NbPolicy:
    """'nihu' = NiHu's per-pair degree tables; 'fixed' = one Dunavant degree.
    near: 'mean_edge' (NiHu, limit 1.5), 'diameter' (limit in element diameters) or 'none'.
    precision: 'mixed' (regular pairs in FP32 with the phase reduced in FP64)
    or 'fp64' (every pair in FP64; collocation on every mesh type and
    field; not the Galerkin formalism)."""
    r = {"nihu": NB_RULE_NIHU, "fixed": NB_RULE_FIXED}.get(rule, rule)
    n = {"mean_edge": NB_NEAR_MEAN_EDGE, "diameter": NB_NEAR_DIAMETER, "none": NB_NEAR_NONE}.get(near, near)
    p = {"mixed": NB_PREC_MIXED, "fp64": NB_PREC_FP64}.get(precision, precision)
    return NbPolicy(int(r), int(degree), int(n), float(limit), int(p))


NIHU_POLICY = dict(rule="nihu", near="mean_edge", limit=1.5)
# the headline workload's policy: 7-point rule far, FP64 within 2 element diameters
CONFIG_POLICY = dict(rule="fixed", degree=5, near="diameter", limit=2.0)


def _stream_ptr(stream):
    if stream is None:
        try:
            import torch
            if torch.cuda.is_available() and torch.cuda.is_initialized():
                return C.c_void_p(torch.cuda.current_stream().cuda_stream)
        except Exception:
            pass
        return None
    if hasattr(stream, "cuda_stream"):
        return C.c_void_p(stream.cuda_stream)
    return C.c_void_p(int(stream))


class Mesh:
    """A mesh resident on the current GPU (nb_mesh): elements with 3 node
    indices (flat triangles) or 4 (bilinear quadrilaterals); from_elements
    takes NiHu's element table, which may hold both."""

    def __init__(self, nodes, tris):
        self.nodes = np.ascontiguousarray(nodes, dtype=np.float64)
        self.tris = np.ascontiguousarray(tris, dtype=np.int32)
        if self.nodes.ndim != 2 or self.nodes.shape[1] != 3 or self.tris.ndim != 2 or self.tris.shape[1] not in (3, 4):
            raise ValueError("nodes must be n x 3 and elements m x 3 (triangles) or m x 4 (quadrilaterals)")
        self._npe = self.tris.shape[1]
        h = C.c_void_p()
        create = lib().nb_mesh_create if self._npe == 3 else lib().nb_mesh_create_quads
        check(create(self.nodes.ctypes.data, len(self.nodes), self.tris.ctypes.data, len(self.tris), C.byref(h)))
        self._h = h

    @classmethod
    def from_elements(cls, nodes, elements):
        """NiHu's element table: rows [type id, n0, n1, n2(, n3)] with type id
        32303 (tria_1) or 32404 (quad_1), element id = row number whatever the
        type (nihu/core/mesh.hpp:137-193).  A table of one type gives the mesh
        of that type (tris: its node indices); with both types, tris is the
        table itself (m x 5) and nodes_per_element is 0."""
        el = np.asarray(elements)
        if el.ndim != 2 or el.shape[1] not in (4, 5) or len(el) == 0:
            raise ValueError("elements must be a non-empty m x 4 or m x 5 table [type id, node ids]")
        table = np.zeros((len(el), 5), dtype=np.int32)
        table[:, :el.shape[1]] = el
        types = set(np.unique(table[:, 0]).tolist())
        if not types <= {NB_TRIA_1_ID, NB_QUAD_1_ID}:
            raise ValueError("unknown element type id (tria_1 32303 and quad_1 32404 are supported)")
        if types == {NB_QUAD_1_ID} and el.shape[1] < 5:
            raise ValueError("quad_1 rows need four node indices")
        if len(types) == 1:
            return cls(nodes, table[:, 1:4] if types == {NB_TRIA_1_ID} else table[:, 1:5])
        self = cls.__new__(cls)
        self.nodes = np.ascontiguousarray(nodes, dtype=np.float64)
        if self.nodes.ndim != 2 or self.nodes.shape[1] != 3:
            raise ValueError("nodes must be n x 3")
        self.tris = table
        self._npe = 0
        h = C.c_void_p()
        check(lib().nb_mesh_create_elements(self.nodes.ctypes.data, len(self.nodes), table.ctypes.data, len(table),
                                             C.byref(h)))
        self._h = h
        return self

    @property
    def nodes_per_element(self):
        """3 (triangles), 4 (quadrilaterals) or 0 (both)"""
        return self._npe

    @property
    def handle(self):
        return self._h

    @property
    def n(self):
        return len(self.tris)

    @property
    def n_nodes(self):
        return len(self.nodes)

    def dofs(self, formalism="collocation", side="test"):
        """rows (side "test") or columns (side "trial") of this mesh's side of
        a matrix: nodes for an isoparametric Galerkin field, 4 per element for
        the gauss field, elements otherwise"""
        f = FORMALISMS[formalism]
        if f == NB_GALERKIN_LINEAR or (f == NB_GALERKIN_CONST_LINEAR and side == "trial") or \
                (f == NB_GALERKIN_LINEAR_CONST and side == "test"):
            return self.n_nodes
        return 4 * self.n if f == NB_COLLOCATION_GAUSS else self.n

    def device_bytes(self):
        b = C.c_int64()
        check(lib().nb_mesh_info(self._h, None, None, C.byref(b)))
        return b.value

    def close(self):
        if getattr(self, "_h", None):
            lib().nb_mesh_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_TORCH_DT = {"complex64": NB_C64, "complex128": NB_C128, "float32": NB_F32, "float64": NB_F64}
_NP_DT = {np.dtype(np.complex64): NB_C64, np.dtype(np.complex128): NB_C128, np.dtype(np.float32): NB_F32,
          np.dtype(np.float64): NB_F64}


def out_desc(arr, accumulate=True) -> NbOut:
    """Describe a column-major result: a torch GPU tensor (2-D, column-major,
    i.e. arr.t() contiguous or a transposed view) or a Fortran-ordered numpy array."""
    if isinstance(arr, np.ndarray):
        if arr.ndim != 2:
            raise ValueError("result must be 2-D")
        if arr.dtype not in _NP_DT:
            raise ValueError("unsupported result dtype %s" % arr.dtype)
        isz = arr.itemsize
        if arr.strides[0] != isz or arr.strides[1] % isz:
            raise ValueError("host result must be column-major (Fortran order)")
        ld = arr.strides[1] // isz if arr.shape[1] > 1 else arr.shape[0]
        return NbOut(arr.ctypes.data, arr.shape[0], arr.shape[1], ld, _NP_DT[arr.dtype], NB_HOST, int(accumulate))
    import torch
    if not isinstance(arr, torch.Tensor) or not arr.is_cuda:
        raise ValueError("device result must be a CUDA/HIP torch tensor")
    if arr.dim() != 2:
        raise ValueError("result must be 2-D")
    dt = _TORCH_DT.get(str(arr.dtype).replace("torch.", ""))
    if dt is None:
        raise ValueError("unsupported result dtype %s" % arr.dtype)
    if arr.stride(0) != 1:
        raise ValueError("device result must be column-major (use a transposed view of a row-major tensor)")
    ld = arr.stride(1) if arr.shape[1] > 1 else arr.shape[0]
    return NbOut(arr.data_ptr(), arr.shape[0], arr.shape[1], ld, dt, NB_DEVICE, int(accumulate))


def colmajor(rows, cols, dtype="complex64", device="cuda"):
    """A zeroed column-major torch matrix (rows x cols), as a transposed view."""
    import torch
    return torch.zeros((cols, rows), dtype=getattr(torch, dtype), device=device).t()


def assemble(test: Mesh, trial: Mesh, out, family=NB_HELMHOLTZ, nx=0, ny=0, k=0.0, alpha=1.0, pol=None,
             same_mesh=None, accumulate=True, stream=None, formalism="collocation"):
    """out (+)= alpha * operator; formalism "collocation" (dirac(constant) x
    constant), "galerkin" (constant x constant), "galerkin_linear"
    (isoparametric x isoparametric, rows / columns = mesh nodes),
    "galerkin_const_linear" (constant test x isoparametric trial) or
    "galerkin_linear_const" (the reverse).  k may be
    complex (collocation on every mesh type, Galerkin on meshes of triangles
    or of quadrilaterals)."""
    pol = pol or policy(**NIHU_POLICY)
    same = (test is trial) if same_mesh is None else bool(same_mesh)
    a = complex(alpha)
    kc = complex(k)
    o = out_desc(out, accumulate)
    check(lib().nb_assemble(test.handle, trial.handle, int(same), FORMALISMS[formalism],
                            NbKernel(family, nx, ny, kc.real, kc.imag), NbScalar(a.real, a.imag), C.byref(pol),
                            C.byref(o), _stream_ptr(stream)))
    return out


def assemble_bm(mesh: Mesh, k, A, B, alpha=None, pol=None, accumulate=False, stream=None):
    """Burton-Miller A and B; alpha defaults to -i/k.  A complex k goes to
    nb_assemble_bm_ck."""
    pol = pol or policy(**NIHU_POLICY)
    ck = isinstance(k, (complex, np.complexfloating))
    if alpha is None:
        a = complex(-1j / k) if ck else complex(0.0, -1.0 / k)
    else:
        a = complex(alpha)
    oa, ob = out_desc(A, accumulate), out_desc(B, accumulate)
    if ck:
        check(lib().nb_assemble_bm_ck(mesh.handle, NbScalar(k.real, k.imag), NbScalar(a.real, a.imag), C.byref(pol),
                                      C.byref(oa), C.byref(ob), _stream_ptr(stream)))
    else:
        check(lib().nb_assemble_bm(mesh.handle, float(k), NbScalar(a.real, a.imag), C.byref(pol), C.byref(oa),
                                   C.byref(ob), _stream_ptr(stream)))
    return A, B


def assemble_bm_gauss(mesh: Mesh, k, A, B, alpha=None, pol=None, accumulate=False, stream=None):
    """Burton-Miller A and B (4 n x 4 n) with the quadrilateral gauss field;
    alpha defaults to -i/k; k real or complex."""
    pol = pol or policy(**NIHU_POLICY)
    kc = complex(k)
    a = complex(-1j / kc) if alpha is None else complex(alpha)
    oa, ob = out_desc(A, accumulate), out_desc(B, accumulate)
    check(lib().nb_assemble_bm_gauss(mesh.handle, NbScalar(kc.real, kc.imag), NbScalar(a.real, a.imag),
                                     C.byref(pol), C.byref(oa), C.byref(ob), _stream_ptr(stream)))
    return A, B


def assemble_sweep(mesh: Mesh, ks, A, B, pol=None, launch=None, stream=None, noalloc=False):
    """A, B: torch complex64 tensors of shape (n_k, n, n) holding column-major
    matrices (e.g. a transposed view per frequency); alpha = -i/k each.
    noalloc: nb_assemble_sweep_noalloc (nothing allocated; the mesh's tables
    for the policy must be resident from an earlier sweep)."""
    pol = pol or policy(**NIHU_POLICY)
    ks = np.ascontiguousarray(ks, dtype=np.float64)
    n = mesh.n
    if A.shape != (len(ks), n, n) or B.shape != A.shape:
        raise ValueError("A and B must be (n_k, n, n)")
    if A.stride(1) != 1 or B.stride(1) != 1 or A.stride() != B.stride():
        raise ValueError("A and B must hold column-major matrices with equal strides")
    ld, kst = A.stride(2), A.stride(0)
    lp = C.byref(launch) if launch is not None else None
    fn = lib().nb_assemble_sweep_noalloc if noalloc else lib().nb_assemble_sweep
    check(fn(mesh.handle, ks.ctypes.data, len(ks), C.byref(pol), A.data_ptr(), B.data_ptr(), ld, kst, lp,
             _stream_ptr(stream)))
    return A, B


def classify(test: Mesh, trial: Mesh, nx, ny, pol=None, same_mesh=None, stream=None):
    import torch
    pol = pol or policy(**NIHU_POLICY)
    same = (test is trial) if same_mesh is None else bool(same_mesh)
    cls = torch.empty((test.n, trial.n), dtype=torch.int8, device="cuda")
    check(lib().nb_classify(test.handle, trial.handle, int(same), nx, ny, C.byref(pol), C.c_void_p(cls.data_ptr()),
                            _stream_ptr(stream)))
    return cls


def autotune(mesh: Mesh, n_k=1, pol=None):
    pol = pol or policy(**NIHU_POLICY)
    best = NbLaunch()
    check(lib().nb_autotune(mesh.handle, int(n_k), C.byref(pol), C.byref(best)))
    return best


def set_launch(tile_cols=0, k_batch=0, persistent=0):
    check(lib().nb_set_launch(C.byref(NbLaunch(tile_cols, k_batch, persistent))))


def gemv_workspace(A):
    """The work buffer nb_gemv_c64 needs for the matrix A (a torch tensor)."""
    import torch
    b = C.c_int64()
    check(lib().nb_gemv_workspace(A.shape[0], A.shape[1], C.byref(b)))
    return torch.empty(max(1, b.value // 16), dtype=torch.complex128, device=A.device)


def gemv(A, x, y=None, work=None, stream=None):
    """y = A x in FP64 for a column-major complex64 GPU matrix A (stride(0)
    == 1) and a complex128 GPU vector x (nb_gemv_c64); returns y."""
    import torch
    if A.dtype != torch.complex64 or not A.is_cuda or A.dim() != 2:
        raise ValueError("A must be a complex64 GPU matrix")
    if A.stride(0) != 1 and A.shape[0] > 1:
        raise ValueError("A must be column-major")
    rows, cols = A.shape
    x = x.to(torch.complex128).contiguous()
    if x.shape != (cols,):
        raise ValueError("x must have one entry per column of A")
    if y is None:
        y = torch.empty(rows, dtype=torch.complex128, device=A.device)
    if work is None:
        work = gemv_workspace(A)
    ld = A.stride(1) if cols > 1 else rows
    check(lib().nb_gemv_c64(C.c_void_p(A.data_ptr()), rows, cols, ld, C.c_void_p(x.data_ptr()),
                            C.c_void_p(y.data_ptr()), C.c_void_p(work.data_ptr()), work.numel() * 16,
                            _stream_ptr(stream)))
    return y


MATH_ROUTINES = {"sincos": 0, "log": 1, "rsqrt": 2, "rcp": 3}


def eval_math(fn, x):
    """The engine's FP64 device routine `fn` ("sincos", "log", "rsqrt",
    "rcp") at the values of the torch float64 GPU tensor x; returns (n, 2)
    (sin, cos) or (n,) values."""
    import torch
    x = x.contiguous()
    if x.dtype != torch.float64 or not x.is_cuda:
        raise ValueError("x must be a float64 GPU tensor")
    y = torch.empty((x.numel(), 2), dtype=torch.float64, device=x.device)
    check(lib().nb_eval_math(MATH_ROUTINES[fn], C.c_void_p(x.data_ptr()), C.c_void_p(y.data_ptr()), x.numel(),
                             _stream_ptr(None)))
    return y if fn == "sincos" else y[:, 0]


def synchronize(stream=None):
    check(lib().nb_synchronize(_stream_ptr(stream)))


PASS_NAMES = ("regular", "near", "diagonal", "sweep")


def profile(enable=True):
    """Record per-pass device times with HIP events on the launch stream."""
    check(lib().nb_profile(int(bool(enable))))


def profile_read(reset=True):
    """{pass: (total_ms, launches)} since the last reset (waits for the events)."""
    ms = (C.c_double * 4)()
    n = (C.c_int64 * 4)()
    check(lib().nb_profile_read(ms, n, int(bool(reset))))
    return {PASS_NAMES[p]: (ms[p], n[p]) for p in range(4)}


def device_info():
    name = C.create_string_buffer(64)
    ncu, hbm = C.c_int32(), C.c_int64()
    check(lib().nb_device_info(name, 64, C.byref(ncu), C.byref(hbm)))
    return name.value.decode(), ncu.value, hbm.value
