"""Meshes holding both element types (NiHu's mesh over tria_1 and quad_1,
built from its element table: element id = table row), in collocation with
constant fields, through the C ABI, against the FP64 oracle.

The expected matrix is put together from the oracle's single-type
assemblies by NiHu's rules for such a mesh: a type against itself is a
same-mesh assembly (diagonal pairs included); a type against the other one
pairs different elements, for which NiHu's same-mesh loop finds no element
match and does what its different-mesh loop does (near-singular check, else
the regular rule); entries land at the elements' table rows.  The surface is
the jittered cubed sphere with every third quadrilateral split into two
triangles, the two types interleaved in the table.

Bars as for single-type meshes: rows to 2e-6 of their largest entry and
1e-5 relative Frobenius.  Also: the Burton-Miller pair against its single
operators, host outputs with accumulation, a permutation-sensitive check that
needs no oracle (Gauss's identity for the Laplace double layer and the
transparent point-source test on the closed surface), the sweep (one
Burton-Miller assembly per wavenumber) and the refusals (Galerkin, sweep
autotuning, pair classes, near-singular rules beyond d / R = 1.5).
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROW_TOL = 2e-6
NIHU = (dict(rule="nihu", near="mean_edge", limit=1.5), dict(rule=0, fixed_degree=0, near_mode=0, near_limit=1.5))
TRIA, QUAD = 32303, 32404


@pytest.fixture(scope="module")
def nb():
    import torch
    assert torch.cuda.is_available(), "GPU test without a GPU"
    from nihu_amd import nbem
    return nbem


@pytest.fixture(scope="module")
def mixed():
    from nihu_amd.mesh import mixed_cubed_sphere
    return mixed_cubed_sphere(5, seed=3)


def rowscaled(got, ref):
    scale = np.abs(ref).max(axis=1, keepdims=True)
    scale[scale == 0] = 1.0
    return (np.abs(got - ref) / scale).max()


def relfro(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def oracle_mixed(orc, nodes, table, fam, nx, ny, k, opol):
    parts = [(np.flatnonzero(table[:, 0] == TRIA), 3), (np.flatnonzero(table[:, 0] == QUAD), 4)]
    n = len(table)
    out = np.zeros((n, n), dtype=np.complex128)
    for a, (ia, na) in enumerate(parts):
        for b, (ib, nb_) in enumerate(parts):
            ea = table[ia, 1:1 + na]
            eb = table[ib, 1:1 + nb_]
            blk = orc.assemble(nodes, ea, fam, nx, ny, k=k, pol=orc.policy(**opol),
                               trial=None if a == b else (nodes, eb))
            out[np.ix_(ia, ib)] = blk
    return out


def test_mixed_mesh_handle(nb, mixed):
    nodes, table = mixed
    m = nb.Mesh.from_elements(nodes, table)
    assert m.nodes_per_element == 0 and m.n == len(table)
    # a table of one type gives that type's plain mesh
    tri_only = nb.Mesh.from_elements(nodes[:3] * 0 + np.eye(3), np.array([[TRIA, 0, 1, 2]]))
    assert tri_only.nodes_per_element == 3 and tri_only.n == 1


@pytest.mark.parametrize("fam,nx,ny,k", [(1, 0, 0, 3.0), (1, 0, 1, 3.0), (1, 1, 0, 3.0), (1, 1, 1, 3.0),
                                         (0, 0, 0, 0.0), (0, 1, 1, 0.0)])
def test_mixed_operators(nb, orc, mixed, fam, nx, ny, k):
    import torch
    nodes, table = mixed
    m = nb.Mesh.from_elements(nodes, table)
    n = len(table)
    out = nb.colmajor(n, n, "complex128")
    out.fill_(complex("nan+nanj"))   # an entry no block writes stays NaN
    nb.assemble(m, m, out, fam, nx, ny, k, pol=nb.policy(**NIHU[0]), accumulate=False)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    assert np.isfinite(got).all()
    ref = oracle_mixed(orc, nodes, table, fam, nx, ny, k, NIHU[1])
    rs, fro = rowscaled(got, ref), relfro(got, ref)
    print("mixed %d elements fam %d (%d, %d): row-scaled %.2e rel fro %.2e" % (n, fam, nx, ny, rs, fro), flush=True)
    assert rs <= ROW_TOL and fro <= 1e-5


def test_mixed_field_points(nb, orc, mixed):
    """Mixed trial mesh, separate triangle test mesh (field points): no
    diagonal, near-singular pairs where close."""
    import torch
    from nihu_amd.mesh import jittered_icosphere
    nodes, table = mixed
    m = nb.Mesh.from_elements(nodes, table)
    fn, ft = jittered_icosphere(2, seed=9)
    fn = fn * 1.04
    f = nb.Mesh(fn, ft)
    out = nb.colmajor(len(ft), len(table), "complex128")
    nb.assemble(f, m, out, 1, 0, 1, 2.5, pol=nb.policy(**NIHU[0]), accumulate=False)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    ref = np.zeros_like(got)
    for tid, npe in ((TRIA, 3), (QUAD, 4)):
        ib = np.flatnonzero(table[:, 0] == tid)
        ref[:, ib] = orc.assemble(fn, ft, 1, 0, 1, k=2.5, pol=orc.policy(**NIHU[1]),
                                  trial=(nodes, table[ib, 1:1 + npe]))
    rs = rowscaled(got, ref)
    print("field points on a mixed mesh: row-scaled %.2e" % rs, flush=True)
    assert rs <= ROW_TOL


def test_mixed_burton_miller_and_host_accumulate(nb, mixed):
    """The fused Burton-Miller pair equals its composition from single
    operators on the same mixed mesh; host outputs accumulate."""
    import torch
    nodes, table = mixed
    m = nb.Mesh.from_elements(nodes, table)
    n = len(table)
    k = 4.0
    al = complex(0.0, -1.0 / k)
    pol = nb.policy(**NIHU[0])
    A, B = nb.colmajor(n, n, "complex128"), nb.colmajor(n, n, "complex128")
    nb.assemble_bm(m, k, A, B, pol=pol)
    ops = {}
    for nx, ny in ((0, 0), (0, 1), (1, 0), (1, 1)):
        o = nb.colmajor(n, n, "complex128")
        nb.assemble(m, m, o, 1, nx, ny, k, pol=pol, accumulate=False)
        ops[(nx, ny)] = o
    torch.cuda.synchronize()
    I = np.eye(n)
    Ac = al * ops[(1, 1)].cpu().numpy() + ops[(0, 1)].cpu().numpy() - 0.5 * I
    Bc = ops[(0, 0)].cpu().numpy() + al * ops[(1, 0)].cpu().numpy() + 0.5 * al * I
    ea, eb = relfro(A.cpu().numpy(), Ac), relfro(B.cpu().numpy(), Bc)
    print("mixed Burton-Miller vs composition: A %.2e B %.2e" % (ea, eb), flush=True)
    assert ea < 1e-6 and eb < 1e-6
    # host output, accumulating onto a filled matrix
    H = np.asfortranarray(np.full((n, n), 1.0 + 2.0j))
    nb.assemble(m, m, H, 1, 0, 0, k, pol=pol, accumulate=True)
    assert np.abs(H - (1.0 + 2.0j) - ops[(0, 0)].cpu().numpy()).max() < 1e-12


def test_mixed_identities(nb, mixed):
    """Without the oracle: Gauss's identity (Laplace double-layer rows sum to
    -1/2 on a closed surface) and the transparent point-source test
    ((D - I/2) p = G q for sources inside), which fails at once if a block
    lands on the wrong rows or columns."""
    import torch
    from nihu_amd.serve import _centres_normals
    nodes, table = mixed
    m = nb.Mesh.from_elements(nodes, table)
    n = len(table)
    pol = nb.policy(**NIHU[0])
    D = nb.colmajor(n, n, "float64")
    nb.assemble(m, m, D, 0, 0, 1, 0.0, pol=pol, accumulate=False)
    torch.cuda.synchronize()
    Dn = D.cpu().numpy()
    gauss = np.abs(Dn.sum(axis=1) + 0.5).max()
    # collocation points and normals of each element type
    c = np.zeros((n, 3))
    nu = np.zeros((n, 3))
    for tid, npe in ((TRIA, 3), (QUAD, 4)):
        ib = np.flatnonzero(table[:, 0] == tid)
        c[ib], nu[ib] = _centres_normals(nodes, table[ib, 1:1 + npe])
    k = 1.5
    G, H = nb.colmajor(n, n, "complex128"), nb.colmajor(n, n, "complex128")
    nb.assemble(m, m, G, 1, 0, 0, k, pol=pol, accumulate=False)
    nb.assemble(m, m, H, 1, 0, 1, k, pol=pol, accumulate=False)
    torch.cuda.synchronize()
    x0 = np.array([0.2, 0.1, -0.15])
    rv = c - x0
    r = np.linalg.norm(rv, axis=1)
    p0 = np.exp(-1j * k * r) / (4 * np.pi * r)
    q0 = -p0 * (1 + 1j * k * r) / r * np.einsum("ij,ij->i", rv, nu) / r
    p = np.linalg.solve(H.cpu().numpy() - 0.5 * np.eye(n), G.cpu().numpy() @ q0)
    et = np.linalg.norm(p - p0) / np.linalg.norm(p0)
    print("mixed mesh: Gauss max |row sum + 1/2| %.2e, transparent error %.2e" % (gauss, et), flush=True)
    assert gauss < 1e-3 and et < 0.05


def test_mixed_refusals(nb, mixed):
    import torch
    nodes, table = mixed
    m = nb.Mesh.from_elements(nodes, table)
    n = len(table)
    out = nb.colmajor(n, n)
    with pytest.raises(NotImplementedError):
        nb.assemble(m, m, out, 1, 0, 0, 1.0, formalism="galerkin")
    with pytest.raises(NotImplementedError):
        nb.assemble(m, m, out, 1, 0, 0, 1.0, pol=nb.policy(rule="fixed", degree=5, near="diameter", limit=2.0))
    with pytest.raises(NotImplementedError):
        nb.classify(m, m, 0, 0)
    with pytest.raises(NotImplementedError):
        nb.autotune(m, 8)


def test_mixed_sweep(nb, mixed):
    """The sweep on a mixed mesh: one Burton-Miller assembly per wavenumber,
    equal to nb_assemble_bm at each."""
    import torch
    nodes, table = mixed
    m = nb.Mesh.from_elements(nodes, table)
    n = m.n
    ks = np.array([1.5, 7.0])
    A = torch.zeros((2, n, n), dtype=torch.complex64, device="cuda").transpose(1, 2)
    B = torch.zeros_like(A)
    nb.assemble_sweep(m, ks, A, B, pol=nb.policy(**NIHU[0]))
    for f, k in enumerate(ks):
        A1, B1 = nb.colmajor(n, n), nb.colmajor(n, n)
        nb.assemble_bm(m, float(k), A1, B1, pol=nb.policy(**NIHU[0]))
        torch.cuda.synchronize()
        assert torch.equal(A[f], A1) and torch.equal(B[f], B1)


def test_mixed_several_row_tiles(nb, orc):
    """1152 interleaved elements (576 triangles, 576 quadrilaterals: five
    256-row tiles of the all-rows regular pass): seeded sampled rows of the
    Burton-Miller A and B and of the Laplace double layer against the oracle's
    single-type assemblies placed by NiHu's rules, with every entry of the
    device matrices written (NaN fill)."""
    import torch
    from nihu_amd.mesh import mixed_cubed_sphere
    nodes, table = mixed_cubed_sphere(12, seed=4)
    n = len(table)
    assert n == 1152
    m = nb.Mesh.from_elements(nodes, table)
    k = 6.0
    al = complex(0.0, -1.0 / k)
    pol = nb.policy(**NIHU[0])
    A, B = nb.colmajor(n, n), nb.colmajor(n, n)
    D = nb.colmajor(n, n, "float64")
    for t in (A, B, D):
        t.fill_(float("nan"))
    nb.assemble_bm(m, k, A, B, pol=pol, accumulate=False)
    nb.assemble(m, m, D, nb.NB_LAPLACE, 0, 1, 0.0, pol=pol, accumulate=False)
    torch.cuda.synchronize()
    for t in (A, B, D):
        assert torch.isfinite(torch.view_as_real(t) if t.is_complex() else t).all()
    rows = np.sort(np.random.default_rng(8).choice(n, 96, replace=False))
    parts = [(np.flatnonzero(table[:, 0] == TRIA), 3), (np.flatnonzero(table[:, 0] == QUAD), 4)]
    opol = orc.policy(**NIHU[1])
    ops = {key: np.zeros((len(rows), n), dtype=np.complex128) for key in ((0, 0), (0, 1), (1, 0), (1, 1), "lap")}
    for a_, (ia, na) in enumerate(parts):
        sel = np.flatnonzero(np.isin(rows, ia))
        if sel.size == 0:
            continue
        loc = np.searchsorted(ia, rows[sel]).astype(np.int32)
        for b_, (ib, nb_) in enumerate(parts):
            trial = None if a_ == b_ else (nodes, table[ib, 1:1 + nb_])
            for nx, ny in ((0, 0), (0, 1), (1, 0), (1, 1)):
                ops[(nx, ny)][np.ix_(sel, ib)] = orc.assemble(nodes, table[ia, 1:1 + na], 1, nx, ny, k=k, pol=opol,
                                                              rows=loc, trial=trial)
            ops["lap"][np.ix_(sel, ib)] = orc.assemble(nodes, table[ia, 1:1 + na], 0, 0, 1, pol=opol, rows=loc,
                                                       trial=trial)
    Ar = al * ops[(1, 1)] + ops[(0, 1)]
    Br = ops[(0, 0)] + al * ops[(1, 0)]
    Ar[np.arange(len(rows)), rows] -= 0.5
    Br[np.arange(len(rows)), rows] += 0.5 * al
    ri = torch.from_numpy(rows.astype(np.int64)).cuda()
    Ag, Bg, Dg = A[ri, :].cpu().numpy(), B[ri, :].cpu().numpy(), D[ri, :].cpu().numpy()
    errs = (relfro(Ag, Ar), relfro(Bg, Br), relfro(Dg, ops["lap"].real))
    print("mixed 1152 elements, 96 rows: A %.2e B %.2e Laplace DLP %.2e" % errs, flush=True)
    assert max(errs) < 1e-5
    assert rowscaled(Ag, Ar) <= ROW_TOL and rowscaled(Bg, Br) <= ROW_TOL and rowscaled(Dg, ops["lap"].real) <= ROW_TOL
