"""Parity of the Galerkin passes on quadrilaterals with the FP64 oracle,
through the C ABI: constant and isoparametric (bilinear, nodal) fields.

Bar, as for triangles: with constant fields the entries of node-sharing
pairs come from the FP64 pass alone and agree within 1e-12 of the row's
largest entry (complex128 output); every matrix (FP32 regular pairs plus FP64
singular pairs) agrees within 2e-6 of the row's largest entry and within 1e-5
relative Frobenius error of the oracle under the same quadrature policy; the
mass matrix to 1e-14.  A larger mesh (864 quadrilaterals, many wave tiles
and trial chunks) is compared on whole rows of sampled dofs.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROW_TOL = 2e-6
FRO_TOL = 1e-5
FORMS = {"galerkin": False, "galerkin_linear": True}
POLICIES = {
    "nihu": (dict(rule="nihu", near="mean_edge", limit=1.5), dict(rule=0, fixed_degree=0, near_mode=0, near_limit=1.5)),
    "fixed5": (dict(rule="fixed", degree=5, near="none"), dict(rule=1, fixed_degree=5, near_mode=2, near_limit=1.5)),
}


@pytest.fixture(scope="module")
def nb():
    import torch
    assert torch.cuda.is_available(), "GPU test without a GPU"
    from nihu_amd import nbem
    return nbem


@pytest.fixture(scope="module")
def q4():
    from nihu_amd.mesh import jittered_cubed_sphere
    return jittered_cubed_sphere(4, seed=4)


def rowscaled(got, ref):
    scale = np.abs(ref).max(axis=1, keepdims=True)
    scale[scale == 0] = 1.0
    return (np.abs(got - ref) / scale).max()


def relfro(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _assemble(nb, m_test, m_trial, form, fam, nx, ny, k, alpha, gp, dtype="complex128", same=None):
    import torch
    out = nb.colmajor(m_test.dofs(form), m_trial.dofs(form), dtype)
    nb.assemble(m_test, m_trial, out, fam, nx, ny, k, alpha=alpha, pol=nb.policy(**gp), accumulate=False,
                formalism=form, same_mesh=same)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def _singular_mask(quads):
    n = len(quads)
    sets = [set(q) for q in quads]
    return np.array([[bool(sets[e] & sets[f]) for f in range(n)] for e in range(n)])


@pytest.mark.parametrize("form", ["galerkin", "galerkin_linear"])
@pytest.mark.parametrize("pname", ["nihu", "fixed5"])
@pytest.mark.parametrize("fam,k", [(0, 0.0), (1, 3.0), (1, 10.0), (1, complex(3.0, -0.5))])
@pytest.mark.parametrize("nxy", [(0, 0), (0, 1), (1, 0)])
def test_galerkin_quad_operator(nb, orc, q4, form, pname, fam, k, nxy):
    nodes, quads = q4
    gp, op = POLICIES[pname]
    alpha = 0.8 + 0.25j
    m = nb.Mesh(nodes, quads)
    print("case", form, pname, fam, k, nxy, flush=True)
    got = _assemble(nb, m, m, form, fam, nxy[0], nxy[1], k, alpha, gp)
    ref = orc.assemble_galerkin(nodes, quads, fam, nxy[0], nxy[1], k=k, alpha=alpha, pol=orc.policy(**op),
                                linear=FORMS[form])
    rs, rf = rowscaled(got, ref), relfro(got, ref)
    print("row-scaled %.2e  rel fro %.2e" % (rs, rf), flush=True)
    assert rs <= ROW_TOL and rf <= FRO_TOL
    if form == "galerkin":
        mask = _singular_mask(quads)
        scale = np.abs(ref).max(axis=1, keepdims=True)
        err = (np.abs(got - ref) / scale)[mask].max()
        print("singular entries %.2e" % err, flush=True)
        assert err <= 1e-12


@pytest.mark.parametrize("form", ["galerkin", "galerkin_linear"])
@pytest.mark.parametrize("nxy", [(0, 0), (1, 1)])
def test_other_mesh(nb, orc, q4, form, nxy):
    """Different test and trial meshes: no singular check, so the hypersingular
    kernel assembles too."""
    from nihu_amd.mesh import cubed_sphere
    nodes, quads = q4
    fn, fq = cubed_sphere(3)
    fn = fn * 1.7 + np.array([0.2, -0.1, 0.3])
    gp, op = POLICIES["nihu"]
    mt, ms = nb.Mesh(nodes, quads), nb.Mesh(fn, fq)
    print("other mesh", form, nxy, flush=True)
    got = _assemble(nb, mt, ms, form, 1, nxy[0], nxy[1], 2.5, 1.0, gp)
    ref = orc.assemble_galerkin(nodes, quads, 1, nxy[0], nxy[1], k=2.5, pol=orc.policy(**op), trial=(fn, fq),
                                linear=FORMS[form])
    assert rowscaled(got, ref) <= ROW_TOL and relfro(got, ref) <= FRO_TOL


@pytest.mark.parametrize("form", ["galerkin", "galerkin_linear"])
def test_hypersingular_same_mesh_raises(nb, q4, form):
    nodes, quads = q4
    m = nb.Mesh(nodes, quads)
    out = nb.colmajor(m.dofs(form), m.dofs(form), "complex64")
    with pytest.raises(nb.NbemError, match="Invalid general singular quadrature"):
        nb.assemble(m, m, out, 1, 1, 1, 1.0, formalism=form)


@pytest.mark.parametrize("form", ["galerkin", "galerkin_linear"])
def test_mass_matrix(nb, orc, q4, form):
    nodes, quads = q4
    m = nb.Mesh(nodes, quads)
    got = _assemble(nb, m, m, form, nb.NB_IDENTITY, 0, 0, 0.0, -0.5, POLICIES["nihu"][0])
    ref = orc.assemble_galerkin(nodes, quads, 2, 0, 0, alpha=-0.5, linear=FORMS[form])
    assert np.abs(got - ref).max() <= 1e-14 * np.abs(ref).max()


@pytest.mark.parametrize("form", ["galerkin", "galerkin_linear"])
def test_accumulate_host_and_complex64(nb, orc, q4, form):
    """Host (numpy) output with accumulate: result = previous + alpha M; the
    complex64 device output within FP32 rounding of complex128."""
    nodes, quads = q4
    gp, op = POLICIES["fixed5"]
    m = nb.Mesh(nodes, quads)
    rows = m.dofs(form)
    rng = np.random.default_rng(2)
    host = np.asfortranarray(rng.standard_normal((rows, rows)) + 1j * rng.standard_normal((rows, rows)))
    start = host.copy()
    nb.assemble(m, m, host, 1, 0, 1, 4.0, alpha=-1j, pol=nb.policy(**gp), accumulate=True, formalism=form)
    ref = orc.assemble_galerkin(nodes, quads, 1, 0, 1, k=4.0, alpha=-1j, pol=orc.policy(**op), linear=FORMS[form])
    assert rowscaled(host - start, ref) <= ROW_TOL
    c64 = _assemble(nb, m, m, form, 1, 0, 0, 4.0, 1.0, gp, dtype="complex64")
    c128 = _assemble(nb, m, m, form, 1, 0, 0, 4.0, 1.0, gp, dtype="complex128")
    assert rowscaled(c64.astype(np.complex128), c128) <= 1e-6


@pytest.mark.parametrize("form", ["galerkin", "galerkin_linear"])
def test_real_outputs(nb, orc, q4, form):
    """Laplace double layer into float64 and float32 device outputs (real
    results): float64 within the bar of the complex outputs, float32 within
    FP32 rounding of it."""
    nodes, quads = q4
    gp, op = POLICIES["nihu"]
    m = nb.Mesh(nodes, quads)
    ref = orc.assemble_galerkin(nodes, quads, 0, 0, 1, k=0.0, alpha=-0.7, pol=orc.policy(**op),
                                linear=FORMS[form]).real
    f64 = _assemble(nb, m, m, form, 0, 0, 1, 0.0, -0.7, gp, dtype="float64")
    f32 = _assemble(nb, m, m, form, 0, 0, 1, 0.0, -0.7, gp, dtype="float32")
    assert rowscaled(f64, ref) <= ROW_TOL and relfro(f64, ref) <= FRO_TOL
    assert rowscaled(f32.astype(np.float64), f64) <= 1e-6


@pytest.mark.parametrize("form", ["galerkin", "galerkin_linear"])
def test_many_tiles_sampled_rows(nb, orc, form):
    """864 quadrilaterals (866 nodes): many wave tiles and trial chunks, the
    column-slot evictions of the nodal kernel; whole rows of 12 sampled dofs
    (the oracle integrates only the test elements around them)."""
    from nihu_amd.mesh import jittered_cubed_sphere
    nodes, quads = jittered_cubed_sphere(12, seed=7)
    m = nb.Mesh(nodes, quads)
    gp, op = POLICIES["nihu"]
    rng = np.random.default_rng(11)
    linear = FORMS[form]
    n_dof = len(nodes) if linear else len(quads)
    rows = np.sort(rng.choice(n_dof, 12, replace=False))
    if linear:
        elems = np.unique(np.nonzero(np.isin(quads, rows).any(axis=1))[0])
    else:
        elems = rows
    for nxy, k in (((0, 0), 6.0), ((0, 1), complex(6.0, -0.4))):
        print("864 quads", form, nxy, k, flush=True)
        got = _assemble(nb, m, m, form, 1, nxy[0], nxy[1], k, 1.0, gp)[rows]
        ref = orc.assemble_galerkin(nodes, quads, 1, nxy[0], nxy[1], k=k, pol=orc.policy(**op), linear=linear,
                                    test_elems=elems)[rows]
        rs, rf = rowscaled(got, ref), relfro(got, ref)
        print("row-scaled %.2e  rel fro %.2e" % (rs, rf), flush=True)
        assert rs <= ROW_TOL and rf <= FRO_TOL


@pytest.mark.parametrize("form", ["galerkin", "galerkin_linear"])
def test_pulsating_sphere(nb, form):
    """(M - I/2) p = L q with GPU Galerkin matrices on NiHu's blown-up cube
    sphere: p = -q / (1 + i k) for the pulsating unit sphere."""
    from nihu_amd.mesh import cubed_sphere
    nodes, quads = cubed_sphere(6)
    m = nb.Mesh(nodes, quads)
    gp = POLICIES["nihu"][0]
    k = 1.0
    L = _assemble(nb, m, m, form, 1, 0, 0, k, 1.0, gp)
    M = _assemble(nb, m, m, form, 1, 0, 1, k, 1.0, gp)
    I = _assemble(nb, m, m, form, nb.NB_IDENTITY, 0, 0, 0.0, 1.0, gp)
    p = np.linalg.solve(M - 0.5 * I, L @ np.ones(m.dofs(form)))
    panal = -1.0 / (1.0 + 1j * k)
    err = np.abs(p - panal).max() / abs(panal)
    print(form, "pulsating sphere error %.2e" % err, flush=True)
    assert err < 0.02


def test_mixed_element_types_refused(nb):
    """Galerkin pairs of a triangle and a quadrilateral are not built."""
    from nihu_amd.mesh import cubed_sphere, icosphere
    qn, qq = cubed_sphere(2)
    tn, tt = icosphere(1)
    mq, mt = nb.Mesh(qn, qq), nb.Mesh(tn, tt)
    out = nb.colmajor(mt.n, mq.n)
    with pytest.raises(NotImplementedError, match="one element type"):
        nb.assemble(mt, mq, out, 1, 0, 0, 1.0, formalism="galerkin")
