"""Parity of the HIP engine with the FP64 oracle, through the C ABI.

Bar: pair classes bit-exact; FP64 (near-singular and diagonal) entries within
1e-10 of the row's largest FP64 entry when written as complex128 (acos and log
near their singular points amplify last-bit differences between the device
and host math libraries to ~1e-12); FP32 regular entries within 2e-6 of the
row's largest regular entry; every matrix within
1e-5 relative Frobenius error of the oracle under the same quadrature policy
(FP32 regular pairs).  Meshes are seeded jittered icospheres.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FRO_TOL = 1e-5


@pytest.fixture(scope="module")
def nb():
    import torch
    assert torch.cuda.is_available(), "GPU test without a GPU"
    from nihu_amd import nbem
    name, ncu, hbm = nbem.device_info()
    print("device", name, ncu, "CUs", hbm / 2 ** 30, "GiB", flush=True)
    return nbem


@pytest.fixture(scope="module")
def l3():
    from nihu_amd.mesh import jittered_icosphere
    return jittered_icosphere(3, seed=0)


def relfro(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def to_np(t):
    return t.cpu().numpy()


POLICIES = {
    "nihu": (dict(rule="nihu", near="mean_edge", limit=1.5), dict(rule=0, fixed_degree=0, near_mode=0, near_limit=1.5)),
    "config": (dict(rule="fixed", degree=5, near="diameter", limit=2.0),
               dict(rule=1, fixed_degree=5, near_mode=1, near_limit=2.0)),
}


def test_device_is_gfx950(nb):
    name, ncu, _ = nb.device_info()
    assert name.startswith("gfx950"), name


@pytest.mark.parametrize("pname", ["nihu", "config"])
@pytest.mark.parametrize("nxy", [(0, 0), (0, 1), (1, 1)])
def test_classification_bitwise(nb, orc, l3, pname, nxy):
    nodes, tris = l3
    gp, op = POLICIES[pname]
    m = nb.Mesh(nodes, tris)
    cls = nb.classify(m, m, nxy[0], nxy[1], nb.policy(**gp))
    ref = orc.classify(nodes, tris, nxy[0], nxy[1], orc.policy(**op))
    got = to_np(cls)
    assert np.array_equal(got, ref), np.argwhere(got != ref)[:5]


@pytest.mark.parametrize("pname", ["nihu", "config"])
@pytest.mark.parametrize("fam,k", [(0, 0.0), (1, 4.0), (1, 10.0)])
@pytest.mark.parametrize("nxy", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_single_operator(nb, orc, l3, pname, fam, k, nxy):
    import torch
    nodes, tris = l3
    gp, op = POLICIES[pname]
    n = len(tris)
    m = nb.Mesh(nodes, tris)
    alpha = 0.7 - 0.3j
    out = nb.colmajor(n, n, "complex128")
    nb.assemble(m, m, out, fam, nxy[0], nxy[1], k, alpha=alpha, pol=nb.policy(**gp), accumulate=False)
    torch.cuda.synchronize()
    got = to_np(out)
    ref = orc.assemble(nodes, tris, fam, nxy[0], nxy[1], k, alpha=alpha, pol=orc.policy(**op))
    cls = orc.classify(nodes, tris, nxy[0], nxy[1], orc.policy(**op))
    e = relfro(got, ref)
    print(pname, fam, k, nxy, "rel fro", e, flush=True)
    assert e < FRO_TOL
    # FP64 entries: relative to the largest FP64 entry of the same row (near
    # entries of DLP/DLPt on nearly coplanar pairs are close to zero)
    fp64 = cls < 0
    scale = np.max(np.where(fp64, np.abs(ref), 0.0), axis=1, keepdims=True)
    err64 = np.where(fp64, np.abs(got - ref), 0.0) / scale
    assert err64.max() < 1e-10, err64.max()
    # FP32 entries: absolute error against the largest regular entry of the row
    reg = ~fp64
    rscale = np.max(np.where(reg, np.abs(ref), 0.0), axis=1, keepdims=True)
    err32 = np.where(reg, np.abs(got - ref), 0.0) / rscale
    assert err32.max() < 2e-6, err32.max()


def test_c64_device_output_and_accumulate(nb, orc, l3):
    import torch
    nodes, tris = l3
    n = len(tris)
    m = nb.Mesh(nodes, tris)
    pol = nb.policy(**POLICIES["nihu"][0])
    base = (torch.randn(n, n, dtype=torch.complex64, device="cuda")).t()
    out = base.clone()
    nb.assemble(m, m, out, 1, 0, 1, 3.0, alpha=2.0, pol=pol, accumulate=True)
    nb.assemble(m, m, out, 2, 0, 0, 0.0, alpha=-0.5, pol=pol, accumulate=True)   # identity
    torch.cuda.synchronize()
    ref = orc.assemble(nodes, tris, 1, 0, 1, 3.0, alpha=2.0) - 0.5 * np.eye(n) + to_np(base)
    assert relfro(to_np(out), ref) < FRO_TOL


def test_host_outputs(nb, orc, l3):
    nodes, tris = l3
    n = len(tris)
    m = nb.Mesh(nodes, tris)
    pol = nb.policy(**POLICIES["config"][0])
    H = np.asfortranarray(np.ones((n, n), dtype=np.complex128))
    nb.assemble(m, m, H, 1, 1, 1, 6.0, alpha=1j, pol=pol, accumulate=True)
    ref = 1.0 + orc.assemble(nodes, tris, 1, 1, 1, 6.0, alpha=1j, pol=orc.policy(**POLICIES["config"][1]))
    assert relfro(H, ref) < FRO_TOL
    L = np.asfortranarray(np.zeros((n, n), dtype=np.float64))
    nb.assemble(m, m, L, 0, 0, 0, 0.0, pol=pol, accumulate=False)
    refl = orc.assemble(nodes, tris, 0, 0, 0, 0.0, pol=orc.policy(**POLICIES["config"][1])).real
    assert relfro(L, refl) < FRO_TOL


@pytest.mark.parametrize("pname", ["nihu", "config"])
@pytest.mark.parametrize("k", [1.0, 10.0])
def test_burton_miller(nb, orc, l3, pname, k):
    import torch
    nodes, tris = l3
    gp, op = POLICIES[pname]
    n = len(tris)
    m = nb.Mesh(nodes, tris)
    A, B = nb.colmajor(n, n), nb.colmajor(n, n)
    nb.assemble_bm(m, k, A, B, pol=nb.policy(**gp))
    torch.cuda.synchronize()
    Ar, Br = orc.assemble_bm(nodes, tris, k, pol=orc.policy(**op))
    ea, eb = relfro(to_np(A), Ar), relfro(to_np(B), Br)
    print(pname, k, "A", ea, "B", eb, flush=True)
    assert ea < FRO_TOL and eb < FRO_TOL


def test_field_points_other_mesh(nb, orc, l3):
    """test mesh != trial mesh: no diagonal pass, near pass still applies
    (nihu/core/double_integral.hpp:543-554)."""
    import torch
    from nihu_amd.mesh import icosphere
    nodes, tris = l3
    fn, ft = icosphere(2, radius=1.05, jitter=0.01, seed=9)
    trial = nb.Mesh(nodes, tris)
    test = nb.Mesh(fn, ft)
    out = nb.colmajor(len(ft), len(tris))
    pol = nb.policy(**POLICIES["nihu"][0])
    nb.assemble(test, trial, out, 1, 0, 1, 5.0, pol=pol, accumulate=False)
    torch.cuda.synchronize()
    ref = orc.assemble(fn, ft, 1, 0, 1, 5.0, trial=(nodes, tris))
    cls = orc.classify(fn, ft, 0, 1, trial=(nodes, tris))
    assert (cls == -2).sum() > 0
    assert relfro(to_np(out), ref) < FRO_TOL


def test_error_paths(nb, l3):
    nodes, tris = l3
    m = nb.Mesh(nodes, tris)
    n = len(tris)
    out = nb.colmajor(n, n)
    with pytest.raises(IndexError):
        nb.assemble(m, m, out, 1, 0, 0, 1.0, pol=nb.policy("fixed", 15))
    with pytest.raises(ValueError):
        nb.assemble(m, m, nb.colmajor(n, n - 1), 1, 0, 0, 1.0)
    with pytest.raises(ValueError):
        nb.assemble(m, m, nb.colmajor(n, n, "float32"), 1, 0, 0, 1.0)


@pytest.mark.parametrize("pname", ["config", "nihu"])
def test_headline_workload_rows(nb, orc, pname):
    """20480-triangle jittered icosphere, ka = 10, Burton-Miller A and B: 48
    seeded rows of the full GPU matrices against the oracle.  "config": the
    7-point rule far and FP64 within 2 element diameters; "nihu": the
    distance-driven degrees, whose 1-point rule beyond 22.3 mean edges (the
    centroid fast path of whole far waves) only occurs on meshes this fine."""
    import torch
    from nihu_amd.mesh import jittered_icosphere
    nodes, tris = jittered_icosphere(5, seed=0)
    n = len(tris)
    m = nb.Mesh(nodes, tris)
    A, B = nb.colmajor(n, n), nb.colmajor(n, n)
    # NaN everywhere first: an entry that no pass writes stays NaN
    A.fill_(complex("nan+nanj"))
    B.fill_(complex("nan+nanj"))
    nb.assemble_bm(m, 10.0, A, B, pol=nb.policy(**POLICIES[pname][0]), accumulate=False)
    torch.cuda.synchronize()
    rows = np.sort(np.random.default_rng(1).choice(n, 48, replace=False)).astype(np.int32)
    ri = torch.from_numpy(rows.astype(np.int64)).cuda()
    Ag, Bg = to_np(A[ri, :]), to_np(B[ri, :])
    Ar, Br = orc.assemble_bm(nodes, tris, 10.0, pol=orc.policy(**POLICIES[pname][1]), rows=rows)
    ea, eb = relfro(Ag, Ar), relfro(Bg, Br)
    print("L5 rows (%s): A" % pname, ea, "B", eb, "max abs", np.abs(Ag - Ar).max(), np.abs(Bg - Br).max(),
          flush=True)
    assert ea < FRO_TOL and eb < FRO_TOL
    # every entry was written (the NaN fill is gone everywhere)
    assert torch.isfinite(torch.view_as_real(A)).all() and torch.isfinite(torch.view_as_real(B)).all()


def test_full_size_properties(nb):
    """Size-independent properties on the whole 20480-triangle matrices (no
    oracle needed): Gauss's identity for the Laplace double layer (every row of
    D sums to -1/2 on a closed surface; the FP64 oracle's rows deviate by
    3e-6), the hypersingular operator annihilating constants (oracle: 1e-7 of
    the diagonal), and accumulation doubling every entry exactly."""
    import torch
    from nihu_amd.mesh import jittered_icosphere
    nodes, tris = jittered_icosphere(5, seed=0)
    n = len(tris)
    m = nb.Mesh(nodes, tris)
    pol = nb.policy(**POLICIES["config"][0])
    D = nb.colmajor(n, n, "float32")
    nb.assemble(m, m, D, nb.NB_LAPLACE, 0, 1, 0.0, pol=pol, accumulate=False)
    torch.cuda.synchronize()
    rs = D.to(torch.float64).sum(dim=1)
    dev = (rs + 0.5).abs().max().item()
    del D
    H = nb.colmajor(n, n, "float32")
    nb.assemble(m, m, H, nb.NB_LAPLACE, 1, 1, 0.0, pol=pol, accumulate=False)
    torch.cuda.synchronize()
    hs = H.to(torch.float64).sum(dim=1).abs().max().item()
    hd = torch.diagonal(H).abs().max().item()
    del H
    print("DLP row sums: max |sum + 1/2| %.2e; HSP row sums %.2e of the diagonal" % (dev, hs / hd), flush=True)
    assert dev < 1e-4
    assert hs < 1e-5 * hd
    G = nb.colmajor(n, n)
    nb.assemble(m, m, G, nb.NB_HELMHOLTZ, 0, 0, 10.0, pol=pol, accumulate=False)
    G1 = G.clone()
    nb.assemble(m, m, G, nb.NB_HELMHOLTZ, 0, 0, 10.0, pol=pol, accumulate=True)
    torch.cuda.synchronize()
    assert torch.equal(torch.view_as_real(G), 2 * torch.view_as_real(G1))
