"""Frequency sweep (one persistent launch for many wavenumbers) against the
single-wavenumber Burton-Miller path and the FP64 oracle."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CONFIG = dict(rule="fixed", degree=5, near="diameter", limit=2.0)
CONFIG_O = dict(rule=1, fixed_degree=5, near_mode=1, near_limit=2.0)


@pytest.fixture(scope="module")
def nb():
    import torch
    assert torch.cuda.is_available()
    from nihu_amd import nbem
    return nbem


@pytest.fixture(scope="module")
def l3():
    from nihu_amd.mesh import jittered_icosphere
    return jittered_icosphere(3, seed=2)


def stack(nk, n):
    import torch
    return torch.zeros((nk, n, n), dtype=torch.complex64, device="cuda").transpose(1, 2)


def relfro(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("kb", [2, 4, 8])
def test_sweep_matches_single_and_oracle(nb, orc, l3, kb):
    import torch
    nodes, tris = l3
    n = len(tris)
    m = nb.Mesh(nodes, tris)
    ks = np.array([0.7, 3.3, 10.0, 17.5, 31.9])
    A, B = stack(len(ks), n), stack(len(ks), n)
    nb.assemble_sweep(m, ks, A, B, pol=nb.policy(**CONFIG), launch=nb.NbLaunch(32, kb, 4))
    torch.cuda.synchronize()
    for f, k in enumerate(ks):
        A1, B1 = nb.colmajor(n, n), nb.colmajor(n, n)
        nb.assemble_bm(m, k, A1, B1, pol=nb.policy(**CONFIG))
        torch.cuda.synchronize()
        a, b = A[f].cpu().numpy(), B[f].cpu().numpy()
        assert relfro(a, A1.cpu().numpy()) < 1e-6 and relfro(b, B1.cpu().numpy()) < 1e-6, f
    for f in (0, 4):
        Ar, Br = orc.assemble_bm(nodes, tris, ks[f], pol=orc.policy(**CONFIG_O))
        ea, eb = relfro(A[f].cpu().numpy(), Ar), relfro(B[f].cpu().numpy(), Br)
        print("k", ks[f], "A", ea, "B", eb, flush=True)
        assert ea < 1e-5 and eb < 1e-5


def test_sweep_uniform_wavenumbers(nb, orc, l3):
    """Uniformly spaced wavenumbers (9, not a multiple of the 4-wavenumber
    block): the near-singular pairs rotate e^{-ikr} from one wavenumber to the
    next; every matrix matches its single-wavenumber assembly, the FP64 near
    entries to one complex64 rounding, and the oracle."""
    import torch
    nodes, tris = l3
    n = len(tris)
    m = nb.Mesh(nodes, tris)
    ks = np.linspace(1.0, 17.0, 9)
    A, B = stack(len(ks), n), stack(len(ks), n)
    nb.assemble_sweep(m, ks, A, B, pol=nb.policy(**CONFIG))
    torch.cuda.synchronize()
    cls = nb.classify(m, m, 1, 1, nb.policy(**CONFIG)).cpu().numpy()
    near = cls == -2
    for f, k in enumerate(ks):
        A1, B1 = nb.colmajor(n, n), nb.colmajor(n, n)
        nb.assemble_bm(m, float(k), A1, B1, pol=nb.policy(**CONFIG))
        torch.cuda.synchronize()
        a, b, a1, b1 = A[f].cpu().numpy(), B[f].cpu().numpy(), A1.cpu().numpy(), B1.cpu().numpy()
        assert relfro(a, a1) < 1e-6 and relfro(b, b1) < 1e-6, f
        for x, y in ((a, a1), (b, b1)):
            d = np.abs(x[near] - y[near]) / np.abs(y[near]).max()
            assert d.max() < 2e-7, (f, d.max())
    for f in (0, 8):
        Ar, Br = orc.assemble_bm(nodes, tris, ks[f], pol=orc.policy(**CONFIG_O))
        assert relfro(A[f].cpu().numpy(), Ar) < 1e-5 and relfro(B[f].cpu().numpy(), Br) < 1e-5


@pytest.mark.parametrize("kb", [2, 4, 8])
def test_sweep_nihu_policy(nb, orc, l3, kb):
    """NiHu's distance-driven degrees, five wavenumbers (not a multiple of the
    batch) in one launch: the regular entries equal the single-wavenumber
    assembly's bit for bit (same evaluation, the pair's degrees shared by the
    batch), the FP64 near-singular and diagonal entries to one complex64
    rounding; two wavenumbers against the oracle."""
    import torch
    nodes, tris = l3
    n = len(tris)
    m = nb.Mesh(nodes, tris)
    pol = nb.policy("nihu")
    ks = np.array([2.0, 9.0, 0.6, 23.5, 14.25])
    A, B = stack(len(ks), n), stack(len(ks), n)
    A.fill_(complex("nan+nanj"))
    B.fill_(complex("nan+nanj"))
    nb.assemble_sweep(m, ks, A, B, pol=pol, launch=nb.NbLaunch(32, kb, 4))
    torch.cuda.synchronize()
    cls = nb.classify(m, m, 1, 1, pol).cpu().numpy()
    reg, fp64 = cls >= 0, cls < 0
    assert reg.sum() > 0 and fp64.sum() > n
    for f, k in enumerate(ks):
        A1, B1 = nb.colmajor(n, n), nb.colmajor(n, n)
        nb.assemble_bm(m, float(k), A1, B1, pol=pol)
        torch.cuda.synchronize()
        for x, y in ((A[f].cpu().numpy(), A1.cpu().numpy()), (B[f].cpu().numpy(), B1.cpu().numpy())):
            assert np.array_equal(x[reg], y[reg]), (f, k)
            d = np.abs(x[fp64] - y[fp64]) / np.abs(y[fp64]).max()
            assert d.max() < 2e-7, (f, k, d.max())
    for f in (1, 3):
        Ar, Br = orc.assemble_bm(nodes, tris, float(ks[f]))
        assert relfro(A[f].cpu().numpy(), Ar) < 1e-5 and relfro(B[f].cpu().numpy(), Br) < 1e-5


def test_sweep_nihu_policy_level4_rows(nb, orc):
    """The 5120-triangle jittered icosphere (level 4), four wavenumbers in one
    launch under the distance-driven policy, whose 1-point far rule (beyond
    22.3 mean edges) the level-3 mesh never reaches: 24 seeded rows of each
    matrix against the oracle at the same wavenumber."""
    import torch
    from nihu_amd.mesh import jittered_icosphere
    nodes, tris = jittered_icosphere(4, seed=0)
    n = len(tris)
    m = nb.Mesh(nodes, tris)
    ks = np.array([0.5, 7.0, 19.0, 32.0])
    A, B = stack(len(ks), n), stack(len(ks), n)
    nb.assemble_sweep(m, ks, A, B, pol=nb.policy("nihu"))
    torch.cuda.synchronize()
    rows = np.sort(np.random.default_rng(5).choice(n, 24, replace=False)).astype(np.int32)
    ri = torch.from_numpy(rows.astype(np.int64)).cuda()
    for f, k in enumerate(ks):
        Ar, Br = orc.assemble_bm(nodes, tris, float(k), rows=rows)
        ea = relfro(A[f][ri, :].cpu().numpy(), Ar)
        eb = relfro(B[f][ri, :].cpu().numpy(), Br)
        print("k %.1f: A %.2e B %.2e" % (k, ea, eb), flush=True)
        assert ea < 1e-5 and eb < 1e-5


@pytest.mark.parametrize("rule", ["fixed", "nihu"])
def test_autotune(nb, l3, rule):
    nodes, tris = l3
    m = nb.Mesh(nodes, tris)
    best = nb.autotune(m, 8, nb.policy(**CONFIG) if rule == "fixed" else nb.policy("nihu"))
    print(rule, "tuned", best.tile_cols, best.k_batch, best.persistent, flush=True)
    assert best.tile_cols in (16, 32, 64) and best.k_batch in (2, 4, 8) and best.persistent in (2, 4, 8)


def test_sweep_rejects_bad_arguments(nb, l3):
    nodes, tris = l3
    n = len(tris)
    m = nb.Mesh(nodes, tris)
    A, B = stack(2, n), stack(2, n)
    with pytest.raises(ValueError):
        nb.assemble_sweep(m, np.array([1.0, -2.0]), A, B, pol=nb.policy(**CONFIG))
    with pytest.raises(ValueError):
        nb.assemble_sweep(m, np.array([1.0, 2.0]), A, B, pol=nb.policy(**CONFIG), launch=nb.NbLaunch(32, 3, 4))


def test_config4_sweep_64_wavenumbers(nb, orc):
    """The sweep workload itself: Burton-Miller A and B for 64 wavenumbers ka
    uniformly spaced in [0.5, 32] on the 5120-triangle jittered icosphere
    (level 4, seed 0), 7-point rule far, FP64 within 2 element diameters, in
    one persistent launch with the autotuned launch shape (launch=None).
    Every wavenumber's matrices against its single-wavenumber assembly
    (relative Frobenius 1e-6, computed on the device), and 16 seeded rows of
    8 wavenumbers, the first and last among them, against the oracle's sweep
    (1e-5)."""
    import torch
    from nihu_amd.mesh import jittered_icosphere
    nodes, tris = jittered_icosphere(4, seed=0)
    n = len(tris)
    m = nb.Mesh(nodes, tris)
    ks = np.linspace(0.5, 32.0, 64)
    pol = nb.policy(**CONFIG)
    A, B = stack(len(ks), n), stack(len(ks), n)
    A.fill_(complex("nan+nanj"))
    B.fill_(complex("nan+nanj"))
    nb.assemble_sweep(m, ks, A, B, pol=pol)
    torch.cuda.synchronize()
    assert torch.isfinite(torch.view_as_real(A)).all() and torch.isfinite(torch.view_as_real(B)).all()
    A1, B1 = nb.colmajor(n, n), nb.colmajor(n, n)
    worst = 0.0
    for f, k in enumerate(ks):
        nb.assemble_bm(m, float(k), A1, B1, pol=pol)
        ea = (torch.linalg.norm(A[f] - A1) / torch.linalg.norm(A1)).item()
        eb = (torch.linalg.norm(B[f] - B1) / torch.linalg.norm(B1)).item()
        worst = max(worst, ea, eb)
        assert ea < 1e-6 and eb < 1e-6, (f, k, ea, eb)
    print("64 wavenumbers: max rel. Frobenius difference to single-wavenumber assembly %.2e" % worst, flush=True)
    rows = np.sort(np.random.default_rng(44).choice(n, 16, replace=False)).astype(np.int32)
    fsel = [0, 9, 18, 27, 36, 45, 54, 63]
    Ar, Br = orc.sweep_bm(nodes, tris, ks[fsel], pol=orc.policy(**CONFIG_O), rows=rows)
    ri = torch.from_numpy(rows.astype(np.int64)).cuda()
    for t, f in enumerate(fsel):
        ea = relfro(A[f][ri, :].cpu().numpy(), Ar[t])
        eb = relfro(B[f][ri, :].cpu().numpy(), Br[t])
        print("ka %.2f: A %.2e B %.2e" % (ks[f], ea, eb), flush=True)
        assert ea < 1e-5 and eb < 1e-5
