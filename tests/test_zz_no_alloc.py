"""nb_assemble_sweep_noalloc, the allocation-free variant of
nb_assemble_sweep (per-call buffers in the library's static scratch region),
through the C ABI: two calls in a row, the second on a smaller mesh with fewer
wavenumbers, each equal bit for bit to nb_assemble_sweep on the same inputs
(a fixed 7-point rule and NiHu's distance-driven degrees);
requests beyond the region and meshes without resident tables are refused
before anything runs."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CONFIG = dict(rule="fixed", degree=5, near="diameter", limit=2.0)


@pytest.fixture(scope="module")
def nb():
    import torch
    assert torch.cuda.is_available(), "GPU test without a GPU"
    from nihu_amd import nbem
    return nbem


def _stack(nk, n, fill=0.0):
    import torch
    t = torch.empty((nk, n, n), dtype=torch.complex64, device="cuda").transpose(1, 2)
    t.fill_(fill)
    return t


@pytest.mark.parametrize("rule", ["fixed", "nihu"])
def test_noalloc_sweep_equals_sweep(nb, rule):
    import torch
    from nihu_amd.mesh import jittered_icosphere
    pol = nb.policy(**CONFIG) if rule == "fixed" else nb.policy("nihu")
    cases = []
    # first call: 1280 triangles, 8 uniformly spaced wavenumbers (the near
    # pairs' rotation kernel); second: 320 triangles, 3 scattered wavenumbers
    for level, ks in ((3, np.linspace(1.0, 22.0, 8)), (2, np.array([2.5, 0.75, 13.0]))):
        nodes, tris = jittered_icosphere(level, seed=6)
        m = nb.Mesh(nodes, tris)
        n = len(tris)
        Ar, Br = _stack(len(ks), n), _stack(len(ks), n)
        nb.assemble_sweep(m, ks, Ar, Br, pol=pol)       # builds the mesh's tables; reference result
        torch.cuda.synchronize()
        cases.append((m, ks, Ar, Br))
    for m, ks, Ar, Br in cases:
        n = m.n
        print("allocation-free sweep (%s): %d triangles, %d wavenumbers" % (rule, n, len(ks)), flush=True)
        A, B = _stack(len(ks), n, complex("nan+nanj")), _stack(len(ks), n, complex("nan+nanj"))
        nb.assemble_sweep(m, ks, A, B, pol=pol, noalloc=True)
        torch.cuda.synchronize()
        assert torch.equal(torch.view_as_real(A), torch.view_as_real(Ar))
        assert torch.equal(torch.view_as_real(B), torch.view_as_real(Br))


def test_noalloc_sweep_refusals(nb):
    import torch
    from nihu_amd.mesh import jittered_icosphere
    nodes, tris = jittered_icosphere(2, seed=8)
    n = len(tris)
    m = nb.Mesh(nodes, tris)
    pol = nb.policy(**CONFIG)
    A, B = _stack(2, n), _stack(2, n)
    # no tables resident for this mesh yet
    with pytest.raises(NotImplementedError):
        nb.assemble_sweep(m, np.array([1.0, 2.0]), A, B, pol=pol, noalloc=True)
    # nor those of NiHu's degrees
    with pytest.raises(NotImplementedError):
        nb.assemble_sweep(m, np.array([1.0, 2.0]), A, B, pol=nb.policy("nihu"), noalloc=True)
    # every pair near-singular (limit 100 diameters): 320^2 FP64 pairs of static
    # parts fit; on 1280 triangles (1.6M pairs, 52 MB) they do not
    nodes3, tris3 = jittered_icosphere(3, seed=8)
    m3 = nb.Mesh(nodes3, tris3)
    wide = nb.policy(rule="fixed", degree=5, near="diameter", limit=100.0)
    n3 = len(tris3)
    A3, B3 = _stack(2, n3), _stack(2, n3)
    nb.assemble_sweep(m3, np.array([1.0, 3.0]), A3, B3, pol=wide)
    torch.cuda.synchronize()
    with pytest.raises(IndexError, match="scratch"):
        nb.assemble_sweep(m3, np.array([1.0, 3.0]), A3, B3, pol=wide, noalloc=True)
