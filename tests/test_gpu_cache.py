"""The per-mesh cache of near-singular pair lists (one per near policy, at most
eight per trial mesh, the oldest evicted first): the mesh's device byte count
follows every build and eviction exactly (8 bytes per listed pair), and an
assembly under an evicted policy rebuilds its list and gives the same matrix
bit for bit."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_near_list_eviction_and_byte_count():
    import torch
    from nihu_amd import nbem
    from nihu_amd.mesh import jittered_icosphere
    nodes, tris = jittered_icosphere(2, seed=9)
    n = len(tris)
    m = nbem.Mesh(nodes, tris)
    limits = [1.0 + 0.25 * i for i in range(10)]
    pols = [nbem.policy(rule="fixed", degree=5, near="diameter", limit=lim) for lim in limits]
    k = 3.0

    def assemble(pol):
        A, B = nbem.colmajor(n, n), nbem.colmajor(n, n)
        nbem.assemble_bm(m, k, A, B, pol=pol)
        torch.cuda.synchronize()
        return A, B

    def n_near(pol):
        return int((nbem.classify(m, m, 1, 1, pol) == -2).sum().item())

    first = assemble(pols[0])
    cached = [pols[0]]
    for pol in pols[1:]:
        before = m.device_bytes()
        assemble(pol)
        expect = before + 8 * n_near(pol)
        if len(cached) == 8:
            expect -= 8 * n_near(cached.pop(0))
        cached.append(pol)
        print("limit %.2f: %d near pairs, %d bytes" % (pol.near_limit, n_near(pol), m.device_bytes()), flush=True)
        assert m.device_bytes() == expect
    # pols[0] and pols[1] were evicted: rebuilding gives the same matrices
    again = assemble(pols[0])
    assert torch.equal(torch.view_as_real(first[0]), torch.view_as_real(again[0]))
    assert torch.equal(torch.view_as_real(first[1]), torch.view_as_real(again[1]))
