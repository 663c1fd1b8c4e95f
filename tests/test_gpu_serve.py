"""Serving loop: physics (rigid-sphere Mie series), same result alone and in a
mix of concurrent requests, admission by HBM footprint."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def meshes():
    import torch
    assert torch.cuda.is_available()
    from nihu_amd.mesh import jittered_icosphere
    return {"L2": jittered_icosphere(2, seed=3), "L3": jittered_icosphere(3, seed=0), "L4": jittered_icosphere(4, seed=0)}


def test_mie_rigid_sphere():
    """Plane wave on a rigid unit sphere: total surface pressure against the
    Mie series (benchmark/helmholtz/plane_wave_scatter_from_sphere.m:56-108
    is the reference's version of this check)."""
    from nihu_amd import serve
    from nihu_amd.mesh import icosphere
    nodes, tris = icosphere(4)
    srv = serve.Server({"S": (nodes, tris)}, slots=1)
    k = 3.0
    d = np.array([0.0, 0.0, 1.0])
    res = srv.run([serve.Request(0, "S", k, tuple(d))], realtime=False)[0]
    c, _ = serve._centres_normals(nodes, tris)
    pinc = np.exp(-1j * k * (c @ d))
    ptot = pinc + res.p
    rc = np.linalg.norm(c, axis=1)
    mie = serve.mie_rigid_sphere_surface(k, 1.0, (c @ d) / rc)
    err = np.linalg.norm(ptot - mie) / np.linalg.norm(mie)
    print("Mie surface error", err, flush=True)
    assert err < 0.03


def test_same_alone_and_in_mix(meshes):
    from nihu_amd import serve
    srv = serve.Server(meshes, slots=4)
    reqs = serve.poisson_stream(12, rate=1e9, seed=7, meshes=("L2", "L3", "L4"), k_range=(1.0, 40.0))
    mixed = {r.rid: r for r in srv.run(reqs, realtime=False)}
    assert len(mixed) == len(reqs)
    # the mix really overlapped: some request was admitted before an earlier one finished
    spans = sorted((r.admitted, r.done) for r in mixed.values())
    assert any(spans[i + 1][0] < spans[i][1] for i in range(len(spans) - 1))
    for rq in reqs[:6]:
        alone = srv.run([serve.Request(rq.rid, rq.mesh, rq.k, rq.direction)], realtime=False)[0]
        a, b = alone.p, mixed[rq.rid].p
        assert np.all(np.isfinite(a))
        diff = np.abs(a - b).max() / np.abs(a).max()
        print("request", rq.rid, rq.mesh, "k=%.2f" % rq.k, "alone vs mix", diff, flush=True)
        assert diff <= 1e-6


def test_admission_by_footprint(meshes):
    from nihu_amd import serve
    n4 = len(meshes["L4"][1])
    budget = 1.5 * serve.footprint_bytes(n4)
    srv = serve.Server({"L4": meshes["L4"]}, slots=4, hbm_budget=budget)
    reqs = [serve.Request(i, "L4", 2.0 + i, (1.0, 0.0, 0.0)) for i in range(3)]
    res = sorted(srv.run(reqs, realtime=False), key=lambda r: r.admitted)
    # only one L4 request fits: each is admitted after the previous one completed
    for a, b in zip(res, res[1:]):
        assert b.admitted >= a.done - 1e-4


def test_oversized_mesh_rejected(meshes):
    from nihu_amd import serve
    with pytest.raises(ValueError):
        serve.Server({"L4": meshes["L4"]}, slots=1, hbm_budget=1e6)


def test_mie_rigid_sphere_quadrilaterals():
    """The same Mie check on a quadrilateral cubed sphere (collocation at the
    element centres, Guiggiani hypersingular diagonal)."""
    from nihu_amd import serve
    from nihu_amd.mesh import cubed_sphere
    nodes, quads = cubed_sphere(12)
    srv = serve.Server({"Q": (nodes, quads)}, slots=1)
    k = 3.0
    d = np.array([0.0, 0.0, 1.0])
    res = srv.run([serve.Request(0, "Q", k, tuple(d))], realtime=False)[0]
    c, _ = serve._centres_normals(nodes, quads)
    ptot = np.exp(-1j * k * (c @ d)) + res.p
    mie = serve.mie_rigid_sphere_surface(k, 1.0, (c @ d) / np.linalg.norm(c, axis=1))
    err = np.linalg.norm(ptot - mie) / np.linalg.norm(mie)
    print("Mie surface error, 864 quadrilaterals", err, flush=True)
    assert err < 0.03


def test_iterative_refinement(meshes):
    """FP64 iterative refinement of the complex64 LU solve: the refined field
    approaches the complex128 solution of the same complex64 system; the
    server with refinement agrees with the plain solve to the latter's
    accuracy."""
    import torch
    from nihu_amd import serve, nbem
    nodes, tris = meshes["L3"]
    k, d = 7.0, np.array([0.0, 0.6, 0.8])
    m = nbem.Mesh(nodes, tris)
    n = len(tris)
    A, B = nbem.colmajor(n, n), nbem.colmajor(n, n)
    nbem.assemble_bm(m, k, A, B, pol=nbem.policy(**nbem.CONFIG_POLICY))
    c, nu = serve._centres_normals(nodes, tris)
    q = (1j * k) * (nu @ d) * np.exp(-1j * k * (c @ d))
    b = B @ torch.from_numpy(q.astype(np.complex64)).cuda()
    LU, piv, _ = torch.linalg.lu_factor_ex(A)
    p0 = torch.linalg.lu_solve(LU, piv, b.unsqueeze(1)).squeeze(1)
    p2 = serve._refine(A, LU, piv, b, p0, 2, block=256)
    exact = torch.linalg.solve(A.to(torch.complex128), b.to(torch.complex128).unsqueeze(1)).squeeze(1)
    torch.cuda.synchronize()
    e0 = (torch.linalg.norm(p0.to(torch.complex128) - exact) / torch.linalg.norm(exact)).item()
    e2 = (torch.linalg.norm(p2 - exact) / torch.linalg.norm(exact)).item()
    print("complex64 LU %.2e, refined twice %.2e" % (e0, e2), flush=True)
    assert e2 < 1e-3 * e0 or e2 < 1e-12
    plain = serve.Server({"L3": (nodes, tris)}, slots=1).run([serve.Request(0, "L3", k, tuple(d))], realtime=False)[0].p
    refd = serve.Server({"L3": (nodes, tris)}, slots=1, refine=2).run([serve.Request(0, "L3", k, tuple(d))],
                                                                        realtime=False)[0].p
    assert np.linalg.norm(refd - plain) / np.linalg.norm(refd) < 10 * e0 + 1e-6
