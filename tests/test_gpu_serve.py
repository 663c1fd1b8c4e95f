"""Serving loop: physics (rigid-sphere Mie series), same result alone and in a
mix of concurrent requests, admission by HBM footprint, and requests that fail
(invalid, expired, solver not converged) ending alone with their status."""

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
    budget = 1.5 * serve.footprint_bytes(n4, "gmres")
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
    plain = serve.Server({"L3": (nodes, tris)}, slots=1, solver="lu").run([serve.Request(0, "L3", k, tuple(d))],
                                                                          realtime=False)[0].p
    refd = serve.Server({"L3": (nodes, tris)}, slots=1, refine=2, solver="lu").run(
        [serve.Request(0, "L3", k, tuple(d))], realtime=False)[0].p
    assert np.linalg.norm(refd - plain) / np.linalg.norm(refd) < 10 * e0 + 1e-6


def _oracle_solve(orc, nodes, tris, req):
    """The plane-wave system of one request with the oracle's FP64 matrices
    (the configuration policy), solved in complex128 on the device."""
    import torch
    from nihu_amd import serve
    Ar, Br = orc.assemble_bm(nodes, tris, req.k, pol=orc.policy(rule=1, fixed_degree=5, near_mode=1, near_limit=2.0))
    c, nu = serve._centres_normals(nodes, tris)
    d = np.asarray(req.direction)
    q = (1j * req.k) * (nu @ d) * np.exp(-1j * req.k * (c @ d))
    Ag = torch.from_numpy(Ar).cuda()
    del Ar
    rhs = torch.from_numpy(Br).cuda() @ torch.from_numpy(q).cuda()
    del Br
    p = torch.linalg.solve(Ag, rhs.unsqueeze(1)).squeeze(1).cpu().numpy()
    del Ag, rhs
    torch.cuda.empty_cache()
    return p


def test_config5_stream_with_20480_triangles(orc):
    """The serving workload's meshes (1280, 5120 and 20480-triangle jittered
    icospheres, seed 0), nine requests with ka in [1, 40] arriving together,
    three per mesh, so that requests on one mesh are batched into one sweep
    launch, each solved by GMRES to a relative residual of 1e-8: the first
    20480-triangle request gives the same surface field alone and in the mix
    (1e-6), and one served solution per mesh matches the complex128 solve of
    the oracle's FP64 matrices (relative 2-norm 1e-5)."""
    from nihu_amd import serve
    from nihu_amd.mesh import jittered_icosphere
    meshes = {"L3": jittered_icosphere(3, seed=0), "L4": jittered_icosphere(4, seed=0),
              "L5": jittered_icosphere(5, seed=0)}
    srv = serve.Server(meshes, slots=4)
    base = serve.poisson_stream(9, rate=1e9, seed=21, k_range=(1.0, 40.0))
    reqs = [serve.Request(r.rid, ("L5", "L3", "L4")[i % 3], r.k, r.direction, r.arrival) for i, r in enumerate(base)]
    served = {r.rid: r for r in srv.run(reqs, realtime=False)}
    assert len(served) == len(reqs) and all(np.isfinite(r.p).all() for r in served.values())
    big = reqs[0]
    alone = srv.run([serve.Request(big.rid, big.mesh, big.k, big.direction)], realtime=False)[0]
    diff = np.abs(alone.p - served[big.rid].p).max() / np.abs(alone.p).max()
    print("20480 triangles, ka %.2f: alone vs mix %.2e; latencies %s" % (
        big.k, diff, {r.rid: round(r.latency, 3) for r in served.values()}), flush=True)
    assert diff <= 1e-6
    for key in ("L3", "L4", "L5"):
        req = next(r for r in reqs if r.mesh == key)
        p_ref = _oracle_solve(orc, *meshes[key], req)
        err = np.linalg.norm(served[req.rid].p - p_ref) / np.linalg.norm(p_ref)
        print("%s ka %.2f: served vs oracle complex128 solve %.2e" % (key, req.k, err), flush=True)
        assert err < 1e-5


def test_batched_equals_alone(meshes):
    """Requests on one mesh batched into one sweep launch give the same
    surface field as each request alone (1e-6), on the triangle meshes
    (one-launch sweep) and on a quadrilateral mesh (one assembly per
    wavenumber inside the sweep call)."""
    from nihu_amd import serve
    from nihu_amd.mesh import jittered_cubed_sphere
    ms = dict(meshes)
    ms["Q"] = jittered_cubed_sphere(6, seed=2)
    srv = serve.Server(ms, slots=2, max_batch=8)
    reqs = [serve.Request(i, key, k, (0.0, 0.6, 0.8)) for i, (key, k) in enumerate(
        [("L3", 2.0), ("L3", 9.5), ("L3", 17.25), ("L3", 33.0), ("Q", 3.0), ("Q", 8.0), ("L4", 4.0), ("L4", 21.0)])]
    batched = {r.rid: r for r in srv.run(reqs, realtime=False)}
    one = serve.Server(ms, slots=2, max_batch=1)
    for rq in reqs:
        alone = one.run([rq], realtime=False)[0]
        diff = np.abs(alone.p - batched[rq.rid].p).max() / np.abs(alone.p).max()
        print("request %d %s ka %.2f: batched vs alone %.2e" % (rq.rid, rq.mesh, rq.k, diff), flush=True)
        assert diff <= 1e-6


def test_gemv_and_gmres_against_lu(meshes):
    """nb_gemv_c64 (complex64 matrix, FP64 products and sums) against the
    complex128 product of the same matrix, bit-stable from call to call, and
    the server's GMRES solution against the complex128 LU solve of the same
    complex64 system."""
    import torch
    from nihu_amd import nbem, serve, solve
    nodes, tris = meshes["L4"]
    k, d = 12.0, np.array([0.6, 0.0, 0.8])
    m = nbem.Mesh(nodes, tris)
    n = len(tris)
    A, B = nbem.colmajor(n, n), nbem.colmajor(n, n)
    nbem.assemble_bm(m, k, A, B, pol=nbem.policy(**nbem.CONFIG_POLICY))
    x = torch.randn(n, dtype=torch.complex128, device="cuda")
    y = nbem.gemv(A, x)
    ref = A.to(torch.complex128) @ x
    e = (torch.linalg.norm(y - ref) / torch.linalg.norm(ref)).item()
    assert e < 1e-14, e
    assert torch.equal(nbem.gemv(A, x), y)
    # a sub-block with leading dimension > rows, and an odd number of columns
    As = A[5:1000, 3:1234]
    ys = nbem.gemv(As, x[3:1234])
    assert (torch.linalg.norm(ys - As.to(torch.complex128) @ x[3:1234]) / torch.linalg.norm(ys)).item() < 1e-14
    c, nu = serve._centres_normals(nodes, tris)
    q = torch.from_numpy((1j * k) * (nu @ d) * np.exp(-1j * k * (c @ d))).cuda()
    b = nbem.gemv(B, q)
    r = solve.gmres(A, b, tol=1e-10)
    exact = torch.linalg.solve(A.to(torch.complex128), b.unsqueeze(1)).squeeze(1)
    err = (torch.linalg.norm(r.x - exact) / torch.linalg.norm(exact)).item()
    print("GMRES: %d iterations, residual %.2e, error vs complex128 LU %.2e" % (r.iterations, r.residual, err),
          flush=True)
    assert r.converged and err < 1e-8


def test_gmres_batched_on_the_device(meshes):
    """Three Burton-Miller systems of the 1280-triangle mesh solved in
    lockstep equal each system solved alone (iterations, and solutions to
    1e-12: the batch orthogonalises with batched products)."""
    import torch
    from nihu_amd import nbem, serve, solve
    nodes, tris = meshes["L3"]
    m = nbem.Mesh(nodes, tris)
    n = len(tris)
    ks = np.array([3.0, 11.5, 27.0])
    A = torch.empty((3, n, n), dtype=torch.complex64, device="cuda").transpose(1, 2)
    B = torch.empty_like(A)
    nbem.assemble_sweep(m, ks, A, B, pol=nbem.policy(**nbem.CONFIG_POLICY))
    c, nu = serve._centres_normals(nodes, tris)
    d = np.array([0.0, 0.6, 0.8])
    rhs = torch.stack([nbem.gemv(B[g], torch.from_numpy((1j * k) * (nu @ d) * np.exp(-1j * k * (c @ d))).cuda())
                       for g, k in enumerate(ks)])
    out = solve.gmres_batched([A[g] for g in range(3)], rhs)
    for g in range(3):
        one = solve.gmres(A[g], rhs[g])
        diff = ((out[g].x - one.x).abs().max() / one.x.abs().max()).item()
        print("k %.1f: %d iterations (alone %d), difference %.1e" % (ks[g], out[g].iterations, one.iterations, diff),
              flush=True)
        assert out[g].converged and out[g].iterations == one.iterations and diff < 1e-12


def test_failures_on_the_device_path(meshes):
    """The real launch path with failing requests in the stream: invalid
    requests are rejected before any launch, a request that cannot be
    admitted before its deadline expires, a GMRES that cannot converge in
    its iteration limit (3 iterations) ends its request with an error; every
    other request is solved and equals its result alone."""
    from nihu_amd import serve
    srv = serve.Server({"L2": meshes["L2"], "L3": meshes["L3"]}, slots=1, max_batch=1)
    reqs = [serve.Request(0, "L3", 5.0, (1.0, 0.0, 0.0)), serve.Request(1, "L2", -3.0, (1.0, 0.0, 0.0)),
            serve.Request(2, "L2", 4.0, (0.0, 2.0, 0.0)), serve.Request(3, "L3", 6.5, (0.0, 0.0, 1.0)),
            serve.Request(4, "L2", 2.5, (0.0, 1.0, 0.0), deadline=0.0),
            serve.Request(5, "L2", 3.5, (0.0, 1.0, 0.0))]
    res = {r.rid: r for r in srv.run(reqs, realtime=False)}
    assert sorted(res) == list(range(6))
    assert res[1].status == "rejected" and res[2].status == "rejected"
    assert res[4].status == "expired" and res[4].p is None
    for rid in (0, 3, 5):
        assert res[rid].status == "ok", res[rid].error
        rq = reqs[rid]
        alone = srv.run([serve.Request(rq.rid, rq.mesh, rq.k, rq.direction)], realtime=False)[0]
        assert np.abs(alone.p - res[rid].p).max() <= 1e-6 * np.abs(alone.p).max()
    tight = serve.Server({"L3": meshes["L3"]}, slots=1, maxiter=3)
    bad = tight.run([serve.Request(9, "L3", 30.0, (1.0, 0.0, 0.0))], realtime=False)[0]
    print("3-iteration GMRES:", bad.status, bad.error, flush=True)
    assert bad.status == "error" and "GMRES" in bad.error and bad.p is None
