"""NiHu's published accuracy benchmark, plane-wave scattering by a rigid
sphere (benchmark/helmholtz/plane_wave_scatter_from_sphere.m), rebuilt on the
engine: the conventional (CBIE) collocation system of its MEX application
(nihu/applications/helmholtz/helmholtz_3d_coll_mex.mex.cpp:35-58:
Ls = SLP, Ms = DLP - I/2 on the surface; Lf, Mf from the field points),
NiHu's default quadrature policy, the surface solve Ms \\ (Ls q) (complex128
LU) and the exterior field pf = Mf ps - Lf qs.

NiHu publishes (dox/bench_plane_wave_sphere.md:64-66): series truncation
error 1.29315e-07, surface relative error 0.0140948, field relative error
0.013953.  The setup here reproduces the first to six digits (the mesh and k
are the reference's); the two discretisation errors come from the engine's
matrices and must match NiHu's within 10 %."""

import numpy as np
import pytest

from tests.test_mie_benchmark_setup import benchmark_setup

pytestmark = pytest.mark.gpu

PUBLISHED = {"series": 1.29315e-07, "surface": 0.0140948, "field": 0.013953}


def test_plane_wave_sphere_benchmark():
    import torch
    from nihu_amd import nbem, postproc
    assert torch.cuda.is_available()
    nodes, quads, k, fn, fe = benchmark_setup()
    n = len(quads)
    surf = nbem.Mesh(nodes, quads)
    field = nbem.Mesh(fn, fe)
    pol = nbem.policy(**nbem.NIHU_POLICY)
    Ls, Ms = nbem.colmajor(n, n, "complex128"), nbem.colmajor(n, n, "complex128")
    nbem.assemble(surf, surf, Ls, nbem.NB_HELMHOLTZ, 0, 0, k, pol=pol, accumulate=False)
    nbem.assemble(surf, surf, Ms, nbem.NB_HELMHOLTZ, 0, 1, k, pol=pol, accumulate=False)
    nbem.assemble(surf, surf, Ms, nbem.NB_IDENTITY, 0, 0, 0.0, alpha=-0.5, pol=pol, accumulate=True)
    # incident plane wave along +x at the collocation points (element centres)
    X = nodes[quads]
    c = X.mean(axis=1)
    nv = np.cross(X[:, 1] + X[:, 2] - X[:, 0] - X[:, 3], X[:, 2] + X[:, 3] - X[:, 0] - X[:, 1])
    nu = nv / np.linalg.norm(nv, axis=1)[:, None]
    d = np.array([1.0, 0.0, 0.0])
    ps_inc = np.exp(-1j * k * (c @ d))
    qs_inc = -1j * k * (nu @ d) * ps_inc
    qs = torch.from_numpy(-qs_inc).cuda()
    ps = torch.linalg.solve(Ms, (Ls @ qs).unsqueeze(1)).squeeze(1)
    pf = postproc.field_pressure(field, surf, k, ps, qs, pol=pol, block_rows=4096)
    torch.cuda.synchronize()
    ps, pf = ps.cpu().numpy(), pf.cpu().numpy()
    assert np.isfinite(ps).all() and np.isfinite(pf).all()
    fc = fn[fe].mean(axis=1)
    _, ps_ana = postproc.plane_wave_sphere_series(k, 0.4, c, d)
    _, pf_ana = postproc.plane_wave_sphere_series(k, 0.4, fc, d)
    err_surf = np.linalg.norm(ps_ana - ps) / np.linalg.norm(ps_ana)
    err_field = np.linalg.norm(pf_ana - pf) / np.linalg.norm(pf_ana)
    print("k r0 = %.4f, %d surface and %d field elements: surface relative error %.6g (NiHu 0.0140948), "
          "field relative error %.6g (NiHu 0.013953)" % (k * 0.4, n, len(fe), err_surf, err_field), flush=True)
    assert abs(err_surf / PUBLISHED["surface"] - 1.0) < 0.10
    assert abs(err_field / PUBLISHED["field"] - 1.0) < 0.10
