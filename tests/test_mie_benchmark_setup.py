"""The setup of NiHu's published rigid-sphere benchmark
(benchmark/helmholtz/plane_wave_scatter_from_sphere.m, results in
dox/bench_plane_wave_sphere.md:64-66), rebuilt on the CPU: the blown-up cube
sphere (r0 = 0.4, 12 elements along the radius), k = min(mesh_kmax(., 9)),
the 100 x 100 field slab without the elements that touch R <= 0.404, and the
truncated series.  NiHu publishes the series truncation error of its setup,
1.29315e-07: it depends only on the collocation points and k, so matching it
to all six printed digits shows that the mesh and k are the reference's."""

import numpy as np

from nihu_amd import mesh as M
from nihu_amd import postproc


def benchmark_setup():
    nodes, quads = M.sphere_boundary(0.4, 12)
    k = M.mesh_kmax(nodes, quads, 9).min()
    fn, fe = M.create_slab([[-1, 0, -1], [1, 0, -1], [1, 0, 1], [-1, 0, 1]], 100, 100)
    fn, fe = M.select_elements(fn, fe, lambda x: np.linalg.norm(x, axis=1) > 0.404)
    return nodes, quads, k, fn, fe


def test_benchmark_mesh_and_series_truncation():
    nodes, quads, k, fn, fe = benchmark_setup()
    assert quads.shape == (3456, 4)
    c = nodes[quads].mean(axis=1)
    pinc_ser, _ = postproc.plane_wave_sphere_series(k, 0.4, c)
    pinc = np.exp(-1j * k * c[:, 0])
    err_ser = np.linalg.norm(pinc - pinc_ser) / np.linalg.norm(pinc)
    print("k = %.6f, %d field elements, series truncation error %.6g" % (k, len(fe), err_ser))
    assert abs(err_ser - 1.29315e-07) < 5e-13
    # the slab: 100 x 100 elements of side 0.02 in the y = 0 plane, those with
    # every node outside R = 0.404 kept
    assert fe.shape[1] == 4 and 8000 < len(fe) < 10000
    assert np.all(np.linalg.norm(fn[fe], axis=2) > 0.404)
    assert np.allclose(fn[:, 1], 0.0)


def test_series_solves_the_rigid_sphere_problem():
    """The scattered series cancels the incident normal derivative on the
    sphere (rigid boundary, q = 0), checked by central differences in r, and
    its far field decays as 1/r."""
    k, r0 = 5.0, 0.4
    rng = np.random.default_rng(3)
    v = rng.normal(size=(50, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    h = 1e-6
    pi_p, ps_p = postproc.plane_wave_sphere_series(k, r0, v * (r0 + h), order=30)
    pi_m, ps_m = postproc.plane_wave_sphere_series(k, r0, v * (r0 - h), order=30)
    dtot = (pi_p + ps_p - pi_m - ps_m) / (2 * h)
    dinc = (pi_p - pi_m) / (2 * h)
    assert np.abs(dtot).max() < 1e-6 * np.abs(dinc).max()
    _, far1 = postproc.plane_wave_sphere_series(k, r0, v * 50.0, order=30)
    _, far2 = postproc.plane_wave_sphere_series(k, r0, v * 100.0, order=30)
    assert np.allclose(np.abs(far1), 2.0 * np.abs(far2), rtol=2e-2)
