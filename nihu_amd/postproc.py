"""Exterior field post-processing and the rigid-sphere series solution.

field_pressure: the radiated / scattered pressure at field points from the
surface pressure p and its normal derivative q,
    p_f = M_f p - L_f q,
with L_f, M_f the Helmholtz single- and double-layer collocation matrices
between the field points (centres of the elements of a field mesh) and the
surface elements: NiHu's post-processing step (postproc in
nihu/applications/helmholtz/helmholtz_3d_coll.cpp:124-143 returns
Hf * psurf - Gf * qsurf; benchmark/helmholtz/plane_wave_scatter_from_sphere.m:55,
pf_num = Mf*ps_scat_num - Lf*qs_scat).  The field matrices are assembled by
the engine (different test and trial meshes: regular pairs, and the
near-singular pass where a field point lies close to the surface) one block
of field points at a time into device buffers, so memory stays bounded for
large field meshes, and multiplied on the device.

plane_wave_sphere_series: the incident and scattered fields of a unit plane
wave exp(-i k d.x) on a rigid sphere of radius r0 as the truncated series of
benchmark/helmholtz/plane_wave_scatter_from_sphere.m:57-99 and
dox/bench_plane_wave_sphere.md (time convention exp(+i w t); spherical
Bessel j_l and Hankel h_l^(2) = j_l - i y_l).
"""

from __future__ import annotations

import numpy as np

from . import nbem


def field_pressure(field: "nbem.Mesh", surface: "nbem.Mesh", k: float, p, q, pol=None, block_rows: int = 8192,
                   dtype: str = "complex128"):
    """p_f = M_f p - L_f q at the collocation points of `field` (element
    centres), for surface pressure p and normal derivative q at the surface
    elements (torch GPU vectors or numpy arrays).  Returns a torch complex
    vector on the device."""
    import torch
    pol = pol or nbem.policy(**nbem.NIHU_POLICY)
    dt = getattr(torch, dtype)
    p = torch.as_tensor(np.asarray(p) if not torch.is_tensor(p) else p).to(device="cuda", dtype=dt)
    q = torch.as_tensor(np.asarray(q) if not torch.is_tensor(q) else q).to(device="cuda", dtype=dt)
    if field is surface:
        raise ValueError("field points must come from a mesh other than the surface")
    m, n = field.n, surface.n
    if p.shape != (n,) or q.shape != (n,):
        raise ValueError("p and q must have one value per surface element")
    out = torch.empty(m, dtype=dt, device="cuda")
    # the field mesh's elements in blocks of rows: a sub-mesh per block
    nodes, elems = field.nodes, field.tris
    for r0 in range(0, m, block_rows):
        r1 = min(m, r0 + block_rows)
        sub = field if (r0 == 0 and r1 == m) else nbem.Mesh(nodes, elems[r0:r1])
        L = nbem.colmajor(r1 - r0, n, dtype)
        M = nbem.colmajor(r1 - r0, n, dtype)
        nbem.assemble(sub, surface, L, nbem.NB_HELMHOLTZ, 0, 0, k, pol=pol, same_mesh=False, accumulate=False)
        nbem.assemble(sub, surface, M, nbem.NB_HELMHOLTZ, 0, 1, k, pol=pol, same_mesh=False, accumulate=False)
        out[r0:r1] = M @ p - L @ q
        del L, M
    return out


def plane_wave_sphere_series(k: float, r0: float, points, direction=(1.0, 0.0, 0.0), order: int = 20):
    """(incident series, scattered field) at `points` (n x 3) for the unit
    plane wave exp(-i k d.x) on a rigid sphere of radius r0 centred at the
    origin, summed to `order` as the reference benchmark does:
    p_inc = sum (2l+1) i^l j_l(kr) P_l(cos th),
    p_scat = sum A_l h_l^(2)(kr) P_l(cos th),
    A_l = -(2l+1) i^l (l j_{l-1}(kr0) - (l+1) j_{l+1}(kr0)) / (l h_{l-1}(kr0) - (l+1) h_{l+1}(kr0)),
    with th the angle between x and -d (so that cos th = -d.x / r)."""
    from scipy.special import eval_legendre, spherical_jn, spherical_yn
    x = np.asarray(points, dtype=np.float64)
    d = np.asarray(direction, dtype=np.float64)
    d = d / np.linalg.norm(d)
    r = np.linalg.norm(x, axis=1)
    cth = -(x @ d) / r

    def jl(l, z):
        return spherical_jn(l, z) if l >= 0 else np.cos(z) / z      # j_{-1}(z) = cos z / z

    def hl(l, z):
        if l >= 0:
            return spherical_jn(l, z) - 1j * spherical_yn(l, z)
        return (np.cos(z) - 1j * np.sin(z)) / z                     # h2_{-1}(z) = j_{-1} - i y_{-1}, y_{-1} = sin z / z

    pinc = np.zeros(len(x), dtype=complex)
    pscat = np.zeros(len(x), dtype=complex)
    kr0 = k * r0
    for l in range(order + 1):
        P = eval_legendre(l, cth)
        pinc += (2 * l + 1) * (1j) ** l * spherical_jn(l, k * r) * P
        A = -(2 * l + 1) * (1j) ** l * (l * jl(l - 1, kr0) - (l + 1) * jl(l + 1, kr0)) / (
            l * hl(l - 1, kr0) - (l + 1) * hl(l + 1, kr0))
        pscat += A * hl(l, k * r) * P
    return pinc, pscat
