"""Headline benchmark: dense Burton-Miller collocation assembly on one MI355X.

Workload (one step): the Helmholtz 3D collocation matrices A = alpha*D - I/2 + H
and B = G + alpha*Ht + alpha/2*I (alpha = -i/k) on the 20480-triangle unit
icosphere (level 5, +-1% seeded radial jitter, seed 0), ka = 10: far pairs with
the 7-point Dunavant rule in FP32 (complex64 results), pairs whose centroid
distance is under 2 element diameters and the diagonal in FP64.  Inputs stay
resident on the device; a step is one nb_assemble_bm call (regular, near and
diagonal passes) writing both 20480^2 complex64 matrices.

Metric: matrix entries/s = (number of kernel operators: SLP, DLP, DLPt, HSP =
4) x 20480^2 per step / wall seconds.

Also reported: the dominant kernel's roofline from its HIP-event time; the
relative Frobenius and max errors of both matrices against the FP64 oracle
(NiHu's algorithm restated), which assembles the same matrices on the host
with OpenMP over rows and is timed as the CPU baseline (beside it the serial
pair loop on a row sample, and the host's CPU model); the solved surface field
against the oracle's, and the total field against the Mie series.  The same
line carries two more sections after the timed headline steps: `sweep`, the
64-wavenumber sweep on 5120 triangles in one launch against 64
single-wavenumber assemblies (ms per sweep, p50/p99 sweep latency, errors
against the oracle), and `serve`, a Poisson stream of assembly-plus-solve
requests served with and without batching (p50/p99 latency, solves/s) and a
rate sweep up to the first rate the server does not keep up with; then short
runs (5 timed steps, 32 rows or node groups against the oracle) of the Galerkin workload
(config 3), the Galerkin workload on quadrilaterals, the quadrilateral, the
gauss-field and the mixed-element workloads, each with its roofline:
`galerkin`, `galerkin_quad`, `quad`, `gauss`, `mixed`, and config 1, the
Laplace matrices of 1280 triangles in FP64, `laplace_fp64` (--no-extras skips
all of them).

    python bench.py [--gpus 1] [--steps 10] [--warmup 20]
"""

from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FLOPS_PER_POINT_BM = 65        # SLP+DLP+DLPt+HSP fused, per (collocation point, quadrature point)
PEAK_FP32_TFLOPS = 157.3       # MI355X published FP32 peak (packed); non-packed VALU issue is half
PEAK_HBM_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def read_traffic(path):
    """HBM bytes per launch of the dominant kernel from a committed rocprofv3
    --pmc summary (line 'traffic_bytes_per_launch: <float>'), else None."""
    try:
        txt = open(path).read()
    except OSError:
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", txt)
    return float(m.group(1)) if m else None


def surface_field_error(nodes, tris, k, A, B, Ar, Br):
    """Plane-wave scattering from the rigid surface, A p = B q with
    q = i k (d.n) e^{-ik d.x} at the element centres (seeded direction d),
    solved in complex128 on the GPU once with the engine's matrices and once
    with the oracle's; returns the relative 2-norm and max error of p."""
    import numpy as np
    import torch
    X = nodes[tris]
    c = X.mean(1)
    nv = np.cross(X[:, 1] - X[:, 0], X[:, 2] - X[:, 0])
    nu = nv / np.linalg.norm(nv, axis=1)[:, None]
    d = np.random.default_rng(7).standard_normal(3)
    d /= np.linalg.norm(d)
    q = torch.from_numpy((1j * k) * (nu @ d) * np.exp(-1j * k * (c @ d))).cuda()

    def solve(Am, Bm):
        Am = Am.to(torch.complex128)
        rhs = Bm.to(torch.complex128) @ q
        return torch.linalg.solve(Am, rhs.unsqueeze(1)).squeeze(1)

    pg = solve(A, B).cpu().numpy()
    torch.cuda.empty_cache()
    pr = solve(torch.from_numpy(Ar).cuda(), torch.from_numpy(Br).cuda()).cpu().numpy()
    torch.cuda.empty_cache()
    # discretisation sanity check: total field p_inc + p against the Mie
    # series of the rigid unit sphere (the mesh is the sphere to +-1%)
    from nihu_amd.serve import mie_rigid_sphere_surface
    pinc = np.exp(-1j * k * (c @ d))
    mie = mie_rigid_sphere_surface(k, 1.0, (c @ d) / np.linalg.norm(c, axis=1))
    return {"surface_field_rel": float(np.linalg.norm(pg - pr) / np.linalg.norm(pr)),
            "surface_field_max_abs": float(np.abs(pg - pr).max()),
            "surface_field_solve": "A p = B q, plane wave, complex128 LU on the GPU for both matrix sets",
            "mie_rel": float(np.linalg.norm(pinc + pg - mie) / np.linalg.norm(mie)),
            "mie_rel_oracle": float(np.linalg.norm(pinc + pr - mie) / np.linalg.norm(mie))}


ORACLE_BUILDS = {
    "tuned": "g++ -O3 -march=x86-64-v4 -ffp-contract=fast -fopenmp (AVX-512, FMA contraction)",
    "exact": "g++ -O3 -march=x86-64-v3 -ffp-contract=off -fopenmp (the checker build)",
}


def oracle_tuned_build(oracle):
    """Time the CPU baseline with the tuned oracle build where the host can
    run it (x86-64-v4); returns the build's name."""
    if oracle.tuned_supported():
        oracle.use_build("tuned")
        return "tuned"
    oracle.use_build("exact")
    return "exact"


def host_cpu():
    """CPU model (from /proc/cpuinfo) and the number of cores this process may
    run on, stated beside every CPU number."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count()
    return {"model": model, "cores_available": avail, "logical_cpus": os.cpu_count()}


def copy_calibration(nbytes):
    """One plain device-to-device copy of nbytes: the kernel whose HBM counters
    calibrate FETCH_SIZE / WRITE_SIZE (it reads and writes nbytes)."""
    import torch
    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    log("copy kernel: %d bytes read, %d bytes written" % (nbytes, nbytes))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=20,
                    help="untimed steps first; kernel durations settle over the first few launches of a process")
    ap.add_argument("--level", type=int, default=5)
    ap.add_argument("--ka", type=float, default=10.0)
    ap.add_argument("--policy", choices=["config", "nihu"], default="config")
    ap.add_argument("--cpu-rows", type=int, default=-1,
                    help="CPU baseline: -1 the whole matrices (default; also gives full-matrix and solved-field "
                         "errors), N > 0 a sample of N rows, 0 skip")
    ap.add_argument("--traffic-file", default=None,
                    help="rocprofv3 FETCH/WRITE summary of the regular kernel (default: the one of --policy)")
    ap.add_argument("--profile-copy-bytes", type=int, default=0,
                    help="also run one plain device copy of this many bytes (HBM-counter calibration)")
    ap.add_argument("--workload", choices=["bm", "sweep", "serve", "galerkin", "galerkin_quad", "quad", "gauss", "mixed",
                                           "laplace"],
                    default="bm",
                    help="bm: the headline single-wavenumber workload; sweep: 64 wavenumbers on 5120 triangles; "
                         "serve: Poisson stream of assembly-plus-LU-solve requests; galerkin: SLP and DLP "
                         "Galerkin matrices with linear triangles on the 20480-element ellipsoid; galerkin_quad: "
                         "the same with bilinear quadrilaterals on a 9600-element cube sphere; quad: "
                         "Burton-Miller collocation on a 20184-element quadrilateral cubed sphere; gauss: the "
                         "same with the gauss field (4 dofs per element) on 5400 quadrilaterals; mixed: the "
                         "same on a 20000-element surface of triangles and quadrilaterals (NiHu element table); "
                         "laplace: the Laplace SLP and DLP matrices of 1280 triangles in FP64 (plain method)")
    ap.add_argument("--requests", type=int, default=24, help="serve: requests per stream")
    ap.add_argument("--no-extras", action="store_true",
                    help="headline only: skip the sweep, serving, Galerkin, quadrilateral and mixed-mesh "
                         "sections of the default line")
    ap.add_argument("--rate", type=float, default=4.0, help="serve: mean arrivals per second")
    ap.add_argument("--refine", type=int, default=0,
                    help="serve: steps of FP64 iterative refinement of each complex64 LU solve")
    args = ap.parse_args()
    if args.traffic_file is None:
        args.traffic_file = os.path.join(ROOT, "profiles", "regular_bm_L5_traffic.txt" if args.policy == "config"
                                         else "regular_bm_L5_nihu_traffic.txt")

    if args.gpus != 1:
        print("bench.py: this engine runs on one GPU; the multi-GPU path (row blocks per GPU) is design only. "
              "Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    if args.workload == "sweep":
        return run_sweep(args)
    if args.workload == "serve":
        return run_serve(args)
    if args.workload == "galerkin":
        print(json.dumps(run_galerkin(args)))
        return
    if args.workload == "galerkin_quad":
        print(json.dumps(run_galerkin_quad(args)))
        return
    if args.workload == "quad":
        print(json.dumps(run_quad(args)))
        return
    if args.workload == "gauss":
        print(json.dumps(run_gauss(args)))
        return
    if args.workload == "mixed":
        print(json.dumps(run_mixed(args)))
        return
    if args.workload == "laplace":
        print(json.dumps(run_laplace(args)))
        return

    import numpy as np
    import torch
    from nihu_amd import nbem
    from nihu_amd.mesh import jittered_icosphere

    torch.cuda.set_device(0)
    nodes, tris = jittered_icosphere(args.level, seed=0)
    n = len(tris)
    k = args.ka  # unit radius: ka = k
    alpha = complex(0.0, -1.0 / k)
    pol_kw = nbem.CONFIG_POLICY if args.policy == "config" else nbem.NIHU_POLICY
    pol = nbem.policy(**pol_kw)
    name, ncu, hbm = nbem.device_info()
    log("device %s, %d CUs, %.0f GiB; mesh %d triangles; policy %s" % (name, ncu, hbm / 2 ** 30, n, pol_kw))

    mesh = nbem.Mesh(nodes, tris)
    A, B = nbem.colmajor(n, n), nbem.colmajor(n, n)
    stream = torch.cuda.current_stream()

    # pair census (for the algorithmic work of the regular pass)
    cls = nbem.classify(mesh, mesh, 0, 0, pol)
    n_near = int((cls == -2).sum().item())
    n_diag = n
    n_reg = n * n - n_near - n_diag
    if args.policy == "config":
        pts = 7 * n_reg
    else:
        pts = None
    del cls

    if args.profile_copy_bytes > 0:
        copy_calibration(args.profile_copy_bytes)

    for w in range(args.warmup):
        nbem.assemble_bm(mesh, k, A, B, alpha=alpha, pol=pol, stream=stream)
        torch.cuda.synchronize()
        log("warmup step %d done" % w)

    nbem.profile(True)
    nbem.profile_read(reset=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        nbem.assemble_bm(mesh, k, A, B, alpha=alpha, pol=pol, stream=stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    prof = nbem.profile_read(reset=True)
    nbem.profile(False)
    wall = t1 - t0
    ms_per_step = wall / args.steps * 1e3
    entries = 4 * n * n
    value = entries * args.steps / wall
    log("timed %d steps: %.3f ms/step, %.4g entries/s" % (args.steps, ms_per_step, value))
    pass_ms = {p: (v[0] / v[1] if v[1] else 0.0) for p, v in prof.items()}
    log("per-pass device ms: %s" % pass_ms)

    reg_ms = pass_ms["regular"]
    write_bytes = 2 * 8 * n_reg
    roofline = None
    if reg_ms > 0:
        if pts is not None:
            flops = pts * FLOPS_PER_POINT_BM
            ach = flops / (reg_ms * 1e-3) / 1e12
            roofline = {"bound": "valu", "achieved": round(ach, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                        "frac": round(ach / PEAK_FP32_TFLOPS, 4), "traffic": read_traffic(args.traffic_file),
                        "engine": "VALU fp32 (no MFMA; packed fp32 disabled at build, so half the peak is the "
                                  "issue limit)",
                        "kernel": "regular_kernel<HELM,ALL4,OUT_BM>", "kernel_ms": round(reg_ms, 4),
                        "algorithmic_flops": flops,
                        "hbm_write_GBs": round(write_bytes / (reg_ms * 1e-3) / 1e9, 1)}
        else:
            gbs = write_bytes / (reg_ms * 1e-3) / 1e9
            roofline = {"bound": "hbm", "achieved": round(gbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                        "frac": round(gbs / PEAK_HBM_GBS, 4), "traffic": read_traffic(args.traffic_file),
                        "kernel": "regular_kernel<HELM,ALL4,OUT_BM>", "kernel_ms": round(reg_ms, 4)}

    # CPU baseline (the FP64 oracle) and accuracy against it
    cpu = None
    err = None
    if args.cpu_rows != 0:
        from oracle import oracle
        full = args.cpu_rows < 0 or args.cpu_rows >= n
        rows = None if full else np.sort(np.random.default_rng(123).choice(n, args.cpu_rows, replace=False)).astype(
            np.int32)
        nr = n if full else len(rows)
        opol = oracle.policy(rule=pol.rule, fixed_degree=pol.fixed_degree, near_mode=pol.near_mode,
                             near_limit=pol.near_limit)
        threads = oracle.num_threads()
        build = oracle_tuned_build(oracle)
        log("cpu baseline: oracle (%s build) on %d rows, %d threads" % (build, nr, threads))
        tc = time.perf_counter()
        Ar, Br = oracle.assemble_bm(nodes, tris, k, alpha=alpha, pol=opol, rows=rows, threads=threads)
        dt = time.perf_counter() - tc
        cpu_value = 4 * nr * n / dt
        cpu = {"value": cpu_value, "unit": "matrix entries/s", "cores": threads, "kind": "port",
               "sample": ("the whole workload (A and B, all %d rows; 4 operators)" % n if full else
                          "%d of %d rows of A and B (4 operators, all %d columns)" % (nr, n, n)) +
                         ", FP64 oracle restating NiHu's pair algorithm, OpenMP over rows, %.1f s wall" % dt,
               "build": ORACLE_BUILDS[build]}
        log("cpu baseline: %.4g entries/s on %d threads (%.1fx faster on GPU)" % (cpu_value, threads,
                                                                                 value / cpu_value))
        # NiHu's own single-matrix path is the serial pair loop: the same
        # oracle on one thread, on a bounded row sample
        srows = np.sort(np.random.default_rng(321).choice(n, 24, replace=False)).astype(np.int32)
        tc = time.perf_counter()
        oracle.assemble_bm(nodes, tris, k, alpha=alpha, pol=opol, rows=srows, threads=1)
        dts = time.perf_counter() - tc
        cpu["serial"] = {"value": 4 * len(srows) * n / dts, "cores": 1,
                         "sample": "%d of %d rows of A and B, one thread, %.1f s wall" % (len(srows), n, dts)}
        cpu["host"] = host_cpu()
        oracle.use_build("exact")
        log("cpu serial: %.4g entries/s; host %s" % (cpu["serial"]["value"], cpu["host"]))
        if full:
            Ag, Bg = A.cpu().numpy(), B.cpu().numpy()
        else:
            ri = torch.from_numpy(rows.astype(np.int64)).cuda()
            Ag, Bg = A[ri, :].cpu().numpy(), B[ri, :].cpu().numpy()
        ea = float(np.linalg.norm(Ag - Ar) / np.linalg.norm(Ar))
        eb = float(np.linalg.norm(Bg - Br) / np.linalg.norm(Br))
        err = {"rows": nr, "A_rel_fro": ea, "B_rel_fro": eb,
               "A_max_abs": float(np.abs(Ag - Ar).max()), "B_max_abs": float(np.abs(Bg - Br).max())}
        del Ag, Bg
        if full:
            err.update(surface_field_error(nodes, tris, k, A, B, Ar, Br))
        del Ar, Br
        log("error vs oracle: %s" % err)

    # the second half of the metric: the served frequency sweep (config 4 in
    # one launch against the per-frequency path; config 5 served stream, its
    # p50/p99 and its saturation rate)
    extras = {}
    if not args.no_extras and args.level == 5:
        del A, B
        torch.cuda.empty_cache()
        extras["sweep"] = sweep_summary(steps=5, warmup=2)
        torch.cuda.empty_cache()
        extras["serve"] = serve_summary(n_req=args.requests, rate=args.rate, refine=args.refine)
        torch.cuda.empty_cache()
        # the other configurations at their full size, short: config 3
        # (Galerkin: the test elements of 32 sampled nodes against the
        # oracle), the quadrilateral and the mixed-element workloads (32 rows),
        # each with its roofline
        for key, fn, rows in (("galerkin", run_galerkin, -1), ("galerkin_quad", run_galerkin_quad, -1),
                              ("quad", run_quad, 32), ("gauss", run_gauss, 1024),
                              ("mixed", run_mixed, 32), ("laplace_fp64", run_laplace, -1)):
            rec = fn(argparse.Namespace(steps=5, warmup=5, cpu_rows=rows, profile_copy_bytes=0, level=5))
            extras[key] = {k: rec.get(k) for k in ("metric", "value", "unit", "ms_per_step", "config", "roofline",
                                                   "cpu_baseline", "accuracy", "pass_ms")}
            torch.cuda.empty_cache()

    out = {
        "metric": "matrix entries/s, Helmholtz 3D coll. 20480 tris; p50/p99 sweep latency",
        "value": value,
        "unit": "matrix entries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 (complex64) far pairs, fp64 near/diagonal",
        "data": "synthetic: seeded jittered icosphere",
        "config": {"workload": "Helmholtz 3D collocation, Burton-Miller A and B (SLP, DLP, DLPt, HSP)",
                   "n_tri": n, "icosphere_level": args.level, "ka": k, "mesh_seed": 0,
                   "far_rule": "Dunavant degree 5 (7 points)" if args.policy == "config" else "NiHu degree tables",
                   "near": "centroid distance < 2 element diameters, FP64" if args.policy == "config"
                   else "d/R < 1.5, FP64",
                   "pairs": {"regular": n_reg, "near": n_near, "diagonal": n_diag}},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "accuracy": err,
        "pass_ms": pass_ms,
    }
    out.update(extras)
    print(json.dumps(out))


# FP32 operations per (test point, trial point) of the Galerkin regular pass
# with linear fields (kernel value, phase, three shape-weighted sums, and the
# per-test-point block update spread over the 6 trial points)
FLOPS_PER_POINTPAIR_GAL = {"SLP": 40, "DLP": 51}


def run_galerkin(args):
    """Galerkin workload: Helmholtz single- and double-layer matrices with
    isoparametric linear triangles on the seeded ellipsoid (level 5: 20480
    elements, 10242 nodes, semi-axes U[0.5, 1.5], +-1% jitter), k = 5 (ka = 5
    for the unit parent sphere), 6 x 6-point regular rule (Dunavant degree 4
    on both elements) in FP32, coincident / edge / vertex pairs with the
    Duffy-type rules in FP64, complex64 results.  A step assembles both
    10242^2 matrices."""
    import numpy as np
    import torch
    from nihu_amd import nbem
    from nihu_amd.mesh import ellipsoid

    torch.cuda.set_device(0)
    nodes, tris = ellipsoid(args.level, seed=0)
    ne, nn = len(tris), len(nodes)
    k = 5.0
    pol = nbem.policy(rule="fixed", degree=4, near="none")
    name, ncu, hbm = nbem.device_info()
    log("device %s, %d CUs; ellipsoid %d elements, %d nodes; k = %g" % (name, ncu, ne, nn, k))
    mesh = nbem.Mesh(nodes, tris)
    L, M = nbem.colmajor(nn, nn), nbem.colmajor(nn, nn)
    stream = torch.cuda.current_stream()
    if args.profile_copy_bytes > 0:
        copy_calibration(args.profile_copy_bytes)

    def step():
        nbem.assemble(mesh, mesh, L, 1, 0, 0, k, pol=pol, accumulate=False, stream=stream,
                      formalism="galerkin_linear")
        nbem.assemble(mesh, mesh, M, 1, 0, 1, k, pol=pol, accumulate=False, stream=stream,
                      formalism="galerkin_linear")

    for w in range(args.warmup):
        step()
        torch.cuda.synchronize()
        log("warmup step %d done" % w)
    nbem.profile(True)
    nbem.profile_read(reset=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    prof = nbem.profile_read(reset=True)
    nbem.profile(False)
    ms_per_step = wall / args.steps * 1e3
    value = 2 * nn * nn * args.steps / wall
    pass_ms = {p: (v[0] / v[1] if v[1] else 0.0) for p, v in prof.items()}
    log("timed %d steps: %.3f ms/step, %.4g entries/s; per-launch device ms %s" % (args.steps, ms_per_step, value,
                                                                                 pass_ms))

    # node-sharing pairs go to the FP64 pass; the rest are regular
    node_elems = [[] for _ in range(nn)]
    for e, t in enumerate(tris):
        for v in t:
            node_elems[v].append(e)
    n_sing = 0
    for e, t in enumerate(tris):
        n_sing += len(set(node_elems[t[0]]) | set(node_elems[t[1]]) | set(node_elems[t[2]]))
    n_reg = ne * ne - n_sing
    reg_ms_pair = 2 * pass_ms["regular"]       # SLP and DLP launches per step
    flops = n_reg * 36 * (FLOPS_PER_POINTPAIR_GAL["SLP"] + FLOPS_PER_POINTPAIR_GAL["DLP"])
    ach = flops / (reg_ms_pair * 1e-3) / 1e12 if reg_ms_pair > 0 else 0.0
    roofline = {"bound": "valu", "achieved": round(ach, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                "frac": round(ach / PEAK_FP32_TFLOPS, 4), "traffic": None,
                "engine": "VALU fp32 (no MFMA; packed fp32 disabled at build)",
                "kernel": "galerkin_regular_kernel<HELM,{SLP,DLP},3,6>", "kernel_ms": round(reg_ms_pair, 4),
                "algorithmic_flops": flops}

    cpu = err = None
    if args.cpu_rows != 0:
        from oracle import oracle
        rng = np.random.default_rng(123)
        n_sample = 32 if args.cpu_rows < 0 else max(1, args.cpu_rows // 16)
        sample = np.sort(rng.choice(nn, n_sample, replace=False))
        elems = np.unique(np.nonzero(np.isin(tris, sample))[0]).astype(np.int32)
        opol = oracle.policy(rule=1, fixed_degree=4, near_mode=2, near_limit=1.5)
        threads = oracle.num_threads()
        tc = time.perf_counter()
        Lr = oracle.assemble_galerkin(nodes, tris, 1, 0, 0, k=k, pol=opol, linear=True, test_elems=elems,
                                      threads=threads)[sample]
        Mr = oracle.assemble_galerkin(nodes, tris, 1, 0, 1, k=k, pol=opol, linear=True, test_elems=elems,
                                      threads=threads)[sample]
        dt = time.perf_counter() - tc
        # the oracle integrated len(elems) of ne test elements against every trial element
        cpu_value = 2 * nn * nn / (dt * ne / len(elems))
        cpu = {"value": cpu_value, "unit": "matrix entries/s", "cores": threads, "kind": "port",
               "sample": "test elements around %d sampled nodes (%d of %d elements) against all %d trial "
                         "elements, SLP and DLP, FP64 oracle, %.1f s wall; rate scaled by %d/%d to the whole "
                         "matrices" % (len(sample), len(elems), ne, ne, dt, len(elems), ne),
               "host": host_cpu()}
        si = torch.from_numpy(sample.astype(np.int64)).cuda()
        Lg = L[si, :].cpu().numpy().astype(np.complex128)
        Mg = M[si, :].cpu().numpy().astype(np.complex128)
        err = {"rows": int(len(sample)),
               "L_rel_fro": float(np.linalg.norm(Lg - Lr) / np.linalg.norm(Lr)),
               "M_rel_fro": float(np.linalg.norm(Mg - Mr) / np.linalg.norm(Mr))}
        log("sampled-row error vs oracle: %s; cpu %.4g entries/s on %d threads" % (err, cpu_value, threads))

    out = {
        "metric": "matrix entries/s, Helmholtz 3D Galerkin linear tris, 20480 elems",
        "value": value, "unit": "matrix entries/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 (complex64) regular pairs, fp64 singular pairs",
        "data": "synthetic: seeded ellipsoid",
        "config": {"workload": "Helmholtz 3D Galerkin, isoparametric linear triangles, SLP and DLP matrices",
                   "n_elements": ne, "n_nodes": nn, "k": k, "mesh_seed": 0,
                   "regular_rule": "Dunavant degree 4 x degree 4 (6 x 6 points)",
                   "singular": "coincident/edge/vertex Duffy-type rules, 4-point Gauss per axis, FP64",
                   "pairs": {"regular": n_reg, "singular": n_sing}},
        "roofline": roofline, "cpu_baseline": cpu, "accuracy": err, "pass_ms": pass_ms,
    }
    return out


# FP32 operations per (test point, trial point) of the quadrilateral Galerkin
# regular pass with bilinear nodal fields, as the kernel writes them: r^2
# (3 FMA, 2 adds), rsq, the phase k (r - d) = k u / (r + d) (FMA, rcp, mul,
# FMA), sin and cos, the kernel (SLP 2 muls; DLP the 1/r^3 and k/r^2 terms,
# r.n_y with the point's own normal, 3 FMA, and 2 muls) and the four
# shape-weighted complex sums (8 FMA); the per-test-point block update is not
# counted
FLOPS_PER_POINTPAIR_GALQ = {"SLP": 35, "DLP": 50}


def galq_point_pairs(nodes, quads, complexity=2):
    """Point pairs per kernel of the regular quadrilateral Galerkin pairs under
    NiHu's degrees: N^4 per element pair, N = (complexity + table degree) / 2
    + 1 from d / R_trial (1/r table for SLP, 1/r^2 for DLP); node-sharing
    pairs are the FP64 pass's."""
    import numpy as np
    X = nodes[quads]
    c = X.mean(1)
    R = np.mean([np.linalg.norm(X[:, (a + 1) % 4] - X[:, a], axis=1) for a in range(4)], axis=0)
    tabs = {"SLP": [(1.1, 6), (1.8, 4), (9.1, 2), (np.inf, 0)],
            "DLP": [(1.1, 8), (1.5, 6), (2.8, 4), (15.8, 2), (np.inf, 0)]}
    out = {"SLP": 0, "DLP": 0}
    n = len(quads)
    for i0 in range(0, n, 512):
        d = np.linalg.norm(c[i0:i0 + 512, None, :] - c[None, :, :], axis=2) / R[None, :]
        for key, tab in tabs.items():
            deg = np.full(d.shape, complexity + tab[-1][1])
            for lim, dg in reversed(tab[:-1]):
                deg[d < lim] = complexity + dg
            N = deg // 2 + 1
            out[key] += int((N.astype(np.int64) ** 4).sum())
    # node-sharing pairs (subtract at the smallest distances: the finest rule)
    node_elems = [[] for _ in range(len(nodes))]
    for e, q in enumerate(quads):
        for v in q:
            node_elems[v].append(e)
    for e, q in enumerate(quads):
        near = set()
        for v in q:
            near.update(node_elems[v])
        for f in near:
            dd = np.linalg.norm(c[e] - c[f]) / R[f]
            for key, tab in tabs.items():
                dg = next(g for lim, g in tab if dd < lim)
                out[key] -= ((complexity + dg) // 2 + 1) ** 4
    return out


def run_galerkin_quad(args):
    """Quadrilateral Galerkin workload: Helmholtz single- and double-layer
    matrices with isoparametric (bilinear, nodal) fields on NiHu's blown-up
    cube sphere of 9600 quadrilaterals (40 x 40 per face, 9602 nodes), k = 5,
    NiHu's degree policy (Gauss degree 2 + the distance interval on both
    elements) in FP32, coincident / edge / corner pairs with the quad x quad
    rules in FP64, complex64 results.  A step assembles both 9602^2
    matrices."""
    import numpy as np
    import torch
    from nihu_amd import nbem
    from nihu_amd.mesh import cubed_sphere

    torch.cuda.set_device(0)
    nodes, quads = cubed_sphere(40)
    ne, nn = len(quads), len(nodes)
    k = 5.0
    pol = nbem.policy(rule="nihu")
    name, ncu, hbm = nbem.device_info()
    log("device %s, %d CUs; cube sphere %d quadrilaterals, %d nodes; k = %g" % (name, ncu, ne, nn, k))
    mesh = nbem.Mesh(nodes, quads)
    L, M = nbem.colmajor(nn, nn), nbem.colmajor(nn, nn)
    stream = torch.cuda.current_stream()

    def step():
        nbem.assemble(mesh, mesh, L, 1, 0, 0, k, pol=pol, accumulate=False, stream=stream,
                      formalism="galerkin_linear")
        nbem.assemble(mesh, mesh, M, 1, 0, 1, k, pol=pol, accumulate=False, stream=stream,
                      formalism="galerkin_linear")

    for w in range(args.warmup):
        step()
        torch.cuda.synchronize()
        log("warmup step %d done" % w)
    nbem.profile(True)
    nbem.profile_read(reset=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    prof = nbem.profile_read(reset=True)
    nbem.profile(False)
    ms_per_step = wall / args.steps * 1e3
    value = 2 * nn * nn * args.steps / wall
    pass_ms = {p: (v[0] / v[1] if v[1] else 0.0) for p, v in prof.items()}
    log("timed %d steps: %.3f ms/step, %.4g entries/s; per-launch device ms %s" % (args.steps, ms_per_step, value,
                                                                                 pass_ms))
    pp = galq_point_pairs(nodes, quads)
    reg_ms_pair = 2 * pass_ms["regular"]       # SLP and DLP launches per step
    flops = pp["SLP"] * FLOPS_PER_POINTPAIR_GALQ["SLP"] + pp["DLP"] * FLOPS_PER_POINTPAIR_GALQ["DLP"]
    ach = flops / (reg_ms_pair * 1e-3) / 1e12 if reg_ms_pair > 0 else 0.0
    roofline = {"bound": "valu", "achieved": round(ach, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                "frac": round(ach / PEAK_FP32_TFLOPS, 4), "traffic": None,
                "engine": "VALU fp32 (no MFMA; packed fp32 disabled at build)",
                "kernel": "galq_regular_kernel<HELM,{SLP,DLP},4>", "kernel_ms": round(reg_ms_pair, 4),
                "algorithmic_flops": flops, "point_pairs": pp}

    cpu = err = None
    if args.cpu_rows != 0:
        from oracle import oracle
        rng = np.random.default_rng(123)
        n_sample = 32 if args.cpu_rows < 0 else max(1, args.cpu_rows // 16)
        sample = np.sort(rng.choice(nn, n_sample, replace=False))
        elems = np.unique(np.nonzero(np.isin(quads, sample))[0]).astype(np.int32)
        opol = oracle.policy()
        threads = oracle.num_threads()
        tc = time.perf_counter()
        Lr = oracle.assemble_galerkin(nodes, quads, 1, 0, 0, k=k, pol=opol, linear=True, test_elems=elems,
                                      threads=threads)[sample]
        Mr = oracle.assemble_galerkin(nodes, quads, 1, 0, 1, k=k, pol=opol, linear=True, test_elems=elems,
                                      threads=threads)[sample]
        dt = time.perf_counter() - tc
        cpu_value = 2 * nn * nn / (dt * ne / len(elems))
        cpu = {"value": cpu_value, "unit": "matrix entries/s", "cores": threads, "kind": "port",
               "sample": "test elements around %d sampled nodes (%d of %d elements) against all %d trial "
                         "elements, SLP and DLP, FP64 oracle, %.1f s wall; rate scaled by %d/%d to the whole "
                         "matrices" % (len(sample), len(elems), ne, ne, dt, len(elems), ne),
               "host": host_cpu()}
        si = torch.from_numpy(sample.astype(np.int64)).cuda()
        Lg = L[si, :].cpu().numpy().astype(np.complex128)
        Mg = M[si, :].cpu().numpy().astype(np.complex128)
        err = {"rows": int(len(sample)),
               "L_rel_fro": float(np.linalg.norm(Lg - Lr) / np.linalg.norm(Lr)),
               "M_rel_fro": float(np.linalg.norm(Mg - Mr) / np.linalg.norm(Mr))}
        log("sampled-row error vs oracle: %s; cpu %.4g entries/s on %d threads" % (err, cpu_value, threads))

    return {
        "metric": "matrix entries/s, Helmholtz 3D Galerkin bilinear quads, 9600 elems",
        "value": value, "unit": "matrix entries/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 (complex64) regular pairs, fp64 singular pairs",
        "data": "synthetic: NiHu's blown-up cube sphere",
        "config": {"workload": "Helmholtz 3D Galerkin, isoparametric bilinear quadrilaterals, SLP and DLP matrices",
                   "n_elements": ne, "n_nodes": nn, "k": k,
                   "regular_rule": "NiHu degrees: Gauss degree 2 + distance interval, N x N points on both elements",
                   "singular": "coincident/edge/corner quad x quad rules, 4-point Gauss per axis, FP64"},
        "roofline": roofline, "cpu_baseline": cpu, "accuracy": err, "pass_ms": pass_ms,
    }


def run_laplace(args):
    """Laplace 3D collocation, the plain method in FP64: SLP and DLP as real
    float64 matrices on the 1280-triangle icosphere (level 3, +-1 % radial
    jitter, seed 0), the 7-point rule for every regular pair, no
    near-singular pass, analytic diagonal; every pair in FP64
    (precision="fp64").  A step assembles both 1280^2 matrices.  The CPU
    baseline is the oracle assembling the same two matrices whole."""
    import numpy as np
    import torch
    from nihu_amd import nbem
    from nihu_amd.mesh import jittered_icosphere

    torch.cuda.set_device(0)
    nodes, tris = jittered_icosphere(3, seed=0)
    n = len(tris)
    pol = nbem.policy(rule="fixed", degree=5, near="none", precision="fp64")
    mesh = nbem.Mesh(nodes, tris)
    S, D = nbem.colmajor(n, n, "float64"), nbem.colmajor(n, n, "float64")
    stream = torch.cuda.current_stream()

    def step():
        nbem.assemble(mesh, mesh, S, nbem.NB_LAPLACE, 0, 0, 0.0, pol=pol, accumulate=False, stream=stream)
        nbem.assemble(mesh, mesh, D, nbem.NB_LAPLACE, 0, 1, 0.0, pol=pol, accumulate=False, stream=stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    nbem.profile(True)
    nbem.profile_read(reset=True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    prof = nbem.profile_read(reset=True)
    nbem.profile(False)
    pass_ms = {p: (v[0] / v[1] if v[1] else 0.0) for p, v in prof.items()}
    ms_per_step = wall / args.steps * 1e3
    value = 2 * n * n * args.steps / wall
    log("laplace FP64: timed %d steps: %.3f ms/step, %.4g entries/s; per-launch device ms %s" % (
        args.steps, ms_per_step, value, pass_ms))
    from oracle import oracle
    threads = oracle.num_threads()
    build = oracle_tuned_build(oracle)
    opol = oracle.policy(rule=1, fixed_degree=5, near_mode=2)
    tc = time.perf_counter()
    Sr = oracle.assemble(nodes, tris, 0, 0, 0, pol=opol, threads=threads).real
    Dr = oracle.assemble(nodes, tris, 0, 0, 1, pol=opol, threads=threads).real
    dt = time.perf_counter() - tc
    oracle.use_build("exact")
    cpu_value = 2 * n * n / dt
    Sg, Dg = S.cpu().numpy(), D.cpu().numpy()
    err = {"SLP_rel_fro": float(np.linalg.norm(Sg - Sr) / np.linalg.norm(Sr)),
           "DLP_rel_fro": float(np.linalg.norm(Dg - Dr) / np.linalg.norm(Dr)),
           "SLP_max_abs": float(np.abs(Sg - Sr).max()), "DLP_max_abs": float(np.abs(Dg - Dr).max())}
    log("laplace FP64: error vs oracle %s; cpu %.4g entries/s on %d threads" % (err, cpu_value, threads))
    return {
        "metric": "matrix entries/s, Laplace 3D coll. FP64, 1280 tris",
        "value": value, "unit": "matrix entries/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp64 (every pair)", "data": "synthetic: seeded jittered icosphere",
        "config": {"workload": "Laplace 3D collocation, SLP and DLP real matrices, plain method in FP64",
                   "n_tri": n, "icosphere_level": 3, "regular_rule": "Dunavant degree 5 (7 points), every pair",
                   "near": "none", "diagonal": "analytic"},
        "roofline": None, "accuracy": err, "pass_ms": pass_ms,
        "cpu_baseline": {"value": cpu_value, "unit": "matrix entries/s", "cores": threads, "kind": "port",
                         "sample": "the whole workload (both matrices), FP64 oracle, OpenMP over rows, %.2f s wall"
                                   % dt, "build": ORACLE_BUILDS[build], "host": host_cpu()},
    }


def run_quad(args):
    """Quadrilateral workload: Burton-Miller A and B (SLP, DLP, DLPt, HSP) on
    NiHu's blown-up cube sphere (create_sphere_boundary.m: radius 1, 29
    elements along the radius, 58 x 58 warped bilinear quadrilaterals per
    cube face, 20184 elements), ka = 10, NiHu's policy: Gauss degree 1 + the
    distance table for regular pairs (FP32), polar / Telles near-singular
    pairs below 1.5 mean edge lengths and Duffy / Guiggiani diagonal (FP64).
    (With a +-1 % radial jitter, seed 0, NiHu's inverse mapping does not
    converge for a few near-singular pairs, where NiHu throws and the engine
    returns NB_ERR_RUNTIME.)"""
    import numpy as np
    import torch
    from nihu_amd import nbem
    from nihu_amd.mesh import sphere_boundary

    torch.cuda.set_device(0)
    nodes, quads = sphere_boundary(1.0, 29)
    n = len(quads)
    k = 10.0
    alpha = complex(0.0, -1.0 / k)
    pol = nbem.policy(**nbem.NIHU_POLICY)
    name, ncu, hbm = nbem.device_info()
    log("device %s, %d CUs; cubed sphere %d quadrilaterals; k = %g" % (name, ncu, n, k))
    mesh = nbem.Mesh(nodes, quads)
    A, B = nbem.colmajor(n, n), nbem.colmajor(n, n)
    stream = torch.cuda.current_stream()
    cls = nbem.classify(mesh, mesh, 1, 1, pol)
    n_near = int((cls == -2).sum().item())
    del cls
    n_reg = n * n - n_near - n
    if args.profile_copy_bytes > 0:
        copy_calibration(args.profile_copy_bytes)
    for w in range(args.warmup):
        nbem.assemble_bm(mesh, k, A, B, alpha=alpha, pol=pol, stream=stream)
        torch.cuda.synchronize()
        log("warmup step %d done" % w)
    nbem.profile(True)
    nbem.profile_read(reset=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        nbem.assemble_bm(mesh, k, A, B, alpha=alpha, pol=pol, stream=stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    prof = nbem.profile_read(reset=True)
    nbem.profile(False)
    ms_per_step = wall / args.steps * 1e3
    value = 4 * n * n * args.steps / wall
    pass_ms = {p: (v[0] / v[1] if v[1] else 0.0) for p, v in prof.items()}
    log("timed %d steps: %.3f ms/step, %.4g entries/s; per-pass device ms %s" % (args.steps, ms_per_step, value,
                                                                                pass_ms))
    reg_ms = pass_ms["regular"]
    write_bytes = 2 * 8 * n_reg
    gbs = write_bytes / (reg_ms * 1e-3) / 1e9 if reg_ms > 0 else 0.0
    roofline = {"bound": "hbm", "achieved": round(gbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(gbs / PEAK_HBM_GBS, 4),
                "traffic": read_traffic(os.path.join(ROOT, "profiles", "quad_regular_cubed58_traffic.txt")),
                "kernel": "quad_regular_kernel<HELM,ALL4,OUT_BM>", "kernel_ms": round(reg_ms, 4),
                "algorithmic_bytes": write_bytes}
    cpu = err = None
    if args.cpu_rows != 0:
        from oracle import oracle
        nr = 512 if args.cpu_rows < 0 else args.cpu_rows
        rows = np.sort(np.random.default_rng(123).choice(n, nr, replace=False)).astype(np.int32)
        opol = oracle.policy(rule=0, fixed_degree=0, near_mode=0, near_limit=1.5)
        threads = oracle.num_threads()
        tc = time.perf_counter()
        Ar, Br = oracle.assemble_bm(nodes, quads, k, alpha=alpha, pol=opol, rows=rows, threads=threads)
        dt = time.perf_counter() - tc
        cpu_value = 4 * nr * n / dt
        cpu = {"value": cpu_value, "unit": "matrix entries/s", "cores": threads, "kind": "port",
               "sample": "%d of %d rows of A and B (4 operators), FP64 oracle, OpenMP over rows, %.1f s wall"
                         % (nr, n, dt), "host": host_cpu()}
        ri = torch.from_numpy(rows.astype(np.int64)).cuda()
        Ag, Bg = A[ri, :].cpu().numpy(), B[ri, :].cpu().numpy()
        err = {"rows": nr, "A_rel_fro": float(np.linalg.norm(Ag - Ar) / np.linalg.norm(Ar)),
               "B_rel_fro": float(np.linalg.norm(Bg - Br) / np.linalg.norm(Br)),
               "A_max_abs": float(np.abs(Ag - Ar).max()), "B_max_abs": float(np.abs(Bg - Br).max())}
        log("sampled-row error vs oracle: %s; cpu %.4g entries/s on %d threads" % (err, cpu_value, threads))
    out = {
        "metric": "matrix entries/s, Helmholtz 3D coll. quadrilaterals, 20184 elems",
        "value": value, "unit": "matrix entries/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 (complex64) regular pairs, fp64 near-singular and diagonal",
        "data": "synthetic: NiHu's blown-up cube sphere (create_sphere_boundary.m, r0 = 1, n0 = 29)",
        "config": {"workload": "Helmholtz 3D collocation on bilinear quadrilaterals, Burton-Miller A and B",
                   "n_elements": n, "k": k, "policy": "NiHu (Gauss degree 1 + distance table, "
                   "near-singular d/R < 1.5: polar SLP, Telles DLP/DLPt/HSP; diagonal Duffy, Guiggiani HSP)",
                   "pairs": {"regular": n_reg, "near": n_near, "diagonal": n}},
        "roofline": roofline, "cpu_baseline": cpu, "accuracy": err, "pass_ms": pass_ms,
    }
    return out


def run_gauss(args):
    """Gauss-field workload: Burton-Miller A and B with NiHu's quad_1_gauss_field
    (four dofs per quadrilateral at the 2 x 2 Gauss points, collocated there;
    the default build of nihu/applications/helmholtz/helmholtz_3d_coll.cpp)
    on NiHu's blown-up cube sphere with 30 x 30 quadrilaterals per face
    (create_sphere_boundary.m, n0 = 15: 5400 elements, 21600 dofs), ka = 10,
    NiHu's policy: Gauss degree 2 + the distance table per element pair
    (FP32), polar / Telles near-singular blocks below 1.5 mean edge lengths
    and Duffy / Guiggiani at the four nodes of each diagonal block (FP64)."""
    import numpy as np
    import torch
    from nihu_amd import nbem
    from nihu_amd.mesh import sphere_boundary

    torch.cuda.set_device(0)
    nodes, quads = sphere_boundary(1.0, 15)
    ne = len(quads)
    nd = 4 * ne
    k = 10.0
    alpha = complex(0.0, -1.0 / k)
    pol = nbem.policy(**nbem.NIHU_POLICY)
    name, ncu, hbm = nbem.device_info()
    log("device %s, %d CUs; cubed sphere %d quadrilaterals, %d gauss dofs; k = %g" % (name, ncu, ne, nd, k))
    mesh = nbem.Mesh(nodes, quads)
    A, B = nbem.colmajor(nd, nd), nbem.colmajor(nd, nd)
    stream = torch.cuda.current_stream()
    cls = nbem.classify(mesh, mesh, 1, 1, pol)
    n_near = int((cls == -2).sum().item())
    del cls
    n_reg_blocks = ne * ne - n_near - ne
    if args.profile_copy_bytes > 0:
        copy_calibration(args.profile_copy_bytes)
    for w in range(args.warmup):
        nbem.assemble_bm_gauss(mesh, k, A, B, alpha=alpha, pol=pol, stream=stream)
        torch.cuda.synchronize()
        log("warmup step %d done" % w)
    nbem.profile(True)
    nbem.profile_read(reset=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        nbem.assemble_bm_gauss(mesh, k, A, B, alpha=alpha, pol=pol, stream=stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    prof = nbem.profile_read(reset=True)
    nbem.profile(False)
    ms_per_step = wall / args.steps * 1e3
    value = 4 * nd * nd * args.steps / wall
    pass_ms = {p: (v[0] / v[1] if v[1] else 0.0) for p, v in prof.items()}
    log("timed %d steps: %.3f ms/step, %.4g entries/s; per-pass device ms %s" % (args.steps, ms_per_step, value,
                                                                                pass_ms))
    reg_ms = pass_ms["regular"]
    # A and B, complex64, 16 entries per regular element-pair block
    write_bytes = 2 * 8 * 16 * n_reg_blocks
    gbs = write_bytes / (reg_ms * 1e-3) / 1e9 if reg_ms > 0 else 0.0
    roofline = {"bound": "hbm", "achieved": round(gbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(gbs / PEAK_HBM_GBS, 4),
                "traffic": read_traffic(os.path.join(ROOT, "profiles", "gauss_regular_cubed30_traffic.txt")),
                "kernel": "gauss_regular_kernel<HELM,ALL4,OUT_BM>", "kernel_ms": round(reg_ms, 4),
                "algorithmic_bytes": write_bytes}
    cpu = err = None
    if args.cpu_rows != 0:
        from oracle import oracle
        nr = 512 if args.cpu_rows < 0 else args.cpu_rows
        rows = np.sort(np.random.default_rng(123).choice(nd, nr, replace=False)).astype(np.int32)
        opol = oracle.policy(rule=0, fixed_degree=0, near_mode=0, near_limit=1.5)
        threads = oracle.num_threads()
        tc = time.perf_counter()
        Ar, Br = oracle.assemble_bm_gauss(nodes, quads, k, alpha=alpha, pol=opol, rows=rows, threads=threads)
        dt = time.perf_counter() - tc
        cpu_value = 4 * nr * nd / dt
        cpu = {"value": cpu_value, "unit": "matrix entries/s", "cores": threads, "kind": "port",
               "sample": "%d of %d rows of A and B (4 operators), FP64 oracle, OpenMP over rows, %.1f s wall"
                         % (nr, nd, dt), "host": host_cpu()}
        ri = torch.from_numpy(rows.astype(np.int64)).cuda()
        Ag, Bg = A[ri, :].cpu().numpy(), B[ri, :].cpu().numpy()
        err = {"rows": nr, "A_rel_fro": float(np.linalg.norm(Ag - Ar) / np.linalg.norm(Ar)),
               "B_rel_fro": float(np.linalg.norm(Bg - Br) / np.linalg.norm(Br)),
               "A_max_abs": float(np.abs(Ag - Ar).max()), "B_max_abs": float(np.abs(Bg - Br).max())}
        log("sampled-row error vs oracle: %s; cpu %.4g entries/s on %d threads" % (err, cpu_value, threads))
    return {
        "metric": "matrix entries/s, Helmholtz 3D coll. quadrilateral gauss field, 21600 dofs",
        "value": value, "unit": "matrix entries/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 (complex64) regular pairs, fp64 near-singular and diagonal",
        "data": "synthetic: NiHu's blown-up cube sphere (create_sphere_boundary.m, r0 = 1, n0 = 15)",
        "config": {"workload": "Helmholtz 3D collocation with the quadrilateral gauss field, Burton-Miller A and B",
                   "n_elements": ne, "n_dofs": nd, "k": k,
                   "policy": "NiHu (Gauss degree 2 + distance table per element pair, near-singular d/R < 1.5: "
                             "polar SLP, Telles DLP/DLPt/HSP; diagonal Duffy, Guiggiani HSP at the nodes)",
                   "element_pairs": {"regular": n_reg_blocks, "near": n_near, "diagonal": ne}},
        "roofline": roofline, "cpu_baseline": cpu, "accuracy": err, "pass_ms": pass_ms,
    }


def run_mixed(args):
    """Both element types in one mesh: Burton-Miller A and B on the jittered
    unit blown-up cube sphere (NiHu's create_sphere_boundary.m) with 50 x 50
    quadrilaterals per face and every third one split into two triangles
    (10000 triangles and 10000 quadrilaterals, interleaved in NiHu's element
    table), ka = 10, NiHu's policy.  The engine assembles the four (test
    type, trial type) blocks and scatters them to the element ids."""
    import numpy as np
    import torch
    from nihu_amd import nbem
    from nihu_amd.mesh import mixed_cubed_sphere

    torch.cuda.set_device(0)
    nodes, table = mixed_cubed_sphere(50, jitter=0.0)
    n = len(table)
    k = 10.0
    alpha = complex(0.0, -1.0 / k)
    pol = nbem.policy(**nbem.NIHU_POLICY)
    name, ncu, hbm = nbem.device_info()
    n_tri = int((table[:, 0] == nbem.NB_TRIA_1_ID).sum())
    log("device %s, %d CUs; mixed cubed sphere %d elements (%d triangles); k = %g" % (name, ncu, n, n_tri, k))
    mesh = nbem.Mesh.from_elements(nodes, table)
    A, B = nbem.colmajor(n, n), nbem.colmajor(n, n)
    stream = torch.cuda.current_stream()
    for w in range(args.warmup):
        nbem.assemble_bm(mesh, k, A, B, alpha=alpha, pol=pol, stream=stream)
        torch.cuda.synchronize()
        log("warmup step %d done" % w)
    nbem.profile(True)
    nbem.profile_read(reset=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        nbem.assemble_bm(mesh, k, A, B, alpha=alpha, pol=pol, stream=stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    prof = nbem.profile_read(reset=True)
    nbem.profile(False)
    ms_per_step = wall / args.steps * 1e3
    value = 4 * n * n * args.steps / wall
    pass_ms = {p: (v[0] / v[1] if v[1] else 0.0) for p, v in prof.items()}
    log("timed %d steps: %.3f ms/step, %.4g entries/s; per-pass device ms (per block) %s"
        % (args.steps, ms_per_step, value, pass_ms))
    cpu = err = None
    if args.cpu_rows != 0:
        from oracle import oracle
        nr = 256 if args.cpu_rows < 0 else args.cpu_rows
        rows = np.sort(np.random.default_rng(123).choice(n, nr, replace=False))
        opol = oracle.policy(rule=0, fixed_degree=0, near_mode=0, near_limit=1.5)
        threads = oracle.num_threads()
        parts = [np.flatnonzero(table[:, 0] == t) for t in (nbem.NB_TRIA_1_ID, nbem.NB_QUAD_1_ID)]
        npe = (3, 4)
        Ar = np.zeros((nr, n), dtype=np.complex128)
        Br = np.zeros((nr, n), dtype=np.complex128)
        tc = time.perf_counter()
        # the oracle's single-type operators, placed by NiHu's rules for such a
        # mesh: a type against itself on the same mesh, against the other type
        # as different meshes; element id = table row
        for a_, ia in enumerate(parts):
            sel = np.flatnonzero(np.isin(rows, ia))
            if sel.size == 0:
                continue
            loc = np.searchsorted(ia, rows[sel]).astype(np.int32)
            ea = table[ia, 1:1 + npe[a_]]
            for b_, ib in enumerate(parts):
                eb = table[ib, 1:1 + npe[b_]]
                trial = None if a_ == b_ else (nodes, eb)
                op = {(nx, ny): oracle.assemble(nodes, ea, 1, nx, ny, k=k, pol=opol, rows=loc, trial=trial,
                                                threads=threads)
                      for nx, ny in ((0, 0), (0, 1), (1, 0), (1, 1))}
                Ar[np.ix_(sel, ib)] = alpha * op[(1, 1)] + op[(0, 1)]
                Br[np.ix_(sel, ib)] = op[(0, 0)] + alpha * op[(1, 0)]
        Ar[np.arange(nr), rows] -= 0.5
        Br[np.arange(nr), rows] += 0.5 * alpha
        dt = time.perf_counter() - tc
        cpu_value = 4 * nr * n / dt
        cpu = {"value": cpu_value, "unit": "matrix entries/s", "cores": threads, "kind": "port",
               "sample": "%d of %d rows of A and B (4 operators, per element-type block), FP64 oracle, OpenMP "
                         "over rows, %.1f s wall" % (nr, n, dt), "host": host_cpu()}
        ri = torch.from_numpy(rows.astype(np.int64)).cuda()
        Ag, Bg = A[ri, :].cpu().numpy(), B[ri, :].cpu().numpy()
        err = {"rows": nr, "A_rel_fro": float(np.linalg.norm(Ag - Ar) / np.linalg.norm(Ar)),
               "B_rel_fro": float(np.linalg.norm(Bg - Br) / np.linalg.norm(Br)),
               "A_max_abs": float(np.abs(Ag - Ar).max()), "B_max_abs": float(np.abs(Bg - Br).max())}
        log("sampled-row error vs oracle: %s; cpu %.4g entries/s on %d threads" % (err, cpu_value, threads))
    out = {
        "metric": "matrix entries/s, Helmholtz 3D coll. triangles + quadrilaterals, 20000 elems",
        "value": value, "unit": "matrix entries/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 (complex64) regular pairs, fp64 near-singular and diagonal",
        "data": "synthetic: NiHu's blown-up cube sphere, every third quadrilateral split",
        "config": {"workload": "Helmholtz 3D collocation on a mesh with both element types, Burton-Miller A and B",
                   "n_elements": n, "n_triangles": n_tri, "k": k, "mesh_seed": 0,
                   "policy": "NiHu (distance-table degrees, near-singular d/R < 1.5)"},
        "roofline": None, "cpu_baseline": cpu, "accuracy": err, "pass_ms": pass_ms,
    }
    return out


def sweep_summary(steps=5, warmup=2, log_prefix="sweep", policy="config"):
    """Burton-Miller A and B for 64 wavenumbers ka uniformly spaced in
    [0.5, 32] on the 5120-triangle jittered icosphere (level 4, seed 0) in one
    persistent launch with the autotuned launch shape -- policy "config":
    7-point rule far, FP64 within 2 element diameters; "nihu": NiHu's
    distance-driven degrees, FP64 within 1.5 mean edges -- compared with the
    same matrices from 64 single-wavenumber assemblies, and 16 seeded rows of 8
    wavenumbers (the first and last among them) against the FP64 oracle's
    sweep."""
    import numpy as np
    import torch
    from nihu_amd import nbem
    from nihu_amd.mesh import jittered_icosphere

    level = 4
    nodes, tris = jittered_icosphere(level, seed=0)
    n = len(tris)
    ks = np.linspace(0.5, 32.0, 64)
    pol = nbem.policy(**nbem.CONFIG_POLICY) if policy == "config" else nbem.policy("nihu")
    mesh = nbem.Mesh(nodes, tris)
    A = torch.zeros((len(ks), n, n), dtype=torch.complex64, device="cuda").transpose(1, 2)
    B = torch.zeros_like(A)
    t = time.perf_counter()
    best = nbem.autotune(mesh, len(ks), pol)
    tune_s = time.perf_counter() - t
    log("%s: autotune %.2f s: tile_cols %d, k_batch %d, persistent %d" % (log_prefix, tune_s, best.tile_cols,
                                                                       best.k_batch, best.persistent))
    for _ in range(warmup):
        nbem.assemble_sweep(mesh, ks, A, B, pol=pol)
        torch.cuda.synchronize()
    nbem.profile(True)
    nbem.profile_read(reset=True)
    lat = []
    for _ in range(steps):
        t = time.perf_counter()
        nbem.assemble_sweep(mesh, ks, A, B, pol=pol)
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - t)
    prof = nbem.profile_read(reset=True)
    wall = float(np.sum(lat))
    entries = 4 * n * n * len(ks)
    value = entries * steps / wall
    log("%s: %.3f ms per 64-wavenumber sweep, %.4g entries/s; passes %s" % (log_prefix, wall / steps * 1e3, value,
                                                                          prof))
    # the same matrices, one Burton-Miller assembly per wavenumber
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for f, k in enumerate(ks):
        nbem.assemble_bm(mesh, float(k), A[f], B[f], pol=pol)
    torch.cuda.synchronize()
    per_freq = entries / (time.perf_counter() - t0)
    nbem.profile(False)
    log("%s: per-frequency assemblies %.4g entries/s" % (log_prefix, per_freq))
    nbem.assemble_sweep(mesh, ks, A, B, pol=pol)
    torch.cuda.synchronize()

    from oracle import oracle
    rows = np.sort(np.random.default_rng(5).choice(n, 16, replace=False)).astype(np.int32)
    fsel = [0, 9, 18, 27, 36, 45, 54, 63]
    opol = oracle.policy(rule=1, fixed_degree=5, near_mode=1, near_limit=2.0) if policy == "config" else oracle.policy()
    threads = oracle.num_threads()
    build = oracle_tuned_build(oracle)
    tc = time.perf_counter()
    Ar, Br = oracle.sweep_bm(nodes, tris, ks[fsel], pol=opol, rows=rows, threads=threads)
    dt = time.perf_counter() - tc
    oracle.use_build("exact")
    cpu_value = 4 * len(rows) * n * len(fsel) / dt
    ri = torch.from_numpy(rows.astype(np.int64)).cuda()
    errs = []
    for t_, f in enumerate(fsel):
        a, b = A[f][ri, :].cpu().numpy(), B[f][ri, :].cpu().numpy()
        errs.append(max(np.linalg.norm(a - Ar[t_]) / np.linalg.norm(Ar[t_]),
                        np.linalg.norm(b - Br[t_]) / np.linalg.norm(Br[t_])))
    log("%s: max rel. Frobenius error over %d wavenumbers (16 rows): %.3e" % (log_prefix, len(fsel), max(errs)))
    del A, B
    torch.cuda.empty_cache()
    reg_ms = prof["sweep"][0] / max(prof["sweep"][1], 1)
    return {
        "workload": "Burton-Miller sweep, 64 wavenumbers ka in [0.5, 32], 5120-triangle jittered icosphere, " +
                    ("7-point rule far, FP64 within 2 element diameters" if policy == "config" else
                     "NiHu's distance-driven degrees, FP64 within 1.5 mean edges") + ", one persistent launch",
        "policy": policy,
        "n_tri": n, "n_k": len(ks), "steps": steps,
        "ms_per_sweep": wall / steps * 1e3,
        "latency_ms": {"p50": float(np.percentile(lat, 50)) * 1e3, "p99": float(np.percentile(lat, 99)) * 1e3},
        "entries_per_s": value,
        "per_frequency_entries_per_s": per_freq,
        "speedup_vs_per_frequency": value / per_freq,
        "regular_sweep_kernel_ms": reg_ms,
        "launch": {"tile_cols": best.tile_cols, "k_batch": best.k_batch, "persistent": best.persistent,
                   "autotune_s": tune_s},
        "accuracy": {"max_rel_fro_over_k": float(max(errs)), "rows": len(rows), "wavenumbers": len(fsel)},
        "cpu_baseline": {"value": cpu_value, "unit": "matrix entries/s", "cores": threads, "kind": "port",
                         "sample": "%d of %d rows, %d of the 64 wavenumbers, FP64 oracle (NiHu's algorithm), OpenMP "
                                   "over frequencies, %.1f s wall" % (len(rows), n, len(fsel), dt),
                         "build": ORACLE_BUILDS[build]},
    }


def run_sweep(args):
    import torch
    torch.cuda.set_device(0)
    sw = sweep_summary(steps=args.steps, warmup=max(1, min(args.warmup, 5)), policy=args.policy)
    out = {
        "metric": "matrix entries/s, Helmholtz 3D Burton-Miller sweep, 64 wavenumbers, 5120 tris",
        "value": sw["entries_per_s"], "unit": "matrix entries/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": sw["ms_per_sweep"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32 (complex64) far pairs, fp64 near/diagonal",
        "data": "synthetic: seeded jittered icosphere",
        "config": {"workload": sw["workload"], "n_tri": sw["n_tri"], "n_k": sw["n_k"], "policy": sw["policy"],
                   "launch": sw["launch"]},
        "sweep": sw,
    }
    print(json.dumps(out))


def _serve_stream(srv, n_req, rate, seed):
    from nihu_amd import serve
    reqs = serve.poisson_stream(n_req, rate=rate, seed=seed)
    t0 = time.perf_counter()
    res = srv.run(reqs, realtime=True)
    wall = time.perf_counter() - t0
    failed = [r for r in res if r.status != "ok"]
    for r in failed:
        log("serve: request %d (%s, k %.3f) %s: %s" % (r.rid, r.mesh, r.k, r.status, r.error))
    res = [r for r in res if r.status == "ok"]
    lat = [r.latency for r in res]
    # the server kept up when no backlog built: the last request is done
    # within a quarter of the stream's arrival span (at least 0.5 s) after the
    # last arrival -- a server at capacity C under a rate 1.25 C or more would
    # need that long to drain what piled up -- and every request was solved
    span = max(r.arrival for r in reqs) - min(r.arrival for r in reqs)
    drain = (max(r.done for r in res) - max(r.arrival for r in reqs)) if res else float("inf")
    return {"rate_per_s": rate, "requests": len(res) + len(failed), "failed": len(failed), "wall_s": wall,
            "solves_per_s": len(res) / wall,
            "p50_s": serve.percentile(lat, 50), "p99_s": serve.percentile(lat, 99),
            "drain_s": drain,
            "kept_up": drain <= max(0.25 * span, 0.5) and not failed,
            "by_mesh_p50_s": {k: serve.percentile([r.latency for r in res if r.mesh == k], 50)
                              for k in sorted({r.mesh for r in res})}}


def _stream_len(n_req, rate):
    """requests of a stream at `rate`: at least n_req, else about 3 s of arrivals (at most 400)"""
    return max(n_req, min(int(3 * rate), 400))


def serve_summary(n_req=24, rate=4.0, refine=0, saturation=True, log_prefix="serve"):
    """Config 5: Poisson stream of Burton-Miller assembly + solve requests
    (GMRES to a relative residual of 1e-8, as NiHu's application solves),
    mesh uniform over the 1280 / 5120 / 20480-triangle jittered icospheres
    (resident on the device), ka uniform in [1, 40], plane wave of seeded
    random direction, admission by complex64 footprint, 4 slots.  The
    batching server (requests on one mesh in one sweep launch) and the
    per-request path (one assembly per request) serve the same stream at the
    offered rate and at a loaded rate (8 x the offered one); then, for each, a
    rate sweep (doubling from the offered rate, about 3 s of arrivals each)
    stops at the first rate the server does not keep up with (a backlog left:
    the last completion more than a quarter of the arrival span, and more
    than 0.5 s, after the last arrival)."""
    from nihu_amd import serve
    from nihu_amd.mesh import jittered_icosphere

    meshes = {"L3": jittered_icosphere(3, seed=0), "L4": jittered_icosphere(4, seed=0),
              "L5": jittered_icosphere(5, seed=0)}
    out = {"workload": "Poisson stream of Burton-Miller assembly + GMRES solve requests, meshes 1280 / 5120 / "
                       "20480 triangles, ka U[1, 40], 4 slots, admission by HBM footprint",
           "meshes_n_tri": {k: len(v[1]) for k, v in meshes.items()}, "refine_steps": refine}
    for name, mb in (("batched", 8), ("per_request", 1)):
        srv = serve.Server(meshes, slots=4, refine=refine, max_batch=mb)
        warm = [serve.Request(-1 - i, key, 5.0 + i, (0.0, 0.0, 1.0)) for i, key in enumerate(meshes)]
        srv.run(warm, realtime=False)
        r = _serve_stream(srv, n_req, rate, seed=11)
        log("%s (%s, max_batch %d): %d requests at %.1f/s in %.2f s; p50 %.3f s p99 %.3f s" % (
            log_prefix, name, mb, r["requests"], rate, r["wall_s"], r["p50_s"], r["p99_s"]))
        out[name] = r
        loaded = 8 * rate
        r8 = _serve_stream(srv, _stream_len(n_req, loaded), loaded, seed=13)
        log("%s (%s, max_batch %d): %d requests at %.1f/s in %.2f s; p50 %.3f s p99 %.3f s" % (
            log_prefix, name, mb, r8["requests"], loaded, r8["wall_s"], r8["p50_s"], r8["p99_s"]))
        out[name + "_loaded"] = r8
        if saturation:
            sweep = []
            rt = rate
            while rt <= 64 * rate:
                rt *= 2
                q = _serve_stream(srv, _stream_len(n_req, rt), rt, seed=12)
                log("%s: rate %.1f/s -> %.2f solves/s, p50 %.3f s p99 %.3f s, kept up %s" % (
                    log_prefix, rt, q["solves_per_s"], q["p50_s"], q["p99_s"], q["kept_up"]))
                sweep.append({k: q[k] for k in ("rate_per_s", "solves_per_s", "p50_s", "p99_s", "drain_s", "kept_up")})
                if not q["kept_up"]:
                    break
            kept = [x["rate_per_s"] for x in sweep if x["kept_up"]]
            sfx = "" if name == "batched" else "_" + name
            out["rate_sweep" + sfx] = sweep
            out["max_sustained_rate_per_s" + sfx] = max([rate] + kept) if out[name]["kept_up"] else None
            out["saturation_solves_per_s" + sfx] = max(x["solves_per_s"] for x in sweep)
        del srv
    out["p50_s"], out["p99_s"] = out["batched"]["p50_s"], out["batched"]["p99_s"]
    out["p50_speedup_vs_per_request"] = out["per_request"]["p50_s"] / out["batched"]["p50_s"]
    out["loaded_p50_speedup_vs_per_request"] = out["per_request_loaded"]["p50_s"] / out["batched_loaded"]["p50_s"]
    if saturation:
        out["saturation_speedup_vs_per_request"] = (out["saturation_solves_per_s"] /
                                                    out["saturation_solves_per_s_per_request"])
    out["solver"] = "GMRES, relative residual 1e-8 (NiHu's Burton-Miller application), FP64-accumulating products"
    return out


def run_serve(args):
    import torch
    torch.cuda.set_device(0)
    sv = serve_summary(n_req=args.requests, rate=args.rate, refine=args.refine)
    out = {
        "metric": "solves/s and p50/p99 latency, Poisson stream of Burton-Miller assembly + GMRES solves",
        "value": sv["batched"]["solves_per_s"], "unit": "solves/s", "n_gpus": 1, "steps": sv["batched"]["requests"],
        "warmup": 3, "ms_per_step": sv["batched"]["wall_s"] / max(sv["batched"]["requests"], 1) * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 (complex64) far pairs and LU, fp64 near/diagonal", "data": "synthetic: seeded Poisson stream",
        "config": {"workload": sv["workload"], "meshes_n_tri": sv["meshes_n_tri"], "rate_per_s": args.rate,
                   "slots": 4, "refine_steps": args.refine},
        "latency_s": {"p50": sv["p50_s"], "p99": sv["p99_s"]},
        "serve": sv,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
