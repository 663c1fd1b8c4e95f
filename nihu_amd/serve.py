"""Continuous-batching server for assembly-plus-solve requests on one GPU.

A request asks for the rigid-body scattered surface pressure of a plane wave
p_inc = exp(-i k d.x) (unit amplitude, direction d) on one of the resident
meshes: the Burton-Miller collocation system
    A p = B q,   A = alpha*D - I/2 + H,   B = G + alpha*Ht + alpha/2 I,
    alpha = -i/k,  q = -dp_inc/dn = i k (d.n) exp(-i k d.x)
(the composition of nihu/applications/helmholtz/helmholtz_3d_coll.cpp:35-120)
is assembled by the HIP engine and solved on the device, by default as that
application does, with unpreconditioned GMRES to a relative residual of 1e-8
(nihu_amd/solve.py: products with the complex64 matrices by the engine's
FP64-accumulating kernel, complex128 Krylov basis), or by a complex64 LU
factorisation (solver="lu", optionally with FP64 iterative refinement).

Scheduling: requests arrive over time (their `arrival` offsets); the loop
admits the oldest waiting request as soon as a slot (a HIP stream with a host
worker thread) is free and its HBM footprint fits the budget, together with
up to max_batch - 1 further arrived requests on the same mesh (continuous
batching: their matrices come from one nb_assemble_sweep launch, the
frequency-sweep pattern of benchmark/EAA/sphere/homogeneous/bem/bem.cpp:91-115,
and their GMRES iterations run in lockstep, one host wait per step for the
batch; a byte cap keeps the large meshes, whose assembly is a small part of
their solve, at one request per launch),
hands the batch's whole pipeline to that slot's thread, which launches it on
the slot's stream, and goes on admitting while earlier batches are still on
the device, so requests of very different sizes overlap.  The threads
matter: the LU factorisation call returns to the host only when the
factorisation is done (0.44 s at 20480 unknowns), and launched from the loop
itself it would hold back every other request meanwhile.  Completion is
detected per request by events (each request of a batch is done when its
own solve is); latency = completion time - arrival time.  Results do not
depend on the mix or the batch: every kernel computes each matrix entry in a
fixed order without atomics, the sweep computes each wavenumber's matrices
independently of the others in its launch, and each request's LU runs on
its own buffers.

Failures stay with their request: an invalid request (unknown mesh, a
wavenumber that is not finite and positive, a direction that is not a unit
vector) is rejected before it is queued; one not admitted within its
deadline leaves the queue as expired; a batch whose launch raises (the
engine's error) ends its requests with that error and frees its slot and
memory; a solver that does not converge, or a solution holding NaN or Inf
(scanned on the device), ends that request alone with an error.  Every
request ends with exactly one Result, and the loop goes on with the others.
"""

from __future__ import annotations

import dataclasses
import time
from collections import deque
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import nbem, solve


@dataclasses.dataclass
class Request:
    rid: int
    mesh: str              # key of a resident mesh
    k: float               # wavenumber
    direction: tuple       # unit propagation direction of the plane wave
    arrival: float = 0.0   # seconds after the start of the run
    deadline: float | None = None   # seconds after arrival by which it must be admitted (None: no limit)


@dataclasses.dataclass
class Result:
    rid: int
    mesh: str
    k: float
    n: int
    arrival: float
    admitted: float
    done: float
    p: np.ndarray | None   # scattered surface pressure at the element centres (None unless status is "ok")
    status: str = "ok"     # "ok", "rejected" (invalid request), "expired" (not admitted by its deadline),
                           # "error" (assembly failed, solver did not converge, or a non-finite solution)
    error: str = ""

    @property
    def latency(self):
        return self.done - self.arrival


def check_request(req: Request, meshes) -> str:
    """Why a request cannot be served ("" if it can): unknown mesh, a
    wavenumber that is not finite and positive, or a direction that is not a
    finite unit 3-vector (to 1e-6)."""
    if req.mesh not in meshes:
        return "unknown mesh %r" % (req.mesh,)
    if not (np.isfinite(req.k) and req.k > 0):
        return "wavenumber must be finite and positive, got %r" % (req.k,)
    d = np.asarray(req.direction, dtype=np.float64)
    if d.shape != (3,) or not np.all(np.isfinite(d)) or abs(np.linalg.norm(d) - 1.0) > 1e-6:
        return "direction must be a finite unit 3-vector"
    if req.deadline is not None and not (req.deadline >= 0):
        return "deadline must be non-negative"
    return ""


def footprint_bytes(n: int, solver: str = "lu", restart: int = 3000) -> int:
    """Device bytes a request holds while admitted: A and B (complex64 n x n
    each) and, for LU, the factor (complex64 n x n) plus vectors and pivots;
    for GMRES the complex128 Krylov basis of min(restart, n) + 1 vectors and
    the product's work buffer instead of the factor."""
    if solver == "gmres":
        return 2 * 8 * n * n + 16 * n * (min(restart, n) + 1) + 16 * n * 80 + 64 * n
    return 3 * 8 * n * n + 64 * n


class Server:
    def __init__(self, meshes: dict, slots: int = 4, hbm_budget: float | None = None, policy=None, refine: int = 0,
                 max_batch: int = 8, solver: str = "gmres", tol: float = 1e-8, restart: int = 3000,
                 maxiter: int | None = None, max_batch_bytes: float = 8 * 2 ** 30):
        """solver: "gmres" (NiHu's application: unpreconditioned GMRES, tol
        on the relative residual, restart, at most maxiter iterations, default
        the number of unknowns) or "lu" (complex64 LU).  refine:
        steps of iterative refinement of the complex64 LU solution (residual
        b - A p accumulated in complex128 over column blocks, correction
        solved with the same LU factor).  max_batch: requests on one mesh
        assembled in one sweep launch and solved together (1: one assembly
        per request); max_batch_bytes caps the batch's A and B matrices
        (complex64): large meshes, whose assembly is a small part of their
        solve, go one request per launch instead of holding one huge block."""
        if solver not in ("gmres", "lu"):
            raise ValueError("solver must be 'gmres' or 'lu'")
        self.solver = solver
        self.tol = float(tol)
        self.restart = int(restart)
        self.maxiter = None if maxiter is None else int(maxiter)
        self.refine = int(refine)
        self.max_batch = max(1, int(max_batch))
        self.max_batch_bytes = float(max_batch_bytes)
        import torch
        self.torch = torch
        self.meshes = {}
        self.geom = {}
        for key, (nodes, tris) in meshes.items():
            m = nbem.Mesh(nodes, tris)
            self.meshes[key] = m
            c, nu = _centres_normals(nodes, tris)
            self.geom[key] = (torch.from_numpy(c).cuda(), torch.from_numpy(nu).cuda())
        self.slots = [torch.cuda.Stream() for _ in range(slots)]
        if hbm_budget is None:
            free, _ = torch.cuda.mem_get_info()
            hbm_budget = 0.8 * free
        self.budget = float(hbm_budget)
        # the configuration policy: 7-point rule far, FP64 within 2 element
        # diameters; quadrilaterals keep NiHu's near-singular rule (d/R < 1.5),
        # the only one their near-singular methods accept
        self.policy = policy or nbem.policy(**nbem.CONFIG_POLICY)
        self.quad_policy = policy or nbem.policy(rule="fixed", degree=5, near="mean_edge", limit=1.5)
        for key, m in self.meshes.items():
            if self._footprint(m.n) > self.budget:
                raise ValueError("mesh %s does not fit the HBM budget" % key)
        # launch shape of the sweep per mesh, tuned once here for batches of
        # up to max_batch wavenumbers (so that no request pays for tuning)
        self.launch = {}
        for key, m in self.meshes.items():
            pol = self._policy(m)
            if m.nodes_per_element == 3:
                self.launch[key] = nbem.autotune(m, self.max_batch, pol)
            else:
                self.launch[key] = None
        torch.cuda.synchronize()

    def _sync(self):
        self.torch.cuda.synchronize()

    def _policy(self, mesh):
        return self.quad_policy if mesh.nodes_per_element != 3 else self.policy

    def _footprint(self, n):
        return footprint_bytes(n, self.solver, self.restart)

    def _launch(self, reqs, stream):
        """One batch of requests on one mesh: the Burton-Miller matrices of
        all their wavenumbers in one sweep launch, the right-hand sides, then
        the solves (GMRES: the batch's systems in lockstep; LU: one after the
        other).  Returns per request (event after its solve, its buffers
        with the solution last, device flag "the solution is finite", error
        message or "").  An exception fails the batch."""
        torch = self.torch
        key = reqs[0].mesh
        mesh = self.meshes[key]
        n = mesh.n
        c, nu = self.geom[key]
        ks = np.array([r.k for r in reqs], dtype=np.float64)
        out = []
        with torch.cuda.stream(stream):
            # column-major matrices, one per wavenumber
            A = torch.empty((len(reqs), n, n), dtype=torch.complex64, device="cuda").transpose(1, 2)
            B = torch.empty((len(reqs), n, n), dtype=torch.complex64, device="cuda").transpose(1, 2)
            # (the engine serialises calls on one mesh itself)
            nbem.assemble_sweep(mesh, ks, A, B, pol=self._policy(mesh), launch=self.launch[key], stream=stream)
            qs = []
            for req in reqs:
                d = torch.tensor(req.direction, dtype=torch.float64, device="cuda")
                phase = -req.k * (c @ d)
                pinc = torch.polar(torch.ones_like(phase), phase)
                qs.append((1j * req.k) * (nu @ d) * pinc)
            if self.solver == "gmres":
                # right-hand sides B q, then the batch's systems in lockstep
                # (one host wait per Arnoldi step for the whole batch)
                rhs = torch.stack([nbem.gemv(B[g], q) for g, q in enumerate(qs)])
                if len(reqs) > 1:
                    sols = solve.gmres_batched([A[g] for g in range(len(reqs))], rhs, tol=self.tol,
                                               restart=self.restart, maxiter=self.maxiter)
                else:
                    sols = [solve.gmres(A[0], rhs[0], tol=self.tol, restart=self.restart, maxiter=self.maxiter)]
                for res in sols:
                    err = "" if res.converged else (
                        "GMRES did not reach the relative residual %g (%g after %d iterations)"
                        % (self.tol, res.residual, res.iterations))
                    p = res.x
                    # NaN or Inf anywhere in the matrices or the right-hand
                    # side reaches the solution: one scan of it catches it
                    finite = torch.isfinite(torch.view_as_real(p)).all()
                    done = torch.cuda.Event()
                    done.record(stream)
                    # keep the buffers alive until the event has passed
                    out.append((done, (A, B, rhs, p), finite, err))
                return out
            for g, q in enumerate(qs):
                rhs = (B[g] @ q.to(torch.complex64)).unsqueeze(1)
                LU, piv, _ = torch.linalg.lu_factor_ex(A[g])
                p = torch.linalg.lu_solve(LU, piv, rhs).squeeze(1)
                if self.refine > 0:
                    p = _refine(A[g], LU, piv, rhs.squeeze(1), p, self.refine)
                finite = torch.isfinite(torch.view_as_real(p)).all()
                ev = torch.cuda.Event()
                ev.record(stream)
                out.append((ev, (A, B, LU, piv, rhs, p), finite, ""))
        return out

    def run(self, requests, realtime: bool = True):
        """Serve the requests; returns the results in completion order.
        realtime: wait for each arrival time (Poisson streams); otherwise all
        requests are available at once (arrival offsets still set latency).
        A request that fails -- invalid, not admitted by its deadline, its
        batch's launch raising, its solver not converging, a non-finite
        solution -- ends with that status and message; the others go on."""
        pending = deque()
        results = []
        for req in sorted(requests, key=lambda r: r.arrival):
            why = check_request(req, self.meshes)
            if why:
                results.append(Result(req.rid, req.mesh, req.k, 0, req.arrival, float("nan"), req.arrival, None,
                                      "rejected", why))
            else:
                pending.append(req)
        free_slots = list(range(len(self.slots)))
        in_use = 0.0
        active = []   # [batch, slot, future of per-request work, bytes, admitted, retired flags]
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0
        with ThreadPoolExecutor(max_workers=len(self.slots), thread_name_prefix="nbem-slot") as pool:
            while pending or active:
                # requests whose deadline for admission has passed leave the queue
                for req in [r for r in pending if r.deadline is not None and now() > r.arrival + r.deadline]:
                    pending.remove(req)
                    results.append(Result(req.rid, req.mesh, req.k, self.meshes[req.mesh].n, req.arrival,
                                          float("nan"), now(), None, "expired",
                                          "not admitted within %g s of its arrival" % req.deadline))
                # admit as many arrived requests as slots and memory allow: the
                # oldest one, with the next arrived requests on its mesh
                while pending and free_slots:
                    head = pending[0]
                    if realtime and head.arrival > now():
                        break
                    fp = self._footprint(self.meshes[head.mesh].n)
                    if in_use + fp > self.budget:
                        break
                    batch = [head]
                    n_head = self.meshes[head.mesh].n
                    cap = max(1, min(self.max_batch, int(self.max_batch_bytes // (16.0 * n_head * n_head))))
                    for req in list(pending)[1:]:
                        if len(batch) >= cap or (realtime and req.arrival > now()):
                            break
                        if req.mesh == head.mesh and in_use + fp * (len(batch) + 1) <= self.budget:
                            batch.append(req)
                    for req in batch:
                        pending.remove(req)
                    slot = free_slots.pop(0)
                    t_adm = now()
                    fut = pool.submit(self._launch, batch, self.slots[slot])
                    in_use += fp * len(batch)
                    active.append([batch, slot, fut, fp * len(batch), t_adm, [False] * len(batch)])
                # retire finished requests; a batch frees its slot and memory
                # when all of it is done (or its launch has failed)
                still = []
                for item in active:
                    batch, slot, fut, nbytes, t_adm, done = item
                    if fut.done():
                        exc = fut.exception()
                        if exc is not None:
                            for g, req in enumerate(batch):
                                results.append(Result(req.rid, req.mesh, req.k, self.meshes[req.mesh].n, req.arrival,
                                                      t_adm, now(), None, "error",
                                                      "%s: %s" % (type(exc).__name__, exc)))
                                done[g] = True
                        else:
                            for g, (ev, bufs, finite, err) in enumerate(fut.result()):
                                if not done[g] and ev.query():
                                    req = batch[g]
                                    p = bufs[-1].cpu().numpy()
                                    if not err and not bool(finite):
                                        err = "non-finite values in the solution"
                                    results.append(Result(req.rid, req.mesh, req.k, len(p), req.arrival, t_adm,
                                                          now(), None if err else p, "error" if err else "ok", err))
                                    done[g] = True
                        if all(done):
                            in_use -= nbytes
                            free_slots.append(slot)
                            continue
                    still.append(item)
                if len(still) == len(active) and still:
                    time.sleep(2e-4)
                active = still
                if not active and pending and realtime and pending[0].arrival > now():
                    time.sleep(min(1e-3, max(0.0, pending[0].arrival - now())))
        self._sync()
        return results


def _refine(A, LU, piv, b, p, steps, block=2048):
    """Iterative refinement of A p = b (complex64 A and LU factor): the
    residual b - A p in complex128, A upcast one column block at a time so
    that no complex128 copy of A is held, the correction solved in complex64
    with the same factor; returns p in complex128."""
    import torch
    p = p.to(torch.complex128)
    b = b.to(torch.complex128)
    n = A.shape[1]
    for _ in range(steps):
        r = b.clone()
        for j0 in range(0, n, block):
            j1 = min(n, j0 + block)
            r -= A[:, j0:j1].to(torch.complex128) @ p[j0:j1]
        d = torch.linalg.lu_solve(LU, piv, r.to(torch.complex64).unsqueeze(1)).squeeze(1)
        p = p + d.to(torch.complex128)
    return p


def _centres_normals(nodes, elems):
    """Collocation points and unit normals: triangle centroids and flat
    normals, or for bilinear quadrilaterals x(0, 0) and dx/dxi x dx/deta there."""
    X = nodes[elems]
    c = X.mean(axis=1)
    if elems.shape[1] == 3:
        nv = np.cross(X[:, 1] - X[:, 0], X[:, 2] - X[:, 0])
    else:
        nv = np.cross(X[:, 1] + X[:, 2] - X[:, 0] - X[:, 3], X[:, 2] + X[:, 3] - X[:, 0] - X[:, 1])
    return c, nv / np.linalg.norm(nv, axis=1)[:, None]


def poisson_stream(n_req: int, rate: float, seed: int = 0, meshes=("L3", "L4", "L5"), k_range=(1.0, 40.0)):
    """Seeded Poisson arrivals (rate per second), mesh uniform over `meshes`,
    k uniform in k_range, direction uniform on the sphere."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=n_req))
    reqs = []
    for i in range(n_req):
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        reqs.append(Request(i, str(rng.choice(list(meshes))), float(rng.uniform(*k_range)), tuple(v), float(t[i])))
    return reqs


def percentile(values, q):
    return float(np.percentile(np.asarray(values), q)) if len(values) else float("nan")


def mie_rigid_sphere_surface(k, a, cos_theta, n_terms=None):
    """Total surface pressure of a unit plane wave exp(-i k d.x) on a rigid
    sphere of radius a at angle theta from d (time convention exp(+i w t)):
    p = sum (2n+1) (-i)^n (-i / ((ka)^2 h_n'(ka))) P_n(cos theta),
    h_n = j_n - i y_n (spherical Hankel function of the second kind)."""
    from scipy.special import eval_legendre, spherical_jn, spherical_yn
    ka = k * a
    if n_terms is None:
        n_terms = int(ka + 4 * ka ** (1 / 3) + 10)
    p = np.zeros_like(np.asarray(cos_theta, dtype=float), dtype=complex)
    for n in range(n_terms + 1):
        hp = spherical_jn(n, ka, derivative=True) - 1j * spherical_yn(n, ka, derivative=True)
        p += (2 * n + 1) * (-1j) ** n * (-1j / (ka ** 2 * hp)) * eval_legendre(n, cos_theta)
    return p
