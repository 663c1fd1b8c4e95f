"""Failure handling of the serving loop (nihu_amd/serve.py), on the CPU: the
scheduler runs with its device launch replaced by a host stand-in that
injects faults into chosen requests -- a launch that raises, a solution with
NaN, a solver that does not converge -- beside invalid requests and expired
deadlines.  Every request must end with exactly one result of the right
status, the healthy ones with their solution, and the loop must free every
slot and byte of its budget (it would not terminate otherwise)."""

import time
import types

import numpy as np
import pytest
import torch

from nihu_amd import serve


class _Event:
    def __init__(self, delay):
        self.t = time.perf_counter() + delay

    def query(self):
        return time.perf_counter() >= self.t


class _HostServer(serve.Server):
    """The scheduler of serve.Server with a host stand-in for _launch: a
    request's 'solution' is k at every element (of the mesh's size)."""

    def __init__(self, faults, slots=2, max_batch=1, budget_requests=2, delay=2e-3):
        self.torch = torch
        self.solver = "gmres"
        self.restart = 3000
        self.max_batch = max_batch
        self.max_batch_bytes = 8 * 2 ** 30
        self.meshes = {"a": types.SimpleNamespace(n=10, nodes_per_element=3),
                       "b": types.SimpleNamespace(n=12, nodes_per_element=3)}
        self.slots = [None] * slots
        self.budget = budget_requests * float(serve.footprint_bytes(12, "gmres"))
        self.faults = faults
        self.delay = delay
        self.batches = []

    def _sync(self):
        pass

    def _launch(self, reqs, stream):
        self.batches.append([r.rid for r in reqs])
        if any(self.faults.get(r.rid) == "raise" for r in reqs):
            raise RuntimeError("injected launch failure")
        out = []
        for r in reqs:
            p = torch.full((self.meshes[r.mesh].n,), complex(r.k, 0.0), dtype=torch.complex128)
            err = ""
            if self.faults.get(r.rid) == "nan":
                p[3] = complex(float("nan"), 0.0)
            if self.faults.get(r.rid) == "inf":
                p[0] = complex(0.0, float("inf"))
            if self.faults.get(r.rid) == "diverge":
                err = "GMRES did not reach the relative residual 1e-08"
            out.append((_Event(self.delay), (p,), torch.isfinite(torch.view_as_real(p)).all(), err))
        return out


def _req(rid, mesh="a", k=None, direction=(1.0, 0.0, 0.0), arrival=0.0, deadline=None):
    return serve.Request(rid, mesh, float(rid + 1) if k is None else k, direction, arrival, deadline)


def _by_rid(results):
    rids = [r.rid for r in results]
    assert len(rids) == len(set(rids)), "a request ended twice"
    return {r.rid: r for r in results}


@pytest.mark.parametrize("max_batch", [1, 4])
def test_faults_stay_with_their_request(max_batch):
    faults = {2: "raise", 4: "nan", 5: "inf", 7: "diverge"}
    reqs = [_req(i, mesh="a" if i % 2 == 0 else "b") for i in range(10)]
    srv = _HostServer(faults, max_batch=max_batch)
    res = _by_rid(srv.run(reqs, realtime=False))
    assert sorted(res) == list(range(10))
    failed_batches = [b for b in srv.batches if 2 in b]
    assert len(failed_batches) == 1
    for rid, r in res.items():
        if rid in failed_batches[0]:
            # the whole batch of the failed launch, and only it
            assert r.status == "error" and "injected launch failure" in r.error and r.p is None
        elif rid in (4, 5):
            assert r.status == "error" and "non-finite" in r.error and r.p is None
        elif rid == 7:
            assert r.status == "error" and "GMRES" in r.error and r.p is None
        else:
            assert r.status == "ok" and r.error == ""
            assert r.n == srv.meshes[r.mesh].n and np.all(r.p == r.k)


def test_invalid_requests_are_rejected():
    reqs = [_req(0), _req(1, mesh="nowhere"), _req(2, k=-1.0), _req(3, k=float("nan")),
            _req(4, direction=(1.0, 1.0, 0.0)), _req(5, direction=(float("inf"), 0.0, 0.0)),
            _req(6, deadline=-1.0), _req(7)]
    srv = _HostServer({})
    res = _by_rid(srv.run(reqs, realtime=False))
    assert sorted(res) == list(range(8))
    assert [rid for rid, r in res.items() if r.status == "rejected"] == [1, 2, 3, 4, 5, 6]
    assert res[0].status == "ok" and res[7].status == "ok"
    assert all(rid not in b for b in srv.batches for rid in (1, 2, 3, 4, 5, 6))


def test_deadline_expiry():
    """One slot and room for one request at a time: each launch holds the
    slot for 20 ms, so a request queued behind three others with a 1 ms
    deadline expires, one with a 10 s deadline is served."""
    reqs = [_req(0), _req(1), _req(2), _req(3, deadline=1e-3), _req(4, deadline=10.0)]
    srv = _HostServer({}, slots=1, budget_requests=1, delay=2e-2)
    res = _by_rid(srv.run(reqs, realtime=False))
    assert res[3].status == "expired" and res[3].p is None and "not admitted" in res[3].error
    assert all(res[i].status == "ok" for i in (0, 1, 2, 4))
    assert all(3 not in b for b in srv.batches)


def test_every_launch_failing_still_terminates():
    reqs = [_req(i, mesh="ab"[i % 2]) for i in range(6)]
    srv = _HostServer({i: "raise" for i in range(6)}, slots=2, max_batch=2)
    res = _by_rid(srv.run(reqs, realtime=False))
    assert sorted(res) == list(range(6)) and all(r.status == "error" for r in res.values())
