"""Admission and completion logic of the circuit server, on the CPU with a fake
executor (no GPU): requests are admitted as memory and slots free up, never
exceed the HBM budget or the slot count, large requests at the head do not
starve small ones, an over-budget request is rejected up front, and every
request completes exactly once."""

import math

import pytest

from qibo_amd.serving import CircuitServer, Request, poisson_stream


class FakeCircuit:
    def __init__(self, n):
        self.nqubits = n
        self.queue = []


class FakeExecutor:
    """Each request takes `2^(n-10)` polls to finish."""

    def __init__(self, server_budget, max_inflight):
        self.live = {}
        self.used = 0
        self.budget = server_budget
        self.max_inflight = max_inflight
        self.order = []
        self.ticks = 0

    def start(self, req):
        assert len(self.live) < self.max_inflight
        self.used += req.footprint()
        assert self.used <= self.budget
        self.live[req.rid] = [req, 2 ** max(req.nqubits - 10, 0)]
        self.order.append(req.rid)
        return req.rid

    def done(self, h):
        entry = self.live[h]
        entry[1] -= 1
        if entry[1] <= 0:
            self.used -= entry[0].footprint()
            del self.live[h]
            return True
        return False

    def result(self, h):
        return h

    def idle(self):
        self.ticks += 1


def _requests(sizes):
    return [Request(i, FakeCircuit(n), "complex128", "expval", arrival=0.0) for i, n in enumerate(sizes)]


def test_budget_and_slots_respected():
    reqs = _requests([20, 22, 18, 24, 20, 16, 23, 21, 19, 24, 12, 15])
    budget = reqs[3].footprint() + reqs[6].footprint()
    ex = FakeExecutor(budget, 3)
    stats = CircuitServer(ex, budget, max_inflight=3).run(reqs, realtime=False)
    assert stats.completed == len(reqs)
    assert sorted(r.result for r in reqs) == list(range(len(reqs)))
    assert stats.peak_bytes <= budget and stats.peak_inflight <= 3
    assert all(not math.isnan(r.completed) and r.completed >= r.admitted for r in reqs)


def test_small_requests_bypass_a_large_head():
    reqs = _requests([26, 26, 10, 10, 10])
    budget = reqs[0].footprint() + reqs[2].footprint() * 3
    ex = FakeExecutor(budget, 8)
    CircuitServer(ex, budget, max_inflight=8, max_bypass=4).run(reqs, realtime=False)
    # the second large request cannot fit next to the first; the small ones go ahead
    assert ex.order.index(2) < ex.order.index(1)


def test_over_budget_request_rejected():
    reqs = _requests([30])
    with pytest.raises(MemoryError):
        CircuitServer(FakeExecutor(1 << 20, 2), 1 << 20).run(reqs, realtime=False)


def test_poisson_stream_shape():
    reqs = poisson_stream(40, qps=50.0, seed=3, nmin=10, nmax=14, dmin=2, dmax=3)
    assert len(reqs) == 40
    assert all(10 <= r.nqubits <= 14 for r in reqs)
    assert all(b.arrival > a.arrival for a, b in zip(reqs, reqs[1:]))
    kinds = {r.kind for r in reqs}
    dtypes = {r.dtype for r in reqs}
    assert kinds == {"expval", "shots"} and dtypes == {"complex64", "complex128"}
    assert all(1000 <= r.nshots <= 10000 for r in reqs if r.kind == "shots")
