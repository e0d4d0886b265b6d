"""Continuous-batching circuit server on one MI355X.

Requests (a circuit, a precision, and either an expectation of sum_i Z_i or a
number of shots) arrive over time. The server admits a waiting request as
soon as an execution slot and enough HBM for its state vector are free, runs
it on a HIP stream of its own so that its kernels overlap with those of the
requests already on the GPU, and completes it when the stream's event fires.
States stay resident in HBM from allocation to completion; nothing goes back
to the host but the final expectation value or shot indices.

* Admission by footprint: a request needs 2^n x itemsize bytes for its state
  plus the workspace of its measurement; the server tracks the bytes of all
  admitted requests against ``hbm_budget`` and admits first-fit from the
  queue (a large request at the head does not block small ones behind it for
  more than ``max_bypass`` admissions).
* Determinism: each request carries its own seed; its uniforms come from
  ``np.random.RandomState(seed)`` and all reductions have a fixed order, so a
  request's result is bitwise the same alone and in any mix.
* Latency: arrival -> completion, reported as p50 / p99 with throughput.

The executor interface is small (``start(request, stream) -> handle``,
``done(handle) -> bool``, ``result(handle)``) so that the admission logic can
be exercised on the CPU with a fake executor.
"""

import collections
import math
import time
from dataclasses import dataclass, field
from typing import Callable, List, Optional

import numpy as np


@dataclass
class Request:
    rid: int
    circuit: object
    dtype: str                 # "complex64" | "complex128"
    kind: str                  # "expval" | "shots"
    nshots: int = 0
    seed: int = 0
    arrival: float = 0.0       # seconds on the server clock
    # filled by the server
    admitted: float = math.nan
    completed: float = math.nan
    result: object = None

    @property
    def nqubits(self) -> int:
        return self.circuit.nqubits

    def footprint(self) -> int:
        itemsize = 16 if self.dtype == "complex128" else 8
        n = self.nqubits
        state = (2**n) * itemsize
        if self.kind == "shots":
            probs = (2**n) * (8 if itemsize == 16 else 4)
            cdf = (2**n) * 8
            return state + probs + cdf + 64 * (2 ** max(n - 12, 0)) + self.nshots * 16
        return state + 8 * 1024 * n


@dataclass
class ServeStats:
    latencies: List[float] = field(default_factory=list)
    wall: float = 0.0
    completed: int = 0
    peak_bytes: int = 0
    peak_inflight: int = 0

    def summary(self) -> dict:
        lat = np.asarray(self.latencies) if self.latencies else np.zeros(1)
        return {
            "circuits": self.completed,
            "wall_s": self.wall,
            "circuits_per_s": self.completed / self.wall if self.wall > 0 else 0.0,
            "p50_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "peak_hbm_bytes": self.peak_bytes,
            "peak_inflight": self.peak_inflight,
        }


class CircuitServer:
    def __init__(self, executor, hbm_budget: int, max_inflight: int = 8, max_bypass: int = 4,
                 clock: Callable[[], float] = time.perf_counter):
        self.executor = executor
        self.hbm_budget = int(hbm_budget)
        self.max_inflight = max_inflight
        self.max_bypass = max_bypass
        self.clock = clock

    def run(self, requests: List[Request], realtime: bool = True) -> ServeStats:
        """Serve ``requests`` (sorted by arrival); returns latency statistics.

        With ``realtime`` the arrival times are honoured against the clock
        (Poisson stream); without it every request is available at t = 0.
        """
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        queue: List[Request] = []
        inflight = []   # (request, handle, bytes)
        used = 0
        bypassed = 0
        stats = ServeStats()
        for r in requests:
            if r.footprint() > self.hbm_budget:
                raise MemoryError(f"request {r.rid} needs {r.footprint()} bytes, budget {self.hbm_budget}")
        t0 = self.clock()
        while pending or queue or inflight:
            now = self.clock() - t0
            while pending and (not realtime or pending[0].arrival <= now):
                queue.append(pending.popleft())
            # completions
            still = []
            for req, h, nbytes in inflight:
                if self.executor.done(h):
                    req.result = self.executor.result(h)
                    req.completed = self.clock() - t0
                    stats.latencies.append(req.completed - (req.arrival if realtime else 0.0))
                    stats.completed += 1
                    used -= nbytes
                else:
                    still.append((req, h, nbytes))
            inflight = still
            # admissions: first fit, bounded bypass of the head of the queue
            admitted_any = True
            while admitted_any and queue and len(inflight) < self.max_inflight:
                admitted_any = False
                for i, req in enumerate(queue):
                    if i > 0 and bypassed >= self.max_bypass:
                        break
                    nbytes = req.footprint()
                    if used + nbytes <= self.hbm_budget:
                        queue.pop(i)
                        bypassed = bypassed + 1 if i > 0 else 0
                        req.admitted = self.clock() - t0
                        h = self.executor.start(req)
                        inflight.append((req, h, nbytes))
                        used += nbytes
                        stats.peak_bytes = max(stats.peak_bytes, used)
                        stats.peak_inflight = max(stats.peak_inflight, len(inflight))
                        admitted_any = True
                        break
            if not admitted_any or not queue:
                self.executor.idle()
        stats.wall = self.clock() - t0
        return stats


class GPUExecutor:
    """Runs each admitted request on its own HIP stream of the MI355X backend."""

    def __init__(self, backends: dict, nstreams: int = 8):
        import torch

        self.torch = torch
        self.backends = backends                      # dtype -> MI355XBackend
        dev = next(iter(backends.values())).torch_device
        self.streams = [torch.cuda.Stream(device=dev) for _ in range(nstreams)]
        self.free_streams = list(range(nstreams))

    def start(self, req: Request):
        torch = self.torch
        backend = self.backends[req.dtype]
        sid = self.free_streams.pop()
        stream = self.streams[sid]
        n = req.nqubits
        with torch.cuda.stream(stream):
            state = backend.zero_state(n)
            state = backend.run_queue([g for g in req.circuit.queue], state, n)
            if req.kind == "expval":
                xm = np.zeros(n, dtype=np.uint64)
                zm = np.asarray([1 << (n - 1 - q) for q in range(n)], dtype=np.uint64)
                ny = np.zeros(n, dtype=np.int32)
                out = backend.expectation_pauli(state, n, xm, zm, ny, np.ones(n))
            else:
                probs = backend.calculate_probabilities(state, list(range(n)), n)
                cdf = backend.exact_cdf(probs)
                u = np.random.RandomState(req.seed).random_sample(req.nshots)
                out = backend.search_cdf(cdf, u)
            host = torch.empty(out.shape, dtype=out.dtype, pin_memory=True)
            host.copy_(out, non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(stream)
        return {"sid": sid, "event": ev, "host": host, "keep": (state, out)}

    def done(self, h) -> bool:
        if h["event"].query():
            if h["sid"] is not None:
                self.free_streams.append(h["sid"])
                h["sid"] = None
                h["keep"] = None
            return True
        return False

    def result(self, h):
        host = h["host"]
        return float(host.item()) if host.dim() == 0 else host.numpy().copy()

    def idle(self):
        time.sleep(50e-6)


def poisson_stream(nrequests: int, qps: float, seed: int, nmin: int = 10, nmax: int = 30,
                   dmin: int = 10, dmax: int = 50) -> List[Request]:
    """The mixed stream of the serving benchmark: n ~ U[nmin, nmax], depth ~
    U[dmin, dmax] layers of the seeded layered generator, half complex64 and
    half complex128, half <sum Z_i> and half nshots ~ U[1000, 10000]."""
    from .models import random_layered_circuit

    rng = np.random.default_rng(seed)
    t = 0.0
    reqs = []
    for i in range(nrequests):
        t += rng.exponential(1.0 / qps)
        n = int(rng.integers(nmin, nmax + 1))
        depth = int(rng.integers(dmin, dmax + 1))
        dtype = "complex64" if rng.random() < 0.5 else "complex128"
        kind = "expval" if rng.random() < 0.5 else "shots"
        nshots = int(rng.integers(1000, 10001)) if kind == "shots" else 0
        circ = random_layered_circuit(n, depth, seed=int(rng.integers(1 << 30)), measure=False)
        reqs.append(Request(i, circ, dtype, kind, nshots, seed=int(rng.integers(1 << 30)), arrival=t))
    return reqs
