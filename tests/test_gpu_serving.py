"""Continuous-batching server on the GPU (needs an MI355X): every request's
result is bitwise the same when served alone and when served in a mixed
stream (tolerance: exact equality), and matches the oracle (expectations
1e-12 complex128 / 1e-4 complex64; shot indices bitwise for complex128).
The executor's three paths -- direct launches, batched launches of requests
below 12 qubits, and replays of graphs captured per circuit structure -- give
bitwise the same results as each other; a graph is replayed several times
with other work between the replays, each replay checked."""

import numpy as np
import pytest

from oracle import OracleBackend
from qibo_amd.backend import MI355XBackend
from qibo_amd import models
from qibo_amd.serving import CircuitServer, GPUExecutor, Request, poisson_stream

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def executor(backend, backend64):
    return GPUExecutor({"complex128": backend, "complex64": backend64}, nstreams=6)


def test_same_alone_and_in_mix(executor):
    reqs = poisson_stream(24, qps=2000.0, seed=5, nmin=10, nmax=17, dmin=2, dmax=6)
    server = CircuitServer(executor, hbm_budget=8 << 30, max_inflight=6)
    stats = server.run(reqs, realtime=True)
    assert stats.completed == len(reqs)
    assert stats.peak_inflight > 1  # requests really overlapped
    mixed = [r.result for r in reqs]
    alone = []
    for r in reqs:
        h = executor.start(r)
        while not executor.done(h):
            pass
        alone.append(executor.result(h))
    for r, a, b in zip(reqs, mixed, alone):
        if r.kind == "expval":
            assert a == b, r.rid
        else:
            np.testing.assert_array_equal(a, b)
    ob = OracleBackend()
    for r, a in zip(reqs[:10], mixed[:10]):
        n = r.nqubits
        psi = ob.execute(r.circuit)
        if r.kind == "expval":
            want = float(np.sum((np.abs(psi) ** 2) * _zsum_weights(n)))
            tol = 1e-12 if r.dtype == "complex128" else 1e-4
            assert abs(a - want) < tol
        elif r.dtype == "complex128":
            p = ob.calculate_probabilities(psi, list(range(n)), n)
            cdf = np.cumsum(p)
            cdf /= cdf[-1]
            u = np.random.RandomState(r.seed).random_sample(r.nshots)
            np.testing.assert_array_equal(a, cdf.searchsorted(u, side="right"))


def _zsum_weights(n):
    idx = np.arange(2**n)
    w = np.zeros(2**n)
    for q in range(n):
        w += 1 - 2 * ((idx >> (n - 1 - q)) & 1)
    return w


def _serve_alone(executor, reqs):
    out = []
    for r in reqs:
        h = executor.start(r)
        while not executor.done(h):
            pass
        out.append(executor.result(h))
    return out


def _same(reqs, a, b):
    for r, x, y in zip(reqs, a, b):
        if r.kind == "expval":
            assert x == y, r.rid
        else:
            np.testing.assert_array_equal(x, y)


def _structured_requests(seed):
    """Requests that repeat circuit structures (same n, depth, gates) with new
    angles and seeds, of both dtypes and kinds, 10..20 qubits."""
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(30):
        n = [10, 11, 13, 16, 20][i % 5]
        depth = 3 + (i % 2)
        dtype = "complex128" if (i // 5) % 2 == 0 else "complex64"
        kind = "shots" if (i // 10) % 2 == 0 else "expval"
        circ = models.random_layered_circuit(n, depth, seed=int(rng.integers(1 << 30)), measure=False)
        reqs.append(Request(i, circ, dtype, kind, 2000 if kind == "shots" else 0,
                            seed=int(rng.integers(1 << 30)), arrival=0.001 * i))
    return reqs


def test_executor_paths_agree(backend, backend64):
    backends = {"complex128": backend, "complex64": backend64}
    plain = GPUExecutor(backends, nstreams=4, batch_small=False, graphs=False)
    fancy = GPUExecutor(backends, nstreams=6, batch_small=True, graphs=True)
    reqs = _structured_requests(11)
    want = _serve_alone(plain, reqs)
    assert plain.stats["direct"] > 0
    # served as a stream: small requests batched, repeated structures replayed
    server = CircuitServer(fancy, hbm_budget=8 << 30, max_inflight=6)
    stats = server.run(reqs, realtime=True)
    assert stats.completed == len(reqs)
    _same(reqs, [r.result for r in reqs], want)
    assert fancy.stats["batched"] > 0
    assert fancy.stats["graph_capture"] > 0
    # replay the graphs again, alone, with other work on the device between
    for _ in range(3):
        got = _serve_alone(fancy, reqs)
        _same(reqs, got, want)
        backend.execute_circuit(models.QFT(14))  # eager work between the replays
    assert fancy.stats["graph_replay"] >= 3 * fancy.stats["graph_capture"]
    plain.close()
    fancy.close()


def test_stream_up_to_24_qubits(backend, backend64):
    executor = GPUExecutor({"complex128": backend, "complex64": backend64}, nstreams=6)
    reqs = poisson_stream(16, qps=500.0, seed=21, nmin=10, nmax=24, dmin=4, dmax=12)
    stats = CircuitServer(executor, hbm_budget=16 << 30, max_inflight=6).run(reqs, realtime=True)
    assert stats.completed == len(reqs)
    ob = OracleBackend()
    for r in reqs:
        if r.nqubits > 20:
            continue
        psi = ob.execute(r.circuit)
        n = r.nqubits
        if r.kind == "expval":
            want = float(np.sum((np.abs(psi) ** 2) * _zsum_weights(n)))
            tol = 1e-12 if r.dtype == "complex128" else 1e-4
            assert abs(r.result - want) < tol
        else:
            p = ob.calculate_probabilities(psi, list(range(n)), n)
            cdf = np.cumsum(p)
            cdf /= cdf[-1]
            u = np.random.RandomState(r.seed).random_sample(r.nshots)
            want = cdf.searchsorted(u, side="right")
            if r.dtype == "complex128":
                np.testing.assert_array_equal(r.result, want)
            else:
                # complex64 probabilities move a few shots across bin edges
                assert np.mean(r.result != want) < 0.01
    executor.close()
