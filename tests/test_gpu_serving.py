"""Continuous-batching server on the GPU (needs an MI355X): every request's
result is bitwise the same when served alone and when served in a mixed
stream (tolerance: exact equality), and matches the oracle (expectations
1e-12 complex128 / 1e-4 complex64; shot indices bitwise for complex128)."""

import numpy as np
import pytest

from oracle import OracleBackend
from qibo_amd.backend import MI355XBackend
from qibo_amd.serving import CircuitServer, GPUExecutor, poisson_stream

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
