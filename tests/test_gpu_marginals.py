"""Marginal probabilities in numpy's summation order (needs an MI355X).

calculate_probabilities of a subset of qubits sums |psi|^2 over the
unmeasured qubits with np.sum(..., axis=unmeasured) (abstract.py:2739-2741).
The kernels reproduce numpy's order (runs of the trailing unmeasured qubits
summed pairwise in buffers of 8192, then in C order; elementwise in C order
when the last qubit is measured), so the marginals equal the oracle's bit
for bit in complex128 and complex64, for measured sets of every shape and
requested orders, and seeded samples of a partial measurement equal the
oracle's shot for shot. Marginals whose serial chain exceeds 2^24 additions
keep the parallel fixed-order sum (checked to 1e-10 relative: numpy's own
sequential rounding over 2^25 values).
"""

import numpy as np
import pytest
import torch

from oracle import OracleBackend
from qibo_amd import Circuit, gates

pytestmark = pytest.mark.gpu

_CASES = [
    (5, [0]), (5, [4]), (5, [2, 0]), (5, [1, 3]), (5, [4, 0, 2]),
    (12, [0]), (12, [11]), (12, [3, 7]), (12, [6, 2, 9]), (12, [0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10]),
    (14, [0]), (14, [13]), (14, [5, 0]), (14, [2, 13, 7]), (14, [1, 12]),
    (17, [0]), (17, [2]), (17, [16]), (17, [3, 9]), (17, [0, 16]), (17, [8, 1, 4, 12]),
    (20, [0]), (20, [1, 18]), (20, [5]),
]


def _state(n, rng, dtype):
    psi = rng.normal(size=2**n) + 1j * rng.normal(size=2**n)
    return (psi / np.linalg.norm(psi)).astype(dtype)


@pytest.mark.parametrize("n,qubits", _CASES)
def test_marginals_bitwise_complex128(backend, n, qubits, rng):
    psi = _state(n, rng, np.complex128)
    want = OracleBackend().calculate_probabilities(psi, qubits, n)
    got = backend.calculate_probabilities(backend.cast(psi), qubits, n).cpu().numpy()
    assert got.dtype == want.dtype
    assert np.array_equal(got, want), np.max(np.abs(got - want))


@pytest.mark.parametrize("n,qubits", _CASES[::2])
def test_marginals_bitwise_complex64(backend64, n, qubits, rng):
    psi = _state(n, rng, np.complex64)
    want = OracleBackend("complex64").calculate_probabilities(psi, qubits, n)
    got = backend64.calculate_probabilities(backend64.cast(psi), qubits, n).cpu().numpy()
    assert got.dtype == np.float32
    assert np.array_equal(got, want), np.max(np.abs(got - want))


def test_long_serial_marginal_keeps_parallel_sum(backend, rng):
    n = 26
    psi = _state(n, rng, np.complex128)
    want = OracleBackend().calculate_probabilities(psi, [25], n)
    got = backend.calculate_probabilities(backend.cast(psi), [25], n).cpu().numpy()
    # numpy adds 2^25 values one by one: its own rounding error is ~sqrt(N) eps
    assert np.max(np.abs(got - want) / want) < 1e-10


def test_partial_measurement_samples_match_oracle(backend, rng):
    n = 14
    c = Circuit(n)
    for q in range(n):
        c.add(gates.RY(q, float(rng.uniform(0, 6))))
    for q in range(0, n - 1, 2):
        c.add(gates.CZ(q, q + 1))
    c.add(gates.M(3, 0, 9))
    ob = OracleBackend()
    psi = ob.execute(c)
    probs = ob.calculate_probabilities(psi, [3, 0, 9], n)
    np.random.seed(77)
    want = ob.sample_shots(probs, 500)
    np.random.seed(77)
    got = backend.execute_circuit(c, nshots=500).samples(binary=False)
    got = got.cpu().numpy() if isinstance(got, torch.Tensor) else np.asarray(got)
    assert np.array_equal(got.ravel(), want)
