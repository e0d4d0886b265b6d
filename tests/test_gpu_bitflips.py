"""Readout bitflips on the GPU against the reference's golden values
(tests/test_measurements_probabilistic.py:110-180 of the reference, numpy
platform, seed 123) and against the oracle on a larger case (needs an MI355X)."""

import numpy as np
import pytest
import torch

from oracle import OracleBackend
from qibo_amd import Circuit, gates
from qibo_amd.result import CircuitResult

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("i,probs", [(0, [0.0, 0.0, 0.0]), (1, [0.1, 0.3, 0.2]), (2, [0.5, 0.5, 0.5])])
def test_post_measurement_bitflips_on_circuit(backend, i, probs):
    backend.set_seed(123)
    c = Circuit(5)
    c.add([gates.X(0), gates.X(2), gates.X(3)])
    c.add(gates.M(0, 1, p0={0: probs[0], 1: probs[1]}))
    c.add(gates.M(3, p0=probs[2]))
    freqs = backend.execute_circuit(c, nshots=30).frequencies(binary=False)
    targets = [{5: 30}, {5: 18, 4: 5, 7: 4, 1: 2, 6: 1}, {4: 8, 2: 6, 5: 5, 1: 3, 3: 3, 6: 2, 7: 2, 0: 1}]
    assert dict(freqs) == targets[i]


def test_bitflips_registers(backend):
    thetas = np.random.default_rng(0).random(4)
    backend.set_seed(123)
    c = Circuit(4)
    c.add(gates.RX(i, theta=float(t)) for i, t in enumerate(thetas))
    c.add(gates.M(0, 1, register_name="a", p0={0: 0.2, 1: 0.4}))
    c.add(gates.M(3, register_name="b", p0=0.3))
    result = backend.execute_circuit(c, nshots=30)
    samples = result.samples(binary=True).cpu().numpy()
    regs = result.samples(binary=True, registers=True)
    assert np.array_equal(np.asarray(regs["a"].cpu() if isinstance(regs["a"], torch.Tensor) else regs["a"]),
                          samples[:, :2])
    assert np.array_equal(np.asarray(regs["b"].cpu() if isinstance(regs["b"], torch.Tensor) else regs["b"]),
                          samples[:, 2:])


@pytest.mark.parametrize("i,p0,p1", [(0, 0.2, None), (1, 0.2, 0.1), (2, (0.1, 0.0, 0.2), None),
                                     (3, {0: 0.2, 1: 0.1, 2: 0.0}, None)])
def test_result_apply_bitflips(backend, i, p0, p1):
    c = Circuit(3)
    c.add(gates.M(*range(3)))
    state = backend.zero_state(3)
    result = CircuitResult(state, c.measurements, backend)
    result._samples = backend.cast(np.zeros((10, 3)), dtype="int64")
    backend.set_seed(123)
    noisy = result.apply_bitflips(p0, p1)
    targets = [[0, 0, 0, 0, 2, 3, 0, 0, 0, 0], [0, 0, 0, 0, 2, 3, 0, 0, 0, 0],
               [0, 0, 0, 0, 0, 1, 0, 0, 0, 0], [0, 0, 0, 0, 2, 0, 0, 0, 0, 0]]
    assert backend.samples_to_decimal(noisy, 3).cpu().numpy().tolist() == targets[i]


def test_bitflips_match_oracle_large(backend):
    rng = np.random.default_rng(4)
    nshots, nbits = 20000, 7
    samples = rng.integers(0, 2, (nshots, nbits)).astype(np.int64)
    probs = rng.random((2, nbits)) * 0.5
    np.random.seed(99)
    want = OracleBackend().apply_bitflips(samples, probs)
    np.random.seed(99)
    got = backend.apply_bitflips(backend.cast(samples, dtype="int64"), probs).cpu().numpy()
    assert np.array_equal(got, want)
