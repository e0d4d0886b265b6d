"""Readout bitflips in the oracle, pinned by the reference's golden values
(tests/test_measurements_probabilistic.py:110-180 of the reference, numpy
platform, seed 123): the flips of all-zero samples for four probability
forms, and the frequencies of the noisy samples of a 5-qubit circuit whose
measurements carry bitflip maps."""

from collections import Counter

import numpy as np
import pytest

from oracle import OracleBackend
from qibo_amd import Circuit, gates


@pytest.mark.parametrize("i,p0,p1", [(0, 0.2, None), (1, 0.2, 0.1), (2, (0.1, 0.0, 0.2), None),
                                     (3, {0: 0.2, 1: 0.1, 2: 0.0}, None)])
def test_apply_bitflips_golden(i, p0, p1):
    targets = [[0, 0, 0, 0, 2, 3, 0, 0, 0, 0], [0, 0, 0, 0, 2, 3, 0, 0, 0, 0],
               [0, 0, 0, 0, 0, 1, 0, 0, 0, 0], [0, 0, 0, 0, 2, 0, 0, 0, 0, 0]]
    m = gates.M(*range(3))
    t0 = m._get_bitflip_tuple(m.qubits, p0)
    t1 = t0 if p1 is None else m._get_bitflip_tuple(m.qubits, p1)
    ob = OracleBackend()
    np.random.seed(123)
    noisy = ob.apply_bitflips(np.zeros((10, 3), dtype=np.int32), np.array([t0, t1], dtype=np.float64))
    assert ob.samples_to_decimal(noisy, 3).tolist() == targets[i]


def bitflip_circuit_frequencies(probs, nshots=30, seed=123):
    """The reference's flow for a circuit with noisy measurements: sample the
    measured qubits' probabilities, convert to bits, flip, count."""
    ob = OracleBackend()
    np.random.seed(seed)
    c = Circuit(5)
    c.add([gates.X(0), gates.X(2), gates.X(3)])
    psi = ob.execute(c)
    p = ob.calculate_probabilities(psi, [0, 1, 3], 5)
    if not any(probs):
        return Counter(ob.sample_frequencies(p, nshots))
    samples = ob.samples_to_binary(ob.sample_shots(p, nshots), 3)
    flips = np.array([[probs[0], probs[1], probs[2]]] * 2, dtype=np.float64)
    noisy = ob.apply_bitflips(samples, flips)
    return Counter(ob.samples_to_decimal(noisy, 3).tolist())


@pytest.mark.parametrize("i,probs", [(0, [0.0, 0.0, 0.0]), (1, [0.1, 0.3, 0.2]), (2, [0.5, 0.5, 0.5])])
def test_bitflip_circuit_golden(i, probs):
    targets = [{5: 30}, {5: 18, 4: 5, 7: 4, 1: 2, 6: 1}, {4: 8, 2: 6, 5: 5, 1: 3, 3: 3, 6: 2, 7: 2, 0: 1}]
    assert dict(bitflip_circuit_frequencies(probs)) == targets[i]


def test_bitflip_maps():
    m = gates.M(0, 1, p0={0: 0.1, 1: 0.3})
    assert m.bitflip_map == ({0: 0.1, 1: 0.3}, {0: 0.1, 1: 0.3}) and m.has_bitflip_noise()
    m = gates.M(3, p0=0.2, p1=0.4)
    assert m.bitflip_map == ({3: 0.2}, {3: 0.4})
    assert not gates.M(1, 2).has_bitflip_noise()
    with pytest.raises(KeyError):
        gates.M(0, p0={1: 0.2})
    with pytest.raises(ValueError):
        gates.M(0, 1, p0=[0.1])
    with pytest.raises(NotImplementedError):
        gates.M(0, collapse=True, p0=0.1)
