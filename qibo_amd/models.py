"""Circuit generators of the benchmark workloads.

* ``QFT`` -- qibo's QFT circuit (``src/qibo/models/qft.py:47-58``): H on qubit
  i1 followed by CU1(i2, i1, pi / 2^(i2-i1)) for i2 > i1, then the SWAPs that
  restore the qubit order. n^2/2 + n gates.
* ``random_layered_circuit`` -- seeded layered circuit: per layer RY(theta) on
  every qubit, RZ(phi) on every qubit, then CZ on the brick pairs
  (l mod 2 + 2j, l mod 2 + 2j + 1); angles U(0, 2 pi) from
  ``np.random.default_rng(seed)``, drawn as one (nlayers, 2, nqubits) array.
* ``hardware_efficient_ansatz`` -- per layer RY(theta) and RZ(phi) on each
  qubit followed by a CNOT ladder CNOT(i, i+1), i = 0..n-2.
"""

import math

import numpy as np

from . import gates
from .circuit import Circuit


def QFT(nqubits: int, with_swaps: bool = True, **kwargs) -> Circuit:
    circuit = Circuit(nqubits, **kwargs)
    for i1 in range(nqubits):
        circuit.add(gates.H(i1))
        for i2 in range(i1 + 1, nqubits):
            circuit.add(gates.CU1(i2, i1, math.pi / 2 ** (i2 - i1)))
    if with_swaps:
        for q in range(nqubits // 2):
            circuit.add(gates.SWAP(q, nqubits - q - 1))
    return circuit


def layered_angles(nqubits: int, nlayers: int, seed: int) -> np.ndarray:
    """Angles of ``random_layered_circuit``: shape (nlayers, 2, nqubits)."""
    rng = np.random.default_rng(seed)
    return rng.uniform(0.0, 2 * math.pi, size=(nlayers, 2, nqubits))


def random_layered_circuit(nqubits: int, nlayers: int = 20, seed: int = 1234,
                           measure: bool = True) -> Circuit:
    angles = layered_angles(nqubits, nlayers, seed)
    circuit = Circuit(nqubits)
    for layer in range(nlayers):
        for q in range(nqubits):
            circuit.add(gates.RY(q, float(angles[layer, 0, q])))
        for q in range(nqubits):
            circuit.add(gates.RZ(q, float(angles[layer, 1, q])))
        for q in range(layer % 2, nqubits - 1, 2):
            circuit.add(gates.CZ(q, q + 1))
    if measure:
        circuit.add(gates.M(*range(nqubits)))
    return circuit


def hardware_efficient_ansatz(nqubits: int, nlayers: int, params=None) -> Circuit:
    """RY/RZ layers with a CNOT ladder; ``params`` has 2 * nqubits * nlayers values."""
    circuit = Circuit(nqubits)
    for _ in range(nlayers):
        for q in range(nqubits):
            circuit.add(gates.RY(q, 0.0))
            circuit.add(gates.RZ(q, 0.0))
        for q in range(nqubits - 1):
            circuit.add(gates.CNOT(q, q + 1))
    if params is not None:
        circuit.set_parameters(np.asarray(params, dtype=np.float64))
    return circuit
