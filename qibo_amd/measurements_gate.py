"""The measurement gate ``M``.

Follows ``src/qibo/gates/measurements.py:11-204``: ``M(*q)`` records which
qubits are measured (``target_qubits``), owns a per-register
``MeasurementResult``, adds basis-rotation gates for non-Z bases, and on a
state vector is a no-op unless ``collapse=True`` (measurements.py:189-204):
then it draws an outcome and collapses the state (circuits with collapsing
measurements run once per shot, ``execute_circuit_repeated``).
"""

from typing import Optional

from . import gates as _g
from .config import raise_error
from .measurements import MeasurementResult


class M(_g.Gate):
    def __init__(
        self,
        *q,
        register_name: Optional[str] = None,
        collapse: bool = False,
        basis=None,
        p0=None,
        p1=None,
    ):
        super().__init__()
        self.name = "measure"
        self.draw_label = "M"
        self.target_qubits = tuple(q)
        self.register_name = register_name
        self.collapse = collapse
        self.result = MeasurementResult(self.target_qubits)
        if basis is None:
            basis = _g.Z
        to_gate = lambda x: getattr(_g, x) if isinstance(x, str) else x
        if not isinstance(basis, list):
            self.basis_gates = len(q) * [to_gate(basis)]
        elif len(basis) != len(self.target_qubits):
            raise_error(
                ValueError,
                f"Given basis list has length {len(basis)} while we are measuring "
                f"{len(self.target_qubits)} qubits.",
            )
        else:
            self.basis_gates = [to_gate(b) for b in basis]
        self.init_args = q
        self.init_kwargs = {
            "register_name": register_name,
            "collapse": collapse,
            "basis": [b.__name__ for b in self.basis_gates],
            "p0": p0,
            "p1": p1,
        }
        if collapse and (p0 is not None or p1 is not None):
            raise_error(
                NotImplementedError, "Bitflip measurement noise is not available when collapsing."
            )
        # bitflip probabilities per qubit (measurements.py:92-97): p0 flips a
        # measured 0, p1 a measured 1; one given stands for both
        if p1 is None:
            p1 = p0
        if p0 is None:
            p0 = p1
        self.bitflip_map = (self._get_bitflip_map(p0), self._get_bitflip_map(p1))
        # basis change gates added by the circuit before this measurement
        self.basis = []
        for qubit, cls in zip(self.target_qubits, self.basis_gates):
            rot = _basis_rotation(cls, qubit)
            if rot is not None:
                self.basis.append(rot)

    @staticmethod
    def _get_bitflip_tuple(qubits, probs):
        """Per-qubit probabilities from a float, a list/tuple (one per qubit)
        or a dict {qubit: p} (measurements.py:121-147)."""
        if isinstance(probs, float):
            if probs < 0 or probs > 1:
                raise_error(ValueError, f"Invalid bitflip probability {probs}.")
            return len(qubits) * (probs,)
        if isinstance(probs, (tuple, list)):
            if len(probs) != len(qubits):
                raise_error(ValueError, f"{len(qubits)} qubits were measured but the given bitflip "
                                        f"probability list contains {len(probs)} values.")
            return tuple(probs)
        if isinstance(probs, dict):
            diff = set(probs.keys()) - set(qubits)
            if diff:
                raise_error(KeyError, f"Bitflip map contains {diff} qubits that are not measured.")
            return tuple(probs[q] if q in probs else 0.0 for q in qubits)
        raise_error(TypeError, f"Invalid type {probs} of bitflip map.")

    def _get_bitflip_map(self, p=None):
        if p is None:
            return {q: 0 for q in self.qubits}
        return dict(zip(self.qubits, self._get_bitflip_tuple(self.qubits, p)))

    def has_bitflip_noise(self):
        return sum(self.bitflip_map[0].values()) > 0 or sum(self.bitflip_map[1].values()) > 0

    def add(self, gate):
        """Merge another measurement's qubits (used for the global measurement gate)."""
        from .kinds import is_measurement

        if not is_measurement(gate):
            raise_error(TypeError, f"cannot merge {type(gate).__name__} into a measurement gate")
        self.target_qubits += gate.target_qubits
        self.bitflip_map[0].update(gate.bitflip_map[0])
        self.bitflip_map[1].update(gate.bitflip_map[1])

    def controlled_by(self, *q):
        raise_error(NotImplementedError, "Measurement gates cannot be controlled.")

    def matrix(self, backend=None):
        raise_error(NotImplementedError, "Measurement gates do not have a matrix representation.")

    def apply(self, backend, state, nqubits):
        """Non-collapsing: no-op. Collapsing (measurements.py:189-204): draw one
        outcome from the marginal probabilities of the sorted target qubits,
        record it, and collapse the state onto it."""
        self.result.backend = backend
        if not self.collapse:
            return state
        qubits = sorted(self.target_qubits)
        dm = len(state.shape) == 2  # density matrix (measurements.py:205-217)
        probs = backend.calculate_probabilities(state, qubits, nqubits, density_matrix=dm)
        shot = self.result.add_shot(probs, backend=backend)
        return backend.collapse_state(state, qubits, shot, nqubits, density_matrix=dm)


def _basis_rotation(cls, qubit):
    """Gate rotating ``cls``'s eigenbasis onto Z (gates.py ``basis_rotation``)."""
    name = cls.__name__
    if name == "Z":
        return None
    if name == "X":
        return _g.H(qubit)
    if name == "Y":
        # (Y + Z) / sqrt(2), as Y.basis_rotation (gates.py:235-241)
        import math

        import numpy as np

        m = (np.array([[0j, -1j], [1j, 0j]]) + np.array([[1, 0], [0, -1]], dtype=complex)) / math.sqrt(2)
        return _g.Unitary(m, qubit, trainable=False)
    raise_error(NotImplementedError, f"Measurement basis {name} is not supported.")
