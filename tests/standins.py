"""Stand-ins for the objects qibo 0.3.5 hands to a backend after
``qibo.set_backend("qibo_amd")``: gates, measurement gates, fused groups and a
circuit with the attributes qibo's own classes carry. None of them derives
from a ``qibo_amd`` class, so a backend that dispatched on its own class
identity would mis-handle them (a measurement would reach ``matrix()``, which
raises in qibo).

Attributes mirrored (reference file:line):
* ``Gate``: src/qibo/gates/abstract.py:61-96 (``_target_qubits``,
  ``_control_qubits``, ``_parameters``, ``symbolic_parameters``,
  ``is_controlled_by``, ``init_args``/``init_kwargs``), :440-442 (``qubits`` =
  controls + targets), :330-345 (``matrix`` asks the backend);
* ``ParametrizedGate``: :680-695 (``matrix`` -> ``matrix_parametrized``);
* ``M``: src/qibo/gates/measurements.py:46-60, 182-197;
* ``MeasurementResult``: src/qibo/measurements.py:161-222;
* ``FusedGate``: src/qibo/gates/special.py:28-47, 93-104;
* ``Circuit``: src/qibo/models/circuit.py (``nqubits``, ``queue``,
  ``measurements``, ``density_matrix``, ``repeated_execution``,
  ``has_collapse``, ``_final_state``).
"""

import collections

import numpy as np


class Gate:
    def __init__(self):
        self.name = None
        self.is_controlled_by = False
        self.init_args = []
        self.init_kwargs = {}
        self._target_qubits = ()
        self._control_qubits = ()
        self._parameters = ()
        self.symbolic_parameters = {}
        self.trainable = True

    @property
    def target_qubits(self):
        return self._target_qubits

    @target_qubits.setter
    def target_qubits(self, q):
        self._target_qubits = tuple(q)

    @property
    def control_qubits(self):
        return tuple(sorted(self._control_qubits))

    @control_qubits.setter
    def control_qubits(self, q):
        self._control_qubits = tuple(q)

    @property
    def qubits(self):
        return self.control_qubits + self.target_qubits

    @property
    def parameters(self):
        return self._parameters

    def controlled_by(self, *q):
        self.is_controlled_by = True
        self.control_qubits = q
        return self

    def matrix(self, backend=None):
        return backend.matrix(self)

    def apply(self, backend, state, nqubits):
        return backend.apply_gate(self, state, nqubits)


class ParametrizedGate(Gate):
    def __init__(self, trainable=True):
        super().__init__()
        self.trainable = trainable
        self.nparams = 1

    @Gate.parameters.setter
    def parameters(self, x):
        self._parameters = tuple(x)

    def matrix(self, backend=None):
        return backend.matrix_parametrized(self)


class H(Gate):
    def __init__(self, q):
        super().__init__()
        self.name = "h"
        self.target_qubits = (q,)
        self.init_args = [q]


class CZ(Gate):
    def __init__(self, q0, q1):
        super().__init__()
        self.name = "cz"
        self.control_qubits = (q0,)
        self.target_qubits = (q1,)
        self.init_args = [q0, q1]


class CNOT(Gate):
    def __init__(self, q0, q1):
        super().__init__()
        self.name = "cx"
        self.control_qubits = (q0,)
        self.target_qubits = (q1,)
        self.init_args = [q0, q1]


class _Rn(ParametrizedGate):
    def __init__(self, q, theta, trainable=True):
        super().__init__(trainable)
        self.target_qubits = (q,)
        self.init_args = [q]
        self.init_kwargs = {"theta": theta, "trainable": trainable}
        self._parameters = (theta,)


class RX(_Rn):
    pass


class RY(_Rn):
    pass


class RZ(_Rn):
    pass


class MeasurementResult:
    def __init__(self, target_qubits):
        self.target_qubits = tuple(target_qubits)
        self.backend = None
        self._samples = None
        self._frequencies = None

    def has_samples(self):
        return self._samples is not None

    def register_samples(self, samples):
        self._samples = samples

    def register_frequencies(self, frequencies):
        self._frequencies = frequencies

    def reset(self):
        self._samples = None
        self._frequencies = None

    def samples(self, binary=True, registers=False, backend=None):
        if binary:
            return self._samples
        s = np.asarray(self._samples.cpu() if hasattr(self._samples, "cpu") else self._samples)
        return s @ (2 ** np.arange(s.shape[1] - 1, -1, -1))

    def frequencies(self, binary=True, registers=False, backend=None):
        s = self.samples(binary=False)
        return collections.Counter(int(x) for x in s)


class M(Gate):
    def __init__(self, *q, register_name=None, collapse=False):
        super().__init__()
        self.name = "measure"
        self.target_qubits = tuple(q)
        self.register_name = register_name
        self.collapse = collapse
        self.result = MeasurementResult(self.target_qubits)
        self.init_args = list(q)
        self.init_kwargs = {"register_name": register_name, "collapse": collapse}
        self.bitflip_map = ({q_: 0.0 for q_ in q}, {q_: 0.0 for q_ in q})

    def matrix(self, backend=None):
        raise NotImplementedError("Measurement gates do not have a matrix representation.")

    def apply(self, backend, state, nqubits):
        self.result.backend = backend
        if not self.collapse:
            return state
        raise NotImplementedError("collapse is not mirrored by these stand-ins")


class FusedGate(Gate):
    def __init__(self, *q):
        super().__init__()
        self.name = "Fused Gate"
        self.target_qubits = tuple(sorted(q))
        self.init_args = list(q)
        self.qubit_set = set(q)
        self.gates = []
        self.marked = False

    def append(self, gate):
        self.qubit_set = self.qubit_set | set(gate.qubits)
        self.init_args = sorted(self.qubit_set)
        self.target_qubits = tuple(self.init_args)
        self.gates.append(gate)

    def matrix(self, backend=None):
        return backend.matrix_fused(self)


class Circuit:
    def __init__(self, nqubits):
        self.nqubits = nqubits
        self.queue = []
        self.measurements = []
        self.density_matrix = False
        self.repeated_execution = False
        self.has_collapse = False
        self._final_state = None

    def add(self, gate):
        if isinstance(gate, M):
            self.measurements.append(gate)
        self.queue.append(gate)


def layered(nqubits, nlayers, seed, measure=True):
    """The RY/RZ/CZ-brick layered circuit built from the stand-ins, gate for
    gate and angle for angle as ``qibo_amd.models.random_layered_circuit``
    (RY on every qubit, RZ on every qubit, then CZ on the pairs starting at
    layer mod 2)."""
    rng = np.random.default_rng(seed)
    angles = rng.uniform(0.0, 2 * np.pi, size=(nlayers, 2, nqubits))
    c = Circuit(nqubits)
    for layer in range(nlayers):
        for q in range(nqubits):
            c.add(RY(q, float(angles[layer, 0, q])))
        for q in range(nqubits):
            c.add(RZ(q, float(angles[layer, 1, q])))
        for q in range(layer % 2, nqubits - 1, 2):
            c.add(CZ(q, q + 1))
    if measure:
        c.add(M(*range(nqubits)))
    return c
