"""``Circuit``: the gate queue the backend executes.

Follows ``src/qibo/models/circuit.py`` for the parts on the state-vector path:
the queue and its fusion view (``_Queue.to_fused/from_fused``, :43-95), ``add``
(:608-703, including measurement registers and basis gates), parameter updates
(``set_parameters``, :788-857), ``fuse`` (:954-1003), ``unitary``
(:1005-1023), ``execute``/``__call__`` (:1071-1110) and circuit addition.
QASM, drawing, noise and distributed execution are not part of this package.
"""

import copy
from collections import Counter
from collections.abc import Iterable
from typing import Dict, List, Union

import numpy as np

from . import gates
from .config import raise_error


class _ParametrizedGates(list):
    def __init__(self, *args, **kwargs):
        super().__init__(*args, **kwargs)
        self.set = set()
        self.nparams = 0

    def append(self, gate):
        super().append(gate)
        self.set.add(gate)
        self.nparams += gate.nparams


class _Queue(list):
    def __init__(self, nqubits: int):
        super().__init__()
        self.nqubits = nqubits

    def to_fused(self):
        """Wrap every gate in a ``FusedGate`` and link neighbours per qubit."""
        last_gate = {}
        queue = self.__class__(self.nqubits)
        for gate in self:
            fgate = gates.FusedGate.from_gate(gate)
            if isinstance(gate, gates.SpecialGate):
                fgate.qubit_set = set(range(self.nqubits))
                fgate.init_args = sorted(fgate.qubit_set)
                fgate.target_qubits = tuple(fgate.init_args)
            for q in fgate.qubits:
                if q in last_gate:
                    neighbor = last_gate[q]
                    fgate.left_neighbors[q] = neighbor
                    neighbor.right_neighbors[q] = fgate
                last_gate[q] = fgate
            queue.append(fgate)
        return queue

    def from_fused(self):
        """Drop absorbed groups; unwrap single-gate groups; keep special gates."""
        queue = self.__class__(self.nqubits)
        for gate in self:
            if not gate.marked:
                queue.append(gate.gates[0] if len(gate.gates) == 1 else gate)
            elif isinstance(gate.gates[0], (gates.SpecialGate, gates.M)):
                queue.append(gate.gates[0])
        return queue

    @property
    def nmeasurements(self) -> int:
        return sum(1 for g in self if isinstance(g, gates.M))

    @property
    def moments(self) -> list:
        moments = [self.nqubits * [None]]
        moment_index = self.nqubits * [0]
        for gate in self:
            qubits = (
                gate.qubits
                if not isinstance(gate, gates.CallbackGate)
                else tuple(range(self.nqubits))
            )
            idx = max(moment_index[q] for q in qubits)
            for q in qubits:
                if idx >= len(moments):
                    moments.append(len(moments[-1]) * [None])
                moments[idx][q] = gate
                moment_index[q] = idx + 1
        return moments


class Circuit:
    """Ordered gate list on ``nqubits`` qubits (qubit 0 = most significant bit)."""

    def __init__(self, nqubits: int, accelerators=None, density_matrix: bool = False,
                 wire_names=None):
        if not isinstance(nqubits, int) or nqubits < 1:
            raise_error(ValueError, f"Number of qubits must be a positive integer, got {nqubits}.")
        if accelerators:
            raise_error(
                NotImplementedError,
                "Distributed circuits are not supported: the MI355X backend runs one GPU.",
            )
        self.nqubits = nqubits
        self.init_kwargs = {
            "nqubits": nqubits,
            "accelerators": None,
            "density_matrix": density_matrix,
            "wire_names": wire_names,
        }
        self.wire_names = wire_names if wire_names is not None else list(range(nqubits))
        self.queue = _Queue(nqubits)
        self.parametrized_gates = _ParametrizedGates()
        self.trainable_gates = _ParametrizedGates()
        self.measurements = []
        self._final_state = None
        self.has_collapse = False
        self.has_unitary_channel = False
        self.density_matrix = density_matrix
        self.accelerators = None

    @property
    def repeated_execution(self) -> bool:
        """Run once per shot: collapsing measurements, or unitary channels on
        state vectors (circuit.py:320-324)."""
        return self.has_collapse or (self.has_unitary_channel and not self.density_matrix)

    # -- composition -------------------------------------------------------
    def __add__(self, circuit):
        for k, v in self.init_kwargs.items():
            if circuit.init_kwargs[k] != v:
                raise_error(
                    ValueError,
                    f"Cannot add circuits with different kwargs. {k} is {v} for first "
                    f"circuit and {circuit.init_kwargs[k]} for the second.",
                )
        new = self.__class__(**self.init_kwargs)
        for gate in self.queue:
            new.add(gate)
        for gate in circuit.queue:
            new.add(gate)
        return new

    def _shallow_copy(self):
        new = self.__class__(**self.init_kwargs)
        new.parametrized_gates = self.parametrized_gates
        new.trainable_gates = self.trainable_gates
        new.measurements = self.measurements
        new.has_collapse = self.has_collapse
        new.has_unitary_channel = self.has_unitary_channel
        return new

    def copy(self, deep: bool = False):
        if deep:
            new = self.__class__(**self.init_kwargs)
            for gate in self.queue:
                new.add(copy.deepcopy(gate))
            return new
        new = self._shallow_copy()
        new.queue = copy.copy(self.queue)
        return new

    def invert(self):
        """Circuit with the daggers of all gates in reverse order."""
        new = self.__class__(**self.init_kwargs)
        for gate in reversed(self.queue):
            if isinstance(gate, gates.M):
                continue
            new.add(gate.dagger())
        for m in self.measurements:
            new.add(m.__class__(*m.init_args, **m.init_kwargs))
        return new

    def add(self, gate):
        """Append a gate (or an iterable of gates) to the queue."""
        if isinstance(gate, Iterable) and not isinstance(gate, gates.Gate):
            result = None
            for g in gate:
                result = self.add(g)
            return result
        if not isinstance(gate, gates.Gate):
            raise_error(TypeError, f"Unknown gate type {type(gate)}.")
        if self._final_state is not None:
            raise_error(RuntimeError, "Cannot add gates to a circuit after it is executed.")
        for q in gate.target_qubits:
            if q >= self.nqubits:
                raise_error(
                    ValueError,
                    f"Attempting to add gate with target qubits {gate.target_qubits} on a "
                    f"circuit of {self.nqubits} qubits.",
                )
        if isinstance(gate, gates.M):
            for base in gate.basis:
                if base not in self.queue:
                    self.add(base)
            self.queue.append(gate)
            if gate.register_name is None:
                gate.register_name = f"register{self.queue.nmeasurements - 1}"
            else:
                for mgate in self.measurements:
                    if mgate.register_name == gate.register_name:
                        raise_error(KeyError, f"Register {gate.register_name} already exists in circuit.")
            gate.result.circuit = self
            if gate.collapse:
                self.has_collapse = True
            else:
                self.measurements.append(gate)
            return gate.result
        self.queue.append(gate)
        for measurement in list(self.measurements):
            if set(measurement.qubits) & set(gate.qubits):
                measurement.collapse = True
                self.has_collapse = True
                self.measurements.remove(measurement)
        if isinstance(gate, gates.UnitaryChannel):
            self.has_unitary_channel = True
        if isinstance(gate, gates.ParametrizedGate):
            self.parametrized_gates.append(gate)
            if gate.trainable:
                self.trainable_gates.append(gate)
        return None

    # -- inspection --------------------------------------------------------
    @property
    def measurement_tuples(self) -> dict:
        return {m.register_name: m.target_qubits for m in self.measurements}

    @property
    def ngates(self) -> int:
        return len(self.queue)

    @property
    def depth(self) -> int:
        return len(self.queue.moments)

    @property
    def gate_types(self) -> Counter:
        return Counter(g.__class__ for g in self.queue)

    @property
    def gate_names(self) -> Counter:
        return Counter(g.name for g in self.queue)

    def gates_of_type(self, gate):
        if isinstance(gate, str):
            return [(i, g) for i, g in enumerate(self.queue) if g.name == gate]
        if isinstance(gate, type) and issubclass(gate, gates.Gate):
            return [(i, g) for i, g in enumerate(self.queue) if isinstance(g, gate)]
        raise_error(TypeError, f"Gate identifier {gate} not recognized.")

    # -- parameters --------------------------------------------------------
    def _set_parameters_list(self, parameters, n: int) -> None:
        if n == len(self.trainable_gates):
            for i, gate in enumerate(self.trainable_gates):
                gate.parameters = parameters[i]
        elif n == self.trainable_gates.nparams:
            parameters = list(parameters)
            k = 0
            for i, gate in enumerate(self.trainable_gates):
                if gate.nparams == 1:
                    gate.parameters = parameters[i + k]
                else:
                    gate.parameters = parameters[i + k: i + k + gate.nparams]
                k += gate.nparams - 1
        else:
            raise_error(
                ValueError,
                f"Given list of parameters has length {n} while the circuit contains "
                f"{len(self.trainable_gates)} parametrized gates.",
            )

    def set_parameters(self, parameters: Union[List, Dict, np.ndarray]) -> None:
        self._final_state = None
        if isinstance(parameters, dict):
            diff = set(parameters.keys()) - self.trainable_gates.set
            if diff:
                raise_error(
                    KeyError,
                    f"Dictionary contains gates {diff} which are not on the list of "
                    "parametrized gates of the circuit.",
                )
            for gate, params in parameters.items():
                gate.parameters = params
        elif isinstance(parameters, Iterable) and not isinstance(parameters, (set, str)):
            try:
                n = int(parameters.shape[0])
            except AttributeError:
                n = len(parameters)
            self._set_parameters_list(parameters, n)
        else:
            raise_error(TypeError, f"Invalid type of parameters {type(parameters)}.")

    def get_parameters(self, output_format: str = "list", include_not_trainable: bool = False):
        glist = self.parametrized_gates if include_not_trainable else self.trainable_gates
        if output_format == "list":
            return [g.parameters for g in glist]
        if output_format == "dict":
            return {g: g.parameters for g in glist}
        if output_format == "flatlist":
            out = []
            for g in glist:
                out.extend(g.parameters)
            return out
        raise_error(ValueError, f"Unknown format {output_format} given in ``get_parameters``.")

    # -- transformation ----------------------------------------------------
    def fuse(self, max_qubits: int = 2):
        """Greedy neighbour fusion into ``FusedGate`` groups of <= ``max_qubits``
        (circuit.py:954-1003): the grouping is planned in the library
        (qsim_fuse_plan, host code); ``_fuse_python`` is the same planner in
        Python, kept as its restatement."""
        from . import _lib
        from .kinds import is_measurement, is_special

        lib = _lib.load()
        n = self.nqubits
        gates_in = list(self.queue)
        ng = len(gates_in)
        masks = np.zeros(max(ng, 1), dtype=np.uint64)
        kind = np.zeros(max(ng, 1), dtype=np.uint8)
        for i, g in enumerate(gates_in):
            m = 0
            for q in g.qubits:
                m |= 1 << int(q)
            masks[i] = m
            kind[i] = 1 if is_measurement(g) else (2 if (is_special(g) or isinstance(g, gates.SpecialGate)) else 0)
        status = np.zeros(max(ng, 1), dtype=np.int32)
        members = np.zeros(max(ng, 1), dtype=np.int32)
        offsets = np.zeros(ng + 1, dtype=np.int32)
        _lib.check(lib.qsim_fuse_plan(n, ng, masks.ctypes.data, kind.ctypes.data, int(max_qubits),
                                      status.ctypes.data, members.ctypes.data, offsets.ctypes.data), lib)
        queue = _Queue(n)
        st, mem, off = status.tolist(), members.tolist(), offsets.tolist()
        for i, g in enumerate(gates_in):
            if st[i] == 1:
                queue.append(g)
            elif st[i] == 2:
                group = [gates_in[j] for j in mem[off[i]: off[i + 1]]]
                qubits = sorted({q for h in group for q in h.qubits})
                fgate = gates.FusedGate(*qubits)
                for h in group:
                    fgate.append(h)
                queue.append(fgate)
        circuit = self._shallow_copy()
        circuit.queue = queue
        return circuit

    def _fuse_python(self, max_qubits: int = 2):
        """Circuit.fuse's planner in Python, step for step as the reference."""
        queue = self.queue.to_fused()
        for gate in queue:
            if not gate.marked:
                for q in gate.qubits:
                    neighbor = gate.right_neighbors.get(q)
                    if gate.can_fuse(neighbor, max_qubits):
                        gate.fuse(neighbor)
                    neighbor = gate.left_neighbors.get(q)
                    if gate.can_fuse(neighbor, max_qubits):
                        neighbor.fuse(gate)
        circuit = self._shallow_copy()
        circuit.queue = queue.from_fused()
        return circuit

    def unitary(self, backend=None):
        from .backends import _check_backend

        backend = _check_backend(backend)
        fgate = gates.FusedGate(*range(self.nqubits))
        for gate in self.queue:
            if not isinstance(gate, (gates.SpecialGate, gates.M)):
                fgate.append(gate)
        return fgate.matrix(backend)

    # -- execution ---------------------------------------------------------
    @property
    def final_state(self):
        if self._final_state is None:
            raise_error(RuntimeError, "Cannot access final state before the circuit is executed.")
        return self._final_state

    def execute(self, initial_state=None, nshots: int = 1000):
        from .backends import _Global

        backend = _Global.backend()
        return backend.execute_circuit(self, initial_state, nshots)

    def __call__(self, initial_state=None, nshots: int = 1000):
        return self.execute(initial_state, nshots)

    def __repr__(self):
        return f"Circuit(nqubits={self.nqubits}, ngates={self.ngates})"
