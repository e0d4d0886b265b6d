"""Result objects returned by circuit execution.

Same behaviour as ``src/qibo/result.py`` for state vectors:

* ``QuantumState`` keeps the final state as the backend's device tensor
  (result.py:30-111); ``state(numpy=True)`` copies it to the host and
  ``probabilities(qubits)`` calls ``backend.calculate_probabilities``.
* ``CircuitResult.__init__`` eagerly computes the probabilities of all measured
  qubits in measurement order (result.py:819-837); samples and frequencies are
  drawn lazily from them (result.py:207-282, 367-441) through
  ``backend.sample_shots`` / ``backend.sample_frequencies`` and registered per
  measurement gate.
"""

import collections
from typing import Optional

import numpy as np

from .config import raise_error
from .measurements import frequencies_to_binary


class QuantumState:
    def __init__(self, state, backend=None):
        from .backends import _check_backend

        self.backend = _check_backend(backend)
        self.density_matrix = len(state.shape) == 2
        self.nqubits = int(np.log2(state.shape[0]))
        self._state = state

    def state(self, numpy: bool = False):
        if numpy:
            return self.backend.to_numpy(self._state)
        return self._state

    def probabilities(self, qubits=None):
        if qubits is None:
            qubits = tuple(range(self.nqubits))
        return self.backend.calculate_probabilities(
            self._state, qubits, self.nqubits, density_matrix=self.density_matrix
        )

    def accuracy(self) -> dict:
        """Infidelity (qibo's 1 - |<state|reference>|^2) and max-abs amplitude
        error of this state against the complex128 execution of the circuit
        that produced it, computed on the GPU (``qibo_amd.accuracy``)."""
        from .accuracy import result_accuracy

        return result_accuracy(self)

    def symbolic(self, decimals: int = 5, cutoff: float = 1e-10, max_terms: int = 20):
        state = self.backend.to_numpy(self._state)
        terms = []
        if self.density_matrix:  # abstract.py:2394-2407
            for j, k in zip(*np.nonzero(state)):
                if abs(state[j, k]) >= cutoff:
                    x = np.round(state[j, k], decimals)
                    terms.append(f"{x}|{int(j):0{self.nqubits}b}><{int(k):0{self.nqubits}b}|")
                if len(terms) >= max_terms:
                    terms.append("...")
                    break
            return " + ".join(terms)
        for j in np.flatnonzero(state):
            if abs(state[j]) >= cutoff:
                terms.append(f"{np.round(state[j], decimals)}|{int(j):0{self.nqubits}b}>")
            if len(terms) >= max_terms:
                terms.append("...")
                break
        return " + ".join(terms)

    def __str__(self):
        return self.symbolic()


class MeasurementOutcomes:
    def __init__(self, measurements, backend=None, probabilities=None, samples=None,
                 nshots=1000, nqubits=None):
        self.backend = backend
        self.measurements = measurements
        self.nshots = nshots
        self._nqubits = nqubits
        self._measurement_gate = None
        self._probs = probabilities
        self._samples = samples
        self._frequencies = None
        self._repeated_execution_frequencies = None  # set by execute_circuit_repeated
        if samples is not None:
            for m in measurements:
                indices = [self.measurement_gate.qubits.index(q) for q in m.qubits]
                m.result.register_samples(samples[:, indices])

    @property
    def measurement_gate(self):
        """Single measurement gate holding all measured qubits (result.py:443-458)."""
        if self._measurement_gate is None:
            from .measurements_gate import M

            for gate in self.measurements:
                if self._measurement_gate is None:
                    kwargs = dict(gate.init_kwargs)
                    kwargs.pop("p0", None)
                    kwargs.pop("p1", None)
                    self._measurement_gate = M(*gate.init_args, **kwargs)
                    self._measurement_gate.bitflip_map[0].update(gate.bitflip_map[0])
                    self._measurement_gate.bitflip_map[1].update(gate.bitflip_map[1])
                else:
                    self._measurement_gate.add(gate)
        return self._measurement_gate

    def has_samples(self):
        return self.measurements[0].result.has_samples() or self._samples is not None

    def frequencies(self, binary: bool = True, registers: bool = False):
        qubits = self.measurement_gate.qubits
        if self._repeated_execution_frequencies is not None:  # result.py:235-241
            if binary:
                return self._repeated_execution_frequencies
            return collections.Counter({int(k, 2): v for k, v in self._repeated_execution_frequencies.items()})
        if self._frequencies is None:
            if self.measurement_gate.has_bitflip_noise() and not self.has_samples():
                self._samples = self.samples()  # frequencies of the noisy samples (result.py:244-245)
            if not self.has_samples():
                self._frequencies = self.backend.sample_frequencies(self._probs, self.nshots)
                qubit_map = {q: i for i, q in enumerate(qubits)}
                binary_frequencies = frequencies_to_binary(self._frequencies, len(qubits))
                for gate in self.measurements:
                    rfreqs = collections.Counter()
                    rqubits = gate.target_qubits
                    for bitstring, freq in binary_frequencies.items():
                        idx = 0
                        for i, q in enumerate(rqubits):
                            if int(bitstring[qubit_map.get(q)]):
                                idx += 2 ** (len(rqubits) - i - 1)
                        rfreqs[idx] += freq
                    gate.result.register_frequencies(rfreqs)
            else:
                self._frequencies = self.backend.calculate_frequencies(self.samples(binary=False))
        if registers:
            return {
                gate.register_name: gate.result.frequencies(binary, backend=self.backend)
                for gate in self.measurements
            }
        if binary:
            return frequencies_to_binary(self._frequencies, len(qubits))
        return self._frequencies

    def samples(self, binary: bool = True, registers: bool = False):
        qubits = self.measurement_gate.target_qubits
        if self._samples is None:
            if self.measurements[0].result.has_samples():
                self._samples = self.backend.concatenate(
                    [gate.result.samples(backend=self.backend) for gate in self.measurements],
                    axis=1,
                )
            else:
                if self._frequencies is not None:
                    freqs = self.frequencies(binary=False)
                    samples = np.concatenate([np.repeat(x, f) for x, f in freqs.items()])
                    np.random.shuffle(samples)
                    samples = self.backend.cast(samples.astype(np.int64), dtype="int64")
                else:
                    samples = self.backend.sample_shots(self._probs, self.nshots)
                samples = self.backend.samples_to_binary(samples, len(qubits))
                if self.measurement_gate.has_bitflip_noise():  # result.py:412-423
                    p0, p1 = self.measurement_gate.bitflip_map
                    probs = np.array([[p0.get(q) for q in qubits], [p1.get(q) for q in qubits]], dtype=np.float64)
                    samples = self.backend.apply_bitflips(samples, probs)
                qubit_map = dict(zip(qubits, range(len(qubits))))
                self._samples = samples
                for gate in self.measurements:
                    rqubits = [qubit_map.get(q) for q in gate.target_qubits]
                    gate.result.register_samples(self._samples[:, rqubits])
        if registers:
            return {
                gate.register_name: gate.result.samples(binary, backend=self.backend)
                for gate in self.measurements
            }
        if binary:
            return self._samples
        return self.backend.samples_to_decimal(self._samples, len(qubits))

    def apply_bitflips(self, p0, p1=None):
        """The samples with bitflips of probabilities ``p0`` (and ``p1``) drawn
        on top (result.py:460-467, measurements.py:19-30)."""
        gate = self.measurement_gate
        t0 = gate._get_bitflip_tuple(gate.qubits, p0)
        t1 = t0 if p1 is None else gate._get_bitflip_tuple(gate.qubits, p1)
        return self.backend.apply_bitflips(self.samples(), np.array([t0, t1], dtype=np.float64))


class CircuitResult(QuantumState, MeasurementOutcomes):
    def __init__(self, final_state, measurements, backend=None, samples=None, nshots=1000):
        QuantumState.__init__(self, final_state, backend)
        qubits = [q for m in measurements for q in m.target_qubits]
        if len(qubits) == 0:
            raise ValueError("Circuit does not contain measurements. Use a `QuantumState` instead.")
        probs = QuantumState.probabilities(self, qubits) if samples is None else None
        MeasurementOutcomes.__init__(
            self, measurements, backend=self.backend, probabilities=probs, samples=samples,
            nshots=nshots, nqubits=self.nqubits,
        )

    def probabilities(self, qubits: Optional[list] = None):
        return QuantumState.probabilities(self, qubits)
