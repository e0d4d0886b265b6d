"""Per-register measurement results (``src/qibo/measurements.py:97-222``).

A ``MeasurementResult`` belongs to one ``M`` gate. The circuit result fills it
with the columns of the global sample matrix that belong to this register, or
with per-register frequencies when only frequencies were sampled.
"""

import collections

import numpy as np

from .config import raise_error


def frequencies_to_binary(frequencies, nqubits):
    return collections.Counter({f"{k:0{nqubits}b}": v for k, v in frequencies.items()})


def _measurement_symbol_class():
    """sympy.Symbol tied to one bit of a measurement's latest outcome
    (measurements.py:33-73); sympy is imported on first use."""
    global _MeasurementSymbol
    if _MeasurementSymbol is None:
        import sympy

        class MeasurementSymbol(sympy.Symbol):
            _counter = 0

            def __new__(cls, *args, **kwargs):
                name = f"m{cls._counter}"
                cls._counter += 1
                return super().__new__(cls=cls, name=name)

            def __init__(self, index, result):
                self.index = index
                self.result = result

            def outcome(self):
                res = self.result
                samples = res._samples if res._samples is not None else res.samples(binary=True)
                return int(_to_numpy(samples)[-1][self.index])

            def evaluate(self, expr):
                return expr.subs(self, self.outcome())

        _MeasurementSymbol = MeasurementSymbol
    return _MeasurementSymbol


_MeasurementSymbol = None


class MeasurementResult:
    def __init__(self, target_qubits):
        self.target_qubits = tuple(target_qubits)
        self.circuit = None
        self.backend = None
        self._samples = None
        self._frequencies = None
        self._symbols = None

    @property
    def symbols(self):
        """One sympy symbol per measured qubit, for gate parameters conditioned
        on this (collapsing) measurement's outcome (measurements.py:147-158)."""
        if self._symbols is None:
            cls = _measurement_symbol_class()
            self._symbols = [cls(i, self) for i in range(len(self.target_qubits))]
        return self._symbols

    def has_samples(self):
        return self._samples is not None

    def register_samples(self, samples):
        self._samples = samples

    def add_shot(self, probs, backend=None):
        """Draw one outcome of the sorted target qubits and append its bits
        (measurements.py:114-123)."""
        from .backends import _check_backend

        backend = _check_backend(backend)
        qubits = sorted(self.target_qubits)
        shot = backend.sample_shots(probs, 1)
        bshot = backend.samples_to_binary(shot, len(qubits))
        if self._samples is not None and len(self._samples):
            self._samples = backend.concatenate([self._samples, bshot[:1]], axis=0)
        else:
            self._samples = bshot[:1]
        return shot

    def register_frequencies(self, frequencies):
        self._frequencies = frequencies

    def reset(self):
        self._samples = None
        self._frequencies = None

    @property
    def nshots(self):
        if self.has_samples():
            return len(self._samples)
        if self._frequencies is not None:
            return sum(self._frequencies.values())
        return None

    def samples(self, binary=True, registers=False, backend=None):
        from .backends import _check_backend

        backend = _check_backend(backend)
        if self._samples is None:
            if self.circuit is None:
                raise_error(RuntimeError, "Cannot calculate samples if circuit is not provided.")
            self.circuit.final_state.samples()
        if binary:
            return self._samples
        return backend.samples_to_decimal(self._samples, len(self.target_qubits))

    def frequencies(self, binary=True, registers=False, backend=None):
        from .backends import _check_backend

        backend = _check_backend(backend)
        if self._frequencies is None:
            self._frequencies = backend.calculate_frequencies(
                self.samples(binary=False, backend=backend)
            )
        if binary:
            return frequencies_to_binary(self._frequencies, len(self.target_qubits))
        return self._frequencies


def _to_numpy(x):
    if hasattr(x, "detach"):
        return x.detach().cpu().numpy()
    return np.asarray(x)
