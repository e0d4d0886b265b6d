"""Pauli-sum Hamiltonians and their expectation values.

A ``SymbolicHamiltonian`` here is a list of terms ``coefficient * P_q0 P_q1 ...``
with P in {X, Y, Z} -- the ``simple_terms`` form that qibo's
``SymbolicHamiltonian.expectation`` hands to
``Backend.exp_value_observable_symbolic``
(``src/qibo/hamiltonians/hamiltonians.py:413-426, 586-627``), built directly
instead of through sympy. ``TFIM`` reproduces the symbolic (``dense=False``)
form of qibo's model (``src/qibo/hamiltonians/models.py:587-662``):
H = -sum_k (Z_k Z_{k+1} + h X_k), open or closed boundary.
"""

from functools import reduce
from typing import List, Sequence, Tuple

import numpy as np

from .config import raise_error

_PAULI = {
    "I": np.eye(2, dtype=complex),
    "X": np.array([[0, 1], [1, 0]], dtype=complex),
    "Y": np.array([[0, -1j], [1j, 0]], dtype=complex),
    "Z": np.array([[1, 0], [0, -1]], dtype=complex),
}


class SymbolicHamiltonian:
    def __init__(self, terms: Sequence[Tuple[float, str, Tuple[int, ...]]], nqubits: int,
                 constant: float = 0.0, backend=None):
        merged: List[list] = []
        for coeff, paulis, qubits in terms:
            if len(paulis) != len(qubits):
                raise_error(ValueError, "Pauli string and qubit tuple lengths differ.")
            keep = [(p, q) for p, q in zip(paulis, qubits) if p != "I"]
            if not keep:
                constant += float(np.real(coeff))
                continue
            key = frozenset(keep)
            for entry in merged:
                if entry[0] == key:
                    entry[1] += coeff
                    break
            else:
                merged.append([key, coeff, keep])
        self.nqubits = nqubits
        self.constant = constant
        self._terms = [(c, "".join(p for p, _ in keep), tuple(q for _, q in keep))
                       for _, c, keep in merged]
        self.backend = backend

    @property
    def simple_terms(self):
        """(coefficients, Pauli strings, qubit tuples), as qibo's simple_terms."""
        return ([c for c, _, _ in self._terms], [s for _, s, _ in self._terms],
                [q for _, _, q in self._terms])

    @property
    def terms(self):
        return list(self._terms)

    def _backend(self):
        from .backends import _check_backend

        return _check_backend(self.backend)

    def expectation(self, circuit, nshots=None):
        """<psi|H|psi> for the final state of ``circuit`` (executing it if needed)."""
        if nshots is not None:
            raise_error(NotImplementedError, "Shot-based expectation values are not supported.")
        from .circuit import Circuit

        backend = self._backend()
        coeffs, strings, qubits = self.simple_terms
        if isinstance(circuit, Circuit):
            return self.constant + backend.exp_value_observable_symbolic(
                circuit, strings, qubits, coeffs, self.nqubits)
        return self.expectation_from_state(circuit)

    def expectation_from_state(self, state):
        backend = self._backend()
        coeffs, strings, qubits = self.simple_terms
        state = backend.cast(state, dtype=state.dtype) if hasattr(state, "dtype") else backend.cast(state)
        xm, zm, ny = backend.pauli_tables(strings, qubits, self.nqubits)
        return self.constant + float(backend.expectation_pauli(state, self.nqubits, xm, zm, ny,
                                                               [float(np.real(c)) for c in coeffs]))

    def pauli_tables(self):
        coeffs, strings, qubits = self.simple_terms
        return self._backend().pauli_tables(strings, qubits, self.nqubits) + (
            np.asarray([float(np.real(c)) for c in coeffs]),)

    def matrix(self) -> np.ndarray:
        """Dense 2^n x 2^n matrix (host, small n only)."""
        if self.nqubits > 14:
            raise_error(ValueError, "Dense Hamiltonian matrices are limited to 14 qubits.")
        dim = 2**self.nqubits
        mat = self.constant * np.eye(dim, dtype=complex)
        for coeff, paulis, qubits in self._terms:
            ops = [_PAULI["I"]] * self.nqubits
            for p, q in zip(paulis, qubits):
                ops[q] = _PAULI[p]
            mat = mat + coeff * reduce(np.kron, ops)
        return mat


def TFIM(nqubits: int, h: float = 0.0, dense: bool = False, closed_boundary: bool = True,
         backend=None) -> SymbolicHamiltonian:
    if nqubits < 2:
        raise_error(ValueError, "Number of qubits must be larger than one.")
    if dense:
        raise_error(NotImplementedError, "Only the symbolic (dense=False) TFIM is provided.")
    terms = []
    for q in range(nqubits - 1):
        terms.append((-1.0, "ZZ", (q, q + 1)))
        terms.append((-h, "X", (q,)))
    if closed_boundary and nqubits > 2:
        terms.append((-1.0, "ZZ", (nqubits - 1, 0)))
    terms.append((-h, "X", (nqubits - 1,)))
    if h == 0:
        terms = [t for t in terms if t[1] != "X"]
    return SymbolicHamiltonian(terms, nqubits, backend=backend)


def _one_body(nqubits: int, pauli: str, backend=None) -> SymbolicHamiltonian:
    return SymbolicHamiltonian([(-1.0, pauli, (q,)) for q in range(nqubits)], nqubits, backend=backend)


def X(nqubits: int, dense: bool = False, backend=None):
    """H = -sum_i X_i."""
    return _one_body(nqubits, "X", backend)


def Y(nqubits: int, dense: bool = False, backend=None):
    """H = -sum_i Y_i."""
    return _one_body(nqubits, "Y", backend)


def Z(nqubits: int, dense: bool = False, backend=None):
    """H = -sum_i Z_i."""
    return _one_body(nqubits, "Z", backend)
