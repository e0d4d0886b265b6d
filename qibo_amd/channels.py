"""Noise channels (src/qibo/gates/channels.py).

The same classes, constructor checks and Kraus / Pauli decompositions as the
reference, and the same dispatch: every channel hands the state to a backend
method (``apply_channel``, ``depolarizing_error_density_matrix``,
``reset_error_density_matrix``, ``thermal_error_density_matrix``), which runs
it on the device. Kraus channels need density matrices; unitary channels
(UnitaryChannel, PauliNoiseChannel, DepolarizingChannel) also run on state
vectors by drawing one of their unitaries per execution (repeated execution).
"""

from itertools import product
from math import exp, sqrt
from typing import Tuple

import numpy as np

from .config import PRECISION_TOL, raise_error
from .gates import Gate, I, Unitary, X, Y, Z
from .special import FusedGate


class Channel(Gate):
    """Abstract channel: ``coefficients`` and ``gates`` (channels.py:16-44)."""

    def __init__(self):
        super().__init__()
        self.coefficients = tuple()
        self.gates = tuple()

    def controlled_by(self, *q):
        raise_error(ValueError, f"Noise channel cannot be controlled on qubits {q}.")

    def on_qubits(self, qubit_map: dict):
        raise_error(NotImplementedError, "`on_qubits` method is not available for the `Channel` gate.")

    def apply(self, backend, state, nqubits):
        if len(state.shape) == 1:
            raise_error(NotImplementedError, f"{self.__class__.__name__} cannot be applied to statevector.")
        return backend.apply_channel(self, state, nqubits)

    def matrix(self, backend=None):
        raise_error(NotImplementedError, "`matrix` method not defined for Channels.")

    def __repr__(self):
        return f"{self.__class__.__name__}({', '.join(map(str, self.target_qubits))})"


class KrausChannel(Channel):
    """rho -> sum_k A_k rho A_k^dagger (channels.py:203-327)."""

    def __init__(self, qubits, operators):
        super().__init__()
        self.name = "KrausChannel"
        self.draw_label = "K"
        if isinstance(qubits, int):
            qubits = [(qubits,)] * len(operators)
        elif isinstance(qubits, tuple):
            qubits = [qubits] * len(operators)
        elif not isinstance(qubits, list):
            raise_error(TypeError, f"``qubits`` must be of type int or tuple or int. Got {type(qubits)} instead")
        elif not all(isinstance(q, tuple) for q in qubits):
            raise_error(TypeError, "All elements of ``qubits`` list must be tuples.")
        if isinstance(operators[0], Gate):
            if qubits:
                operators = [
                    _moved(op, {op.qubits[j]: qubits[k][j] for j in range(len(op.qubits))})
                    for k, op in enumerate(operators)
                ]
            self.gates = tuple(operators)
            self.target_qubits = tuple(sorted({q for g in operators for q in g.target_qubits}))
        elif len(qubits) != len(operators):
            raise_error(ValueError, f"``qubits`` list has length {len(qubits)} while "
                                    f"{len(operators)} operators were given.")
        else:
            gates, qubitset = [], set()
            for qubit_tuple, matrix in zip(qubits, operators):
                rank = 2 ** len(qubit_tuple)
                shape = tuple(np.shape(matrix))
                if shape != (rank, rank):
                    raise_error(ValueError, f"Invalid Kraus operator shape {shape} for "
                                            f"acting on {len(qubit_tuple)} qubits.")
                gates.append(Unitary(matrix, *list(qubit_tuple)))
                qubitset.update(qubit_tuple)
            self.gates = tuple(gates)
            self.target_qubits = tuple(sorted(qubitset))
        self.init_args = [self.gates]
        self.coefficients = len(self.gates) * (1,)
        self.coefficient_sum = 1


def _moved(gate, qubit_map):
    """``gate`` on the qubits qubit_map[q] (Gate.on_qubits for the gate kinds a
    channel holds: matrices and fused Paulis)."""
    if isinstance(gate, FusedGate):
        new = FusedGate(*(qubit_map[q] for q in gate.target_qubits))
        for g in gate.gates:
            new.append(_moved(g, qubit_map))
        return new
    if isinstance(gate, Unitary):
        return Unitary(gate.parameters[0], *(qubit_map[q] for q in gate.target_qubits))
    return gate.__class__(*(qubit_map[q] for q in gate.target_qubits))


class UnitaryChannel(KrausChannel):
    """(1 - sum p_k) rho + sum_k p_k U_k rho U_k^dagger (channels.py:329-388)."""

    def __init__(self, qubits, operators):
        if not all(isinstance(pair, tuple) for pair in operators):
            raise_error(TypeError, "``operators`` must be a list of tuples ``(pk, Uk)``.")
        probabilities = [pair[0] for pair in operators]
        operators = [pair[1] for pair in operators]
        if any((p < 0 or p > 1) for p in probabilities):
            raise_error(ValueError, "Probabilities should be between 0 and 1.")
        super().__init__(qubits, operators)
        self.name = "UnitaryChannel"
        self.draw_label = "U"
        self.coefficients = tuple(probabilities)
        self.coefficient_sum = sum(probabilities)
        if self.coefficient_sum > 1 + PRECISION_TOL or self.coefficient_sum < 0:
            raise_error(ValueError, "UnitaryChannel probability sum should be "
                                    f"between 0 and 1 but is {self.coefficient_sum}.")
        self.init_args = [probabilities, self.gates]
        self._qubits = qubits

    def apply(self, backend, state, nqubits):
        return backend.apply_channel(self, state, nqubits)


class PauliNoiseChannel(UnitaryChannel):
    """Pauli strings P_k with probabilities p_k (channels.py:391-468)."""

    def __init__(self, qubits, operators: list):
        if isinstance(qubits, int):
            qubits = (qubits,)
        probabilities, paulis = [], []
        for pauli, probability in operators:
            probabilities.append(probability)
            paulis.append(pauli)
        single = {"I": I, "X": X, "Y": Y, "Z": Z}
        gates = []
        for pauli in paulis:
            fgate = FusedGate(*qubits)
            for qubit, p in zip(qubits, pauli):
                fgate.append(single[p](qubit))
            gates.append(fgate)
        self.gates = tuple(gates)
        self.coefficients = tuple(probabilities)
        super().__init__([], list(zip(probabilities, gates)))
        self.name = "PauliNoiseChannel"
        self.draw_label = "PN"
        self.init_args = qubits
        self.init_kwargs = dict(operators)


class DepolarizingChannel(PauliNoiseChannel):
    """(1 - lam) rho + lam Tr_q(rho) (x) I / 2^n (channels.py:471-528)."""

    def __init__(self, qubits, lam: float):
        if isinstance(qubits, int):
            qubits = (qubits,)
        num_terms = 4 ** len(qubits)
        max_param = num_terms / (num_terms - 1)
        if lam < 0 or lam > max_param:
            raise_error(ValueError, f"Depolarizing parameter must be in between 0 and {max_param}.")
        params = list(product(["I", "X", "Y", "Z"], repeat=len(qubits)))[1::]
        super().__init__(qubits, list(zip(params, [lam / num_terms] * (num_terms - 1))))
        self.name = "DepolarizingChannel"
        self.draw_label = "D"
        self.target_qubits = qubits
        self.init_args = [qubits]
        self.init_kwargs = {"lam": lam}

    def apply(self, backend, state, nqubits):
        if len(state.shape) == 2:
            return backend.depolarizing_error_density_matrix(self, state, nqubits)
        return super().apply(backend, state, nqubits)


class ThermalRelaxationChannel(KrausChannel):
    """Thermal relaxation with times t_1, t_2, gate time and excited-state
    population (channels.py:531-707)."""

    def __init__(self, qubit: int, parameters: list):
        if len(parameters) not in (3, 4):
            raise_error(ValueError, "``parameters`` list must have 3 or 4 elements "
                                    f"while {len(parameters)} were given.")
        t_1, t_2, time = parameters[:3]
        excited_population = parameters[-1] if len(parameters) == 4 else 0.0
        if excited_population < 0 or excited_population > 1:
            raise_error(ValueError, f"Invalid excited state population {excited_population}.")
        if time < 0.0:
            raise_error(ValueError, f"Invalid gate time: {time} < 0.")
        if t_1 <= 0.0:
            raise_error(ValueError, f"Invalid t_1 relaxation time parameter: {t_1} <= 0.0.")
        if t_2 <= 0.0:
            raise_error(ValueError, f"Invalid t_2 relaxation time parameter: {t_2} <= 0.0.")
        if t_2 > 2 * t_1:
            raise_error(ValueError, "Invalid t_2 relaxation time parameter: t_2 > 2 * t_1.")
        p_reset = 1 - exp(-time / t_1)
        p_0 = p_reset * (1 - excited_population)
        p_1 = p_reset * excited_population
        if t_1 < t_2:
            e_t2 = exp(-time / t_2)
            operators = [
                sqrt(p_0) * sqrt(p_0) * np.array([[0, 1], [0, 0]]),
                sqrt(p_1) * np.array([[0, 0], [1, 0]]),
            ]
            k_term = sqrt(4 * e_t2**2 + (p_0 - p_1) ** 2)
            kraus_coeff = sqrt(1 - (p_0 + p_1 + k_term) / 2)
            operators.append(kraus_coeff * np.array([[(p_0 - p_1 - k_term) / (2 * e_t2), 0], [0, 1]]))
            operators.append(kraus_coeff * np.array([[(p_0 - p_1 + k_term) / (2 * e_t2), 0], [0, 1]]))
            super().__init__([(qubit,)] * len(operators), operators)
            self.init_kwargs["e_t2"] = e_t2
        else:
            p_z = (exp(-time / t_1) - exp(-time / t_2)) / 2
            operators = (
                sqrt(p_0) * np.array([[1, 0], [0, 0]]),
                sqrt(p_0) * np.array([[0, 1], [0, 0]]),
                sqrt(p_1) * np.array([[0, 0], [1, 0]]),
                sqrt(p_1) * np.array([[0, 0], [0, 1]]),
                sqrt(p_z) * np.array([[1, 0], [0, -1]]),
                sqrt(1 - p_0 - p_1 - p_z) * np.eye(2),
            )
            super().__init__([(qubit,)] * len(operators), list(operators))
            self.init_kwargs["p_z"] = p_z
        self.init_args = [qubit, t_1, t_2, time]
        self.t_1, self.t_2 = t_1, t_2
        self.init_kwargs["excited_population"] = excited_population
        self.init_kwargs["p_0"] = p_0
        self.init_kwargs["p_1"] = p_1
        self.name = "ThermalRelaxationChannel"
        self.draw_label = "TR"

    def apply(self, backend, state, nqubits):
        """t_1 < t_2: the 4x4 map on (row qubit q, column qubit q) of the
        vectorized matrix; otherwise reset(rho) - p_z rho + p_z Z rho Z with Z on
        qubit 0, as the reference writes it (channels.py:678-707)."""
        qubit = self.target_qubits[0]
        if self.t_1 < self.t_2:
            p0, p1, e_t2 = self.init_kwargs["p_0"], self.init_kwargs["p_1"], self.init_kwargs["e_t2"]
            matrix = np.array([[1 - p1, 0, 0, p0], [0, e_t2, 0, 0], [0, 0, e_t2, 0], [p1, 0, 0, 1 - p0]])
            gate = Unitary(matrix, qubit, qubit + nqubits)
            return backend.thermal_error_density_matrix(gate, state, nqubits)
        p_z = self.init_kwargs["p_z"]
        out = backend.reset_error_density_matrix(self, backend.cast(state, copy=True), nqubits)
        backend.dm_axpy(out, state, -p_z)
        zrho = backend.apply_gate(Z(0), backend.cast(state, copy=True), nqubits)
        backend.dm_axpy(out, zrho, p_z)
        return out


class AmplitudeDampingChannel(KrausChannel):
    """Kraus operators [[1, 0], [0, sqrt(1 - g)]], [[0, sqrt(g)], [0, 0]] (channels.py:710-747)."""

    def __init__(self, qubit, gamma: float):
        if not isinstance(gamma, float):
            raise_error(TypeError, f"gamma must be type float, but it is type {type(gamma)}.")
        if gamma < 0.0 or gamma > 1.0:
            raise_error(ValueError, "gamma must be a float between 0 and 1.")
        operators = [np.array([[1, 0], [0, sqrt(1 - gamma)]], dtype=complex),
                     np.array([[0, sqrt(gamma)], [0, 0]], dtype=complex)]
        super().__init__([(qubit,)] * len(operators), operators)
        self.name = "AmplitudeDampingChannel"
        self.draw_label = "AD"
        self.init_kwargs["gamma"] = gamma


class PhaseDampingChannel(KrausChannel):
    """Kraus operators [[1, 0], [0, sqrt(1 - g)]], [[0, 0], [0, sqrt(g)]] (channels.py:750-787)."""

    def __init__(self, qubit, gamma: float):
        if not isinstance(gamma, float):
            raise_error(TypeError, f"gamma must be type float, but it is type {type(gamma)}.")
        if gamma < 0.0 or gamma > 1.0:
            raise_error(ValueError, "gamma must be a float between 0 and 1.")
        operators = [np.array([[1, 0], [0, sqrt(1 - gamma)]], dtype=complex),
                     np.array([[0, 0], [0, sqrt(gamma)]], dtype=complex)]
        super().__init__([(qubit,)] * len(operators), operators)
        self.name = "PhaseDampingChannel"
        self.draw_label = "PD"
        self.init_kwargs["gamma"] = gamma


class ReadoutErrorChannel(KrausChannel):
    """Quantum-to-classical readout error: A_jk = sqrt(P[k, j]) |j><k| (channels.py:790-835)."""

    def __init__(self, qubits, probabilities):
        if any(sum(row) < 1 - PRECISION_TOL for row in probabilities) or any(
            sum(row) > 1 + PRECISION_TOL for row in probabilities
        ):
            raise_error(ValueError, "all rows of probabilities must sum to 1.")
        if isinstance(qubits, int):
            qubits = (qubits,)
        if isinstance(probabilities, list):
            probabilities = np.array(probabilities)
        dim = len(probabilities)
        operators = []
        for j in range(dim):
            for k in range(dim):
                operator = np.zeros((dim, dim))
                operator[j, k] = sqrt(probabilities[k, j])
                operators.append(operator)
        super().__init__([qubits] * len(operators), operators)
        self.name = "ReadoutErrorChannel"
        self.draw_label = "RE"
        self.probabilities = probabilities
        self._qubits = qubits


class ResetChannel(KrausChannel):
    """(1 - p0 - p1) rho + Tr_q(rho) (x) (p0 |0><0| + p1 |1><1|) (channels.py:838-890)."""

    def __init__(self, qubit, probabilities):
        if len(probabilities) != 2:
            raise_error(ValueError, f"ResetChannel needs 2 probabilities, got {len(probabilities)} instead.")
        p_0, p_1 = probabilities
        if p_0 < 0:
            raise_error(ValueError, "Invalid p_0 ({p_0} < 0).")
        if p_1 < 0:
            raise_error(ValueError, "Invalid p_1 ({p_1} < 0).")
        if p_0 + p_1 > 1 + PRECISION_TOL:
            raise_error(ValueError, f"Invalid probabilities (p_0 + p_1 = {p_0 + p_1} > 1).")
        operators = [
            sqrt(p_0) * np.array([[1, 0], [0, 0]]),
            sqrt(p_0) * np.array([[0, 1], [0, 0]]),
            sqrt(p_1) * np.array([[0, 0], [1, 0]]),
            sqrt(p_1) * np.array([[0, 0], [0, 1]]),
        ]
        if p_0 + p_1 < 1:
            operators.append(sqrt(np.abs(1 - p_0 - p_1)) * np.eye(2))
        super().__init__([(qubit,)] * len(operators), operators)
        self.init_kwargs = {"p_0": p_0, "p_1": p_1}
        self.name = "ResetChannel"
        self.draw_label = "R"

    def apply(self, backend, state, nqubits):
        return backend.reset_error_density_matrix(self, state, nqubits)
