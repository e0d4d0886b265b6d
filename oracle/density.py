"""numpy restatement of the reference's density-matrix path (test oracle).

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product package.

Restates, with the same numpy calls in the same order:
  apply_gate on matrices ................ abstract.py:2322-2361 + einsum_utils.py:36-55
  _apply_gate_controlled_by_density_matrix abstract.py:3199-3235 + einsum_utils.py:74-92
  apply_channel ......................... abstract.py:2297-2320
  depolarizing_error_density_matrix ...... abstract.py:2095-2131 (partial_trace :2049-2085)
  reset_error_density_matrix ............ abstract.py:2212-2232
  thermal_error_density_matrix .......... abstract.py:2234-2241, channels.py:678-707
  calculate_probabilities (matrices) .... abstract.py:2734-2750, 3371-3381
  exp_value_observable_symbolic (matrices) abstract.py:2946-3005

Channels are duck-typed by class name, as the reference's classes dispatch
in their ``apply`` methods (channels.py:33-41, 383-384, 518-523, 678-707,
886-887).
"""

import numpy as np

from .numpy_backend import EINSUM_CHARS, OracleBackend


def _prepare_strings(qubits, nqubits):
    """einsum_utils.py:13-28."""
    inp = list(EINSUM_CHARS[:nqubits])
    out = inp[:]
    trans = list(EINSUM_CHARS[nqubits: nqubits + len(qubits)])
    for i, q in enumerate(qubits):
        trans.append(inp[q])
        out[q] = trans[i]
    rest = EINSUM_CHARS[nqubits + len(qubits):]
    return "".join(inp), "".join(out), "".join(trans), rest


def dm_strings(qubits, nqubits):
    """einsum_utils.py:36-44: (left, right) contraction strings."""
    inp, out, trans, rest = _prepare_strings(qubits, nqubits)
    trest = rest[:nqubits]
    return f"{inp}{trest},{trans}->{out}{trest}", f"{trest}{inp},{trans}->{trest}{out}"


def dm_controlled_strings(qubits, nqubits):
    """einsum_utils.py:47-55."""
    inp, out, trans, rest = _prepare_strings(qubits, nqubits)
    trest, c = rest[:nqubits], rest[nqubits]
    return f"{c}{inp}{trest},{trans}->{c}{out}{trest}", f"{c}{trest}{inp},{trans}->{c}{trest}{out}"


def _reverse_order(order):
    rorder = len(order) * [0]
    for i, r in enumerate(order):
        rorder[r] = i
    return rorder


class OracleDensity(OracleBackend):
    """The NumpyBackend operations of the density-matrix path."""

    def zero_state_dm(self, nqubits):
        dims = 2**nqubits
        state = np.zeros((dims, dims), dtype=self.dtype)
        state[0, 0] = 1
        return state

    # -- gates -------------------------------------------------------------
    def apply_gate_dm(self, gate, state, nqubits):
        state = np.reshape(state, 2 * nqubits * (2,))
        if gate.is_controlled_by:
            return self._controlled_dm(gate, state, nqubits)
        matrix = self.matrix(gate)
        matrix = np.reshape(matrix, 2 * len(gate.qubits) * (2,))
        matrix_conj = np.conj(matrix)
        left, right = dm_strings(tuple(gate.qubits), nqubits)
        state = np.einsum(right, state, matrix_conj)
        state = np.einsum(left, state, matrix)
        return np.reshape(state, 2 * (2**nqubits,))

    def _controlled_dm(self, gate, state, nqubits):
        matrix = self.matrix(gate)
        matrix = np.reshape(matrix, 2 * len(gate.target_qubits) * (2,))
        matrixc = np.conj(matrix)
        ncontrol = len(gate.control_qubits)
        nactive = nqubits - ncontrol
        dims_ctrl = 2**ncontrol
        order, targets = self.control_order(gate, nqubits)
        additional = [x + len(order) for x in order]
        order = order[:ncontrol] + additional[:ncontrol] + order[ncontrol:] + additional[ncontrol:]
        state = np.transpose(state, order)
        state = np.reshape(state, 2 * (dims_ctrl,) + 2 * nactive * (2,))
        leftc, rightc = dm_controlled_strings(targets, nactive)
        state01 = state[: dims_ctrl - 1, dims_ctrl - 1]
        state01 = np.einsum(rightc, state01, matrixc)
        state10 = state[dims_ctrl - 1, : dims_ctrl - 1]
        state10 = np.einsum(leftc, state10, matrix)
        left, right = dm_strings(targets, nactive)
        state11 = state[dims_ctrl - 1, dims_ctrl - 1]
        state11 = np.einsum(right, state11, matrixc)
        state11 = np.einsum(left, state11, matrix)
        state00 = state[range(dims_ctrl - 1)]
        state00 = state00[:, range(dims_ctrl - 1)]
        state01 = np.concatenate([state00, state01[:, None]], axis=1)
        state10 = np.concatenate([state10, state11[None]], axis=0)
        state = np.concatenate([state01, state10[None]], axis=0)
        state = np.reshape(state, 2 * nqubits * (2,))
        state = np.transpose(state, _reverse_order(order))
        return np.reshape(state, 2 * (2**nqubits,))

    # -- channels ------------------------------------------------------------
    def apply_channel_dm(self, channel, state, nqubits):
        new_state = (1 - channel.coefficient_sum) * state
        for coeff, gate in zip(channel.coefficients, channel.gates):
            new_state += coeff * self.apply_gate_dm(gate, state, nqubits)
        return new_state

    @staticmethod
    def partial_trace(state, traced_qubits):
        nqubits = int(np.log2(state.shape[0]))
        state = np.reshape(state, 2 * nqubits * (2,))
        order = tuple(sorted(traced_qubits))
        order += tuple(set(list(range(nqubits))) ^ set(traced_qubits))
        order += tuple(elem + nqubits for elem in order)
        shape = 2 * (2 ** len(traced_qubits), 2 ** (nqubits - len(traced_qubits)))
        state = np.transpose(state, order)
        state = np.reshape(state, shape)
        return np.einsum("abac->bc", state)

    def depolarize_dm(self, gate, state, nqubits):
        shape = state.shape
        target_qubits = gate.target_qubits
        lam = gate.init_kwargs["lam"]
        trace = self.partial_trace(state, target_qubits)
        trace = np.reshape(trace, 2 * (nqubits - len(target_qubits)) * (2,))
        k = len(target_qubits)
        identity = np.eye(2**k, dtype=self.dtype)
        identity /= 2**k
        identity = np.reshape(identity, 2 * k * (2,))
        identity = np.tensordot(trace, identity, 0)
        qubits = list(range(nqubits))
        for j in target_qubits:
            qubits.pop(qubits.index(j))
        qubits.sort()
        qubits += list(target_qubits)
        qubit_1 = list(range(nqubits - k)) + list(range(2 * (nqubits - k), 2 * nqubits - k))
        qubit_2 = list(range(nqubits - k, 2 * (nqubits - k))) + list(range(2 * nqubits - k, 2 * nqubits))
        order = []
        for qj in (qubit_1, qubit_2):
            order += [qj[qubits.index(q)] for q in range(len(qubits))]
        identity = np.reshape(np.transpose(identity, order), shape)
        return (1 - lam) * state + lam * identity

    def reset_dm(self, gate, state, nqubits):
        shape = state.shape
        qubit = gate.target_qubits[0]
        p_0, p_1 = gate.init_kwargs["p_0"], gate.init_kwargs["p_1"]
        trace = self.partial_trace(state, (qubit,))
        trace = np.reshape(trace, 2 * (nqubits - 1) * (2,))
        zero = self.zero_state_dm(1)
        zero = np.tensordot(trace, zero, 0)
        order = list(range(2 * nqubits - 2))
        order.insert(qubit, 2 * nqubits - 2)
        order.insert(qubit + nqubits, 2 * nqubits - 1)
        zero = np.reshape(np.transpose(zero, order), shape)
        state = (1 - p_0 - p_1) * state + p_0 * zero
        return state + p_1 * self.apply_gate_dm(_Pauli("X", qubit), zero, nqubits)

    def thermal_dm(self, channel, state, nqubits):
        qubit = channel.target_qubits[0]
        if channel.t_1 < channel.t_2:
            p0, p1, e_t2 = channel.init_kwargs["p_0"], channel.init_kwargs["p_1"], channel.init_kwargs["e_t2"]
            matrix = np.array([[1 - p1, 0, 0, p0], [0, e_t2, 0, 0], [0, 0, e_t2, 0], [p1, 0, 0, 1 - p0]],
                              dtype=self.dtype)
            shape = state.shape
            out = self.apply_gate(_Matrix(matrix, (qubit, qubit + nqubits)), state.ravel(), 2 * nqubits)
            return np.reshape(out, shape)
        p_z = channel.init_kwargs["p_z"]
        return (self.reset_dm(channel, state, nqubits) - p_z * state
                + p_z * self.apply_gate_dm(_Pauli("Z", 0), state, nqubits))

    def apply_dm(self, gate, state, nqubits):
        """One queue element on a density matrix, dispatched as the reference's
        gate classes do."""
        name = gate.__class__.__name__
        if name == "M":
            return state
        if name == "DepolarizingChannel":
            return self.depolarize_dm(gate, state, nqubits)
        if name == "ResetChannel":
            return self.reset_dm(gate, state, nqubits)
        if name == "ThermalRelaxationChannel":
            return self.thermal_dm(gate, state, nqubits)
        if hasattr(gate, "coefficients") and hasattr(gate, "coefficient_sum"):
            return self.apply_channel_dm(gate, state, nqubits)
        return self.apply_gate_dm(gate, state, nqubits)

    def execute_dm(self, circuit, initial_state=None):
        nqubits = circuit.nqubits
        if initial_state is None:
            state = self.zero_state_dm(nqubits)
        else:
            state = np.asarray(initial_state, dtype=self.dtype).copy()
        for gate in circuit.queue:
            state = self.apply_dm(gate, state, nqubits)
        return state

    # -- measurement and expectation -----------------------------------------
    def calculate_probabilities_dm(self, state, qubits, nqubits):
        order = tuple(sorted(qubits))
        order += tuple(qubit for qubit in range(nqubits) if qubit not in qubits)
        order = order + tuple(qubit + nqubits for qubit in order)
        shape = 2 * (2 ** len(qubits), 2 ** (nqubits - len(qubits)))
        state = np.reshape(state, 2 * nqubits * (2,))
        state = np.reshape(np.transpose(state, order), shape)
        probs = np.abs(np.einsum("abab->a", state))
        probs = np.reshape(probs, len(qubits) * (2,))
        return self._order_probabilities(probs, qubits, nqubits).ravel()

    def expectation_dm(self, state, terms, term_qubits, term_coefficients, nqubits):
        """sum_t Re(c_t * einsum(P_t factors, rho)) (abstract.py:2979-3005)."""
        from string import ascii_letters

        dims = 2 * nqubits
        state = np.reshape(state, dims * (2,))
        state_indices = [ascii_letters[e] for e in range(dims)]
        state_string = "".join(state_indices)
        expval = 0.0
        for term, qubits, coefficient in zip(terms, term_qubits, term_coefficients):
            mats = {q: getattr(self.matrices, f) for f, q in zip(term, qubits) if f != "I"}
            qubits, matrices = zip(*mats.items())
            obs_string = ",".join(state_indices[i + nqubits] + state_indices[i] for i in qubits)
            new_string = state_string[:]
            for q in set(range(nqubits)) - set(qubits):
                new_string = new_string[:q] + new_string[q + nqubits] + new_string[q + 1:]
            expval += np.real(coefficient * np.einsum(f"{obs_string},{new_string}->", *matrices, state))
        return expval


class _Pauli:
    """A parameter-free one-qubit gate for internal use (X or Z)."""

    def __init__(self, name, qubit):
        self.__class__ = type(name, (_Pauli,), {})
        self.qubits = self.target_qubits = (qubit,)
        self.control_qubits = ()
        self.is_controlled_by = False
        self.parameters = ()


class _Matrix:
    """A matrix on explicit qubits, read as the reference reads ``Unitary``."""

    def __init__(self, matrix, qubits):
        self.__class__ = type("Unitary", (_Matrix,), {})
        self.qubits = self.target_qubits = tuple(qubits)
        self.control_qubits = ()
        self.is_controlled_by = False
        self.parameters = (matrix,)
