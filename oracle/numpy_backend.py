"""numpy restatement of qibo 0.3.5's NumpyBackend hot path (test oracle).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, never by the product package.

Gates are duck-typed: anything with ``__class__.__name__``, ``qubits``,
``target_qubits``, ``control_qubits``, ``is_controlled_by``, ``parameters``
(and ``gates`` for fused groups) works, exactly as qibo's backend reads them.
Each method restates the cited reference lines with the same numpy calls in
the same order, so its arithmetic is the reference's arithmetic.
"""

import cmath
import math
from collections import Counter
from functools import cached_property
from string import ascii_letters

import numpy as np
from scipy.sparse import csr_matrix
from scipy.sparse import eye as eye_sparse

SHOT_BATCH_SIZE = 2**18  # src/qibo/config.py:41
EINSUM_CHARS = "abcdefghijklmnopqrstuvwxyzABCDEFGHIJKLMNOPQRSTUVWXYZ"  # src/qibo/config.py


class OracleMatrices:
    """Gate matrices, restating src/qibo/backends/npmatrices.py:12-577."""

    def __init__(self, dtype):
        self.dtype = dtype

    def _cast(self, array, dtype):
        if isinstance(array, list):
            return np.array(array, dtype=dtype)
        return array.astype(dtype)

    @cached_property
    def H(self):  # npmatrices.py:26-28
        return self._cast([[1, 1], [1, -1]], dtype=self.dtype) / math.sqrt(2)

    @cached_property
    def X(self):  # :30-32
        return self._cast([[0, 1], [1, 0]], dtype=self.dtype)

    @cached_property
    def Y(self):  # :34-36
        return self._cast([[0j, -1j], [1j, 0j]], dtype=self.dtype)

    @cached_property
    def Z(self):  # :38-40
        return self._cast([[1, 0], [0, -1]], dtype=self.dtype)

    @cached_property
    def S(self):  # :50-52
        return self._cast([[1 + 0j, 0j], [0j, 1j]], dtype=self.dtype)

    @cached_property
    def SDG(self):  # :54-56
        return self._cast([[1 + 0j, 0j], [0j, -1j]], dtype=self.dtype)

    @cached_property
    def T(self):  # :58-62
        return self._cast([[1 + 0j, 0], [0, cmath.exp(1j * math.pi / 4.0)]], dtype=self.dtype)

    @cached_property
    def TDG(self):  # :64-68
        return self._cast([[1 + 0j, 0], [0, cmath.exp(-1j * math.pi / 4.0)]], dtype=self.dtype)

    @cached_property
    def SX(self):  # :42-44
        return self._cast([[1 + 1j, 1 - 1j], [1 - 1j, 1 + 1j]], dtype=self.dtype) / 2

    @cached_property
    def SXDG(self):  # :46-48
        return self._cast([[1 - 1j, 1 + 1j], [1 + 1j, 1 - 1j]], dtype=self.dtype) / 2

    def I(self, n=2):  # :70-71
        return np.eye(n, dtype=self.dtype)

    def RX(self, theta):  # :79-82
        cos = np.cos(theta / 2.0) + 0j
        isin = -1j * np.sin(theta / 2.0)
        return self._cast([[cos, isin], [isin, cos]], dtype=self.dtype)

    def RY(self, theta):  # :84-87
        cos = np.cos(theta / 2.0) + 0j
        sin = np.sin(theta / 2.0) + 0j
        return self._cast([[cos, -sin], [sin, cos]], dtype=self.dtype)

    def RZ(self, theta):  # :89-91
        phase = np.exp(0.5j * theta)
        return self._cast([[np.conj(phase), 0], [0, phase]], dtype=self.dtype)

    def U1(self, theta):  # :113-115
        phase = np.exp(1j * theta)
        return self._cast([[1, 0], [0, phase]], dtype=self.dtype)

    def U2(self, phi, lam):  # :117-123
        eplus = np.exp(1j * (phi + lam) / 2.0)
        eminus = np.exp(1j * (phi - lam) / 2.0)
        return self._cast(
            [[np.conj(eplus), -np.conj(eminus)], [eminus, eplus]], dtype=self.dtype
        ) / math.sqrt(2)

    def U3(self, theta, phi, lam):  # :125-136
        cost = np.cos(theta / 2)
        sint = np.sin(theta / 2)
        eplus = np.exp(1j * (phi + lam) / 2.0)
        eminus = np.exp(1j * (phi - lam) / 2.0)
        return self._cast(
            [[np.conj(eplus) * cost, -np.conj(eminus) * sint], [eminus * sint, eplus * cost]],
            dtype=self.dtype,
        )

    @cached_property
    def CNOT(self):  # :143-147
        return self._cast([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]], dtype=self.dtype)

    @cached_property
    def CY(self):  # :149-159
        return self._cast(
            [[1 + 0j, 0j, 0j, 0j], [0j, 1 + 0j, 0j, 0j], [0j, 0j, 0j, -1j], [0j, 0j, 1j, 0j]],
            dtype=self.dtype,
        )

    @cached_property
    def CZ(self):  # :161-165
        return self._cast([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0], [0, 0, 0, -1]], dtype=self.dtype)

    @cached_property
    def CH(self):  # :167-173
        elem = 1 / math.sqrt(2)
        return self._cast(
            [[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, elem, elem], [0, 0, elem, -elem]], dtype=self.dtype
        )

    def CRX(self, theta):  # :203-212
        cos = np.cos(theta / 2.0) + 0j
        isin = -1j * np.sin(theta / 2.0)
        return self._cast(
            [[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, cos, isin], [0, 0, isin, cos]], dtype=self.dtype
        )

    def CRY(self, theta):  # :214-218
        cos = np.cos(theta / 2.0) + 0j
        sin = np.sin(theta / 2.0) + 0j
        return self._cast(
            [[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, cos, -sin], [0, 0, sin, cos]], dtype=self.dtype
        )

    def CRZ(self, theta):  # :220-228
        phase = np.exp(0.5j * theta)
        return self._cast(
            [[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, np.conj(phase), 0], [0, 0, 0, phase]],
            dtype=self.dtype,
        )

    def CU1(self, theta):  # :230-238
        phase = np.exp(1j * theta)
        return self._cast(
            [[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0], [0, 0, 0, phase]], dtype=self.dtype
        )

    @cached_property
    def SWAP(self):  # :264-268
        return self._cast([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]], dtype=self.dtype)

    @cached_property
    def iSWAP(self):  # :270-280
        return self._cast(
            [[1 + 0j, 0j, 0j, 0j], [0j, 0j, 1j, 0j], [0j, 1j, 0j, 0j], [0j, 0j, 0j, 1 + 0j]],
            dtype=self.dtype,
        )

    @cached_property
    def FSWAP(self):  # :306-310
        return self._cast([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, -1]], dtype=self.dtype)

    def fSim(self, theta, phi):  # :312-324
        cost = np.cos(theta) + 0j
        isint = -1j * np.sin(theta)
        phase = np.exp(-1j * phi)
        return self._cast(
            [[1, 0, 0, 0], [0, cost, isint, 0], [0, isint, cost, 0], [0, 0, 0, phase]],
            dtype=self.dtype,
        )

    def RXX(self, theta):  # :352-363
        cos = np.cos(theta / 2.0) + 0j
        isin = -1j * np.sin(theta / 2.0)
        return self._cast(
            [[cos, 0, 0, isin], [0, cos, isin, 0], [0, isin, cos, 0], [isin, 0, 0, cos]],
            dtype=self.dtype,
        )

    def RZZ(self, theta):  # :378-388
        phase = np.exp(0.5j * theta)
        return self._cast(
            [
                [np.conj(phase), 0, 0, 0],
                [0, phase, 0, 0],
                [0, 0, phase, 0],
                [0, 0, 0, np.conj(phase)],
            ],
            dtype=self.dtype,
        )

    @cached_property
    def TOFFOLI(self):  # :484-498
        m = [[1 if i == j else 0 for j in range(8)] for i in range(8)]
        m[6][6], m[6][7], m[7][6], m[7][7] = 0, 1, 1, 0
        return self._cast(m, dtype=self.dtype)

    @cached_property
    def CCZ(self):  # :500-514
        m = [[1 if i == j else 0 for j in range(8)] for i in range(8)]
        m[7][7] = -1
        return self._cast(m, dtype=self.dtype)

    def Unitary(self, u):  # :562-563
        return np.asarray(u, dtype=self.dtype)


class OracleBackend:
    """The NumpyBackend operations of the state-vector path."""

    def __init__(self, dtype="complex128"):
        self.dtype = dtype
        self.matrices = OracleMatrices(dtype)

    # -- matrices (abstract.py:2649-2717) ---------------------------------
    def matrix(self, gate):
        name = gate.__class__.__name__
        if name == "FusedGate":
            return self.matrix_fused(gate)
        _matrix = getattr(self.matrices, name)
        if name == "I":
            _matrix = _matrix(2 ** len(gate.target_qubits))
        elif callable(_matrix):
            return self.matrix_parametrized(gate)
        return np.asarray(_matrix, dtype=_matrix.dtype)

    def matrix_parametrized(self, gate):
        _matrix = getattr(self.matrices, gate.__class__.__name__)(*gate.parameters)
        return np.asarray(_matrix, dtype=_matrix.dtype)

    def matrix_fused(self, fgate):
        """abstract.py:2680-2717: sparse left-products of embedded gate matrices."""
        rank = len(fgate.target_qubits)
        matrix = eye_sparse(2**rank, dtype=self.dtype)
        for gate in fgate.gates:
            gmatrix = self.matrix(gate)
            num_controls = len(gate.control_qubits)
            if num_controls > 0:
                from scipy.linalg import block_diag

                gmatrix = block_diag(np.eye(2 ** len(gate.qubits) - len(gmatrix)), gmatrix)
            eye = np.eye(2 ** (rank - len(gate.qubits)))
            gmatrix = np.kron(gmatrix, eye)
            original_shape = gmatrix.shape
            gmatrix = np.reshape(gmatrix, 2 * rank * (2,))
            qubits = list(gate.qubits)
            indices = qubits + [q for q in fgate.target_qubits if q not in qubits]
            indices = np.argsort(np.asarray(indices, dtype=np.int64))
            transpose_indices = [int(i) for i in indices]
            transpose_indices.extend([int(i) + rank for i in indices])
            gmatrix = np.transpose(gmatrix, transpose_indices)
            gmatrix = np.reshape(gmatrix, original_shape)
            matrix = csr_matrix(gmatrix) @ matrix
        return np.asarray(matrix.toarray(), dtype=matrix.dtype)

    # -- state (abstract.py:2243-2275) --------------------------------------
    def zero_state(self, nqubits):
        state = np.zeros(2**nqubits, dtype=self.dtype)
        state[0] = 1
        return state

    # -- gate application ----------------------------------------------------
    @staticmethod
    def permutations(qubits, nqubits):
        """einsum_utils.py:94-110."""
        fwd_perm = list(qubits) + [q for q in range(nqubits) if q not in qubits]
        inv = sorted(zip(range(nqubits), fwd_perm), key=lambda x: x[1])
        inv_perm = [i for i, _ in inv]
        return fwd_perm, inv_perm

    def apply_gate(self, gate, state, nqubits):
        """abstract.py:2322-2361 (state-vector branch)."""
        shape = nqubits * (2,)
        state = np.reshape(state, shape)
        if gate.is_controlled_by:
            return self._apply_gate_controlled_by(gate, state, nqubits)
        matrix = self.matrix(gate)
        fwd_perm, inv_perm = self.permutations(tuple(gate.qubits), nqubits)
        state = np.transpose(state, fwd_perm)
        state = np.reshape(state, (2 ** len(gate.qubits), -1))
        state = matrix @ state
        state = np.reshape(state, shape)
        state = np.transpose(state, inv_perm)
        return np.reshape(state, (2**nqubits,))

    @staticmethod
    def control_order(gate, nqubits):
        """einsum_utils.py:58-71."""
        loop_start = 0
        order = list(gate.control_qubits)
        targets = list(gate.target_qubits)
        for control in gate.control_qubits:
            for i in range(loop_start, control):
                order.append(i)
            loop_start = control + 1
            for i, t in enumerate(gate.target_qubits):
                if t > control:
                    targets[i] -= 1
        for i in range(loop_start, nqubits):
            order.append(i)
        return order, targets

    @staticmethod
    def apply_gate_string(qubits, nqubits):
        """einsum_utils.py:13-33."""
        inp = list(EINSUM_CHARS[:nqubits])
        out = inp[:]
        trans = list(EINSUM_CHARS[nqubits: nqubits + len(qubits)])
        for i, q in enumerate(qubits):
            trans.append(inp[q])
            out[q] = trans[i]
        return f"{''.join(inp)},{''.join(trans)}->{''.join(out)}"

    def _apply_gate_controlled_by(self, gate, state, nqubits):
        """abstract.py:3176-3197."""
        matrix = self.matrix(gate)
        matrix = np.reshape(matrix, 2 * len(gate.target_qubits) * (2,))
        ncontrol = len(gate.control_qubits)
        nactive = nqubits - ncontrol
        order, targets = self.control_order(gate, nqubits)
        state = np.transpose(state, order)
        state = np.reshape(state, (2**ncontrol,) + nactive * (2,))
        opstring = self.apply_gate_string(targets, nactive)
        updates = np.einsum(opstring, state[-1], matrix)
        state = np.concatenate([state[:-1], updates[None]], axis=0)
        state = np.reshape(state, nqubits * (2,))
        rorder = [0] * len(order)
        for i, r in enumerate(order):
            rorder[r] = i
        state = np.transpose(state, rorder)
        return np.reshape(state, (2**nqubits,))

    def execute(self, circuit, initial_state=None):
        """_execute_circuit's gate loop (abstract.py:3306-3322); returns the state."""
        nqubits = circuit.nqubits
        if initial_state is None:
            state = self.zero_state(nqubits)
        else:
            state = np.asarray(initial_state, dtype=self.dtype).copy()
        for gate in circuit.queue:
            name = gate.__class__.__name__
            if name == "M":
                continue  # non-collapsing measurement: no-op (measurements.py:189-197)
            state = self.apply_gate(gate, state, nqubits)
        return state

    # -- collapse and repeated execution (state vectors) -----------------------
    def collapse_state(self, state, qubits, shot, nqubits, normalize=True):
        """abstract.py:3279-3304 (_collapse_statevector) with _append_zeros
        (abstract.py:3236-3247)."""
        shape = state.shape
        binshot = list(self.samples_to_binary(np.asarray(shot), len(qubits))[0])
        state = np.reshape(state, nqubits * (2,))
        order = list(qubits) + [q for q in range(nqubits) if q not in qubits]
        state = np.transpose(state, order)
        subshape = (2 ** len(qubits),) + (nqubits - len(qubits)) * (2,)
        state = np.reshape(state, subshape)[int(shot[0])]
        if normalize:
            norm = np.sqrt(np.sum(np.abs(state) ** 2))
            state = state / norm
        for q, r in zip(qubits, binshot):
            state = np.expand_dims(state, q)
            if r == 1:
                state = np.concatenate([np.zeros_like(state), state], axis=q)
            else:
                state = np.concatenate([state, np.zeros_like(state)], axis=q)
        return np.reshape(state, shape)

    def execute_repeated(self, circuit, nshots, initial_state=None):
        """execute_circuit_repeated (abstract.py:2532-2636), state-vector branch:
        per shot, a copy of the initial state runs the queue; a collapsing M
        samples one outcome of its sorted qubits (MeasurementResult.add_shot,
        measurements.py:114-123, one draw from np.random) and collapses
        (gates/measurements.py:189-204); the final measurements are sampled
        once (CircuitResult(..., nshots=1).samples()). Returns the (nshots, m)
        final samples and, per collapsing M (in queue order), its (nshots, k)
        outcomes."""
        nqubits = circuit.nqubits
        state_copy = self.zero_state(nqubits) if initial_state is None else np.array(initial_state, copy=True)
        final_qubits = [q for g in circuit.measurements for q in g.target_qubits]
        collapsing = [g for g in circuit.queue if g.__class__.__name__ == "M" and g.collapse]
        records = [[] for _ in collapsing]
        results = []
        for _ in range(nshots):
            state = state_copy.copy()
            k = 0
            for gate in circuit.queue:
                if gate.__class__.__name__ == "M":
                    if not gate.collapse:
                        continue
                    qubits = sorted(gate.target_qubits)
                    probs = self.calculate_probabilities(state, qubits, nqubits)
                    shot = self.sample_shots(probs, 1)
                    bits = self.samples_to_binary(shot, len(qubits))[0]
                    records[k].append(bits)
                    gate.result.register_samples(np.asarray([bits]))  # the outcome symbols read
                    k += 1
                    state = self.collapse_state(state, qubits, shot, nqubits)
                    continue
                if gate.symbolic_parameters:  # abstract.py:2590-2591
                    gate.substitute_symbols()
                state = self.apply_gate(gate, state, nqubits)
            if final_qubits:
                probs = self.calculate_probabilities(state, final_qubits, nqubits)
                sample = self.sample_shots(probs, 1)
                results.append(self.samples_to_binary(sample, len(final_qubits))[0])
        return np.asarray(results), [np.asarray(r) for r in records]

    # -- probabilities (abstract.py:2734-2758, 3371-3381) ---------------------
    def calculate_probabilities(self, state, qubits, nqubits):
        rtype = np.real(state).dtype
        unmeasured_qubits = tuple(set(list(range(nqubits))) ^ set(qubits))
        state = np.reshape(np.abs(state) ** 2, nqubits * (2,))
        probs = state.astype(rtype, copy=False)
        if len(unmeasured_qubits) != 0:
            probs = np.sum(probs, axis=unmeasured_qubits)
        return self._order_probabilities(probs, qubits, nqubits).ravel()

    @staticmethod
    def _order_probabilities(probs, qubits, nqubits):
        unmeasured, reduced = [], {}
        for qubit in range(nqubits):
            if qubit in qubits:
                reduced[qubit] = qubit - len(unmeasured)
            else:
                unmeasured.append(qubit)
        return np.transpose(probs, [reduced.get(qubit) for qubit in qubits])

    # -- sampling (abstract.py:2760-2801, 1265-1309) ---------------------------
    def sample_shots(self, probabilities, nshots):
        result = np.random.choice(np.arange(len(probabilities)), size=nshots, p=probabilities)
        return np.asarray(result, dtype=np.int64)

    def sample_frequencies(self, probabilities, nshots):
        nprobs = probabilities / np.sum(probabilities)
        frequencies = np.zeros(len(nprobs), dtype=np.int64)
        for _ in range(nshots // SHOT_BATCH_SIZE):
            frequencies = self.update_frequencies(frequencies, nprobs, SHOT_BATCH_SIZE)
        frequencies = self.update_frequencies(frequencies, nprobs, nshots % SHOT_BATCH_SIZE)
        return Counter({i: int(f) for i, f in enumerate(frequencies) if f > 0})

    def update_frequencies(self, frequencies, probabilities, nsamples):
        samples = self.sample_shots(probabilities, nsamples)
        res, counts = np.unique(samples, return_counts=True)
        frequencies[res] += counts
        return frequencies

    @staticmethod
    def apply_bitflips(noiseless_samples, bitflip_probabilities):
        """abstract.py:2279-2295: one uniform per sample bit, then the flips."""
        sprobs = np.random.random(noiseless_samples.shape).astype(np.float64)
        flip_0 = (sprobs < bitflip_probabilities[0]).astype(noiseless_samples.dtype)
        flip_1 = (sprobs < bitflip_probabilities[1]).astype(noiseless_samples.dtype)
        noisy_samples = noiseless_samples + (1 - noiseless_samples) * flip_0
        return noisy_samples - noiseless_samples * flip_1

    @staticmethod
    def samples_to_binary(samples, nqubits):
        qrange = np.arange(nqubits - 1, -1, -1, dtype=np.int32)
        return np.right_shift(samples[:, None], qrange) % 2

    @staticmethod
    def samples_to_decimal(samples, nqubits):
        qrange = np.arange(nqubits - 1, -1, -1, dtype=np.int32)
        qrange = (2**qrange)[:, None]
        samples = np.asarray(samples, dtype=np.int32)
        return (samples @ qrange)[:, 0]

    @staticmethod
    def calculate_frequencies(samples):
        res, counts = np.unique(samples, return_counts=True)
        return Counter(dict(zip(res.tolist(), counts.tolist())))

    # -- expectation (abstract.py:2946-3054, state-vector branch) ---------------
    def exp_value_observable_symbolic(self, state, terms, term_qubits, term_coefficients, nqubits):
        dims = nqubits
        state = np.reshape(state, dims * (2,))
        state_indices = [ascii_letters[e] for e in range(2 * dims)]
        state_dag_indices = state_indices[:nqubits]
        state_indices = state_indices[nqubits:]
        state_dag_string = "".join(state_dag_indices)
        state_string = "".join(state_indices)
        expval = 0.0
        for term, qubits, coefficient in zip(terms, term_qubits, term_coefficients):
            term_matrices = {
                qubit: getattr(self.matrices, factor)
                for factor, qubit in zip(term, qubits)
                if factor != "I"
            }
            qubits, matrices = zip(*term_matrices.items())
            obs_indices = [state_dag_indices[q] + state_indices[q] for q in qubits]
            obs_string = ",".join(obs_indices)
            new_string = state_string[:]
            for q in set(range(nqubits)) - set(qubits):
                new_string = new_string[:q] + state_dag_string[q] + new_string[q + 1:]
            expval += np.real(
                coefficient
                * np.einsum(f"{state_dag_string},{obs_string},{new_string}->",
                            np.conj(state), *matrices, state)
            )
        return expval

    # -- gradient rules (derivative.py:7-226) ------------------------------------
    def expectation(self, circuit, hamiltonian_terms, nqubits, initial_state=None):
        coeffs, strings, qubits = hamiltonian_terms
        state = self.execute(circuit, initial_state)
        return self.exp_value_observable_symbolic(state, strings, qubits, coeffs, nqubits)

    def parameter_shift(self, circuit, hamiltonian_terms, parameter_index, generator_eigenval=0.5,
                        scale_factor=1.0):
        """derivative.py:112-156 (exact branch, nshots=None)."""
        s = np.pi / (4 * generator_eigenval)
        original = np.asarray(_flat_parameters(circuit)).copy()
        shifted = original.copy()
        shifted[parameter_index] += s
        circuit.set_parameters(shifted)
        forward = self.expectation(circuit, hamiltonian_terms, circuit.nqubits)
        shifted[parameter_index] -= 2 * s
        circuit.set_parameters(shifted)
        backward = self.expectation(circuit, hamiltonian_terms, circuit.nqubits)
        circuit.set_parameters(original)
        return (generator_eigenval * (forward - backward) * scale_factor).real

    def finite_differences(self, circuit, hamiltonian_terms, parameter_index, step_size=1e-7):
        """derivative.py:195-226."""
        parameters = np.asarray(_flat_parameters(circuit)).copy()
        shifted = parameters.copy()
        shifted[parameter_index] += step_size
        circuit.set_parameters(shifted)
        forward = self.expectation(circuit, hamiltonian_terms, circuit.nqubits)
        shifted[parameter_index] -= 2 * step_size
        circuit.set_parameters(shifted)
        backward = self.expectation(circuit, hamiltonian_terms, circuit.nqubits)
        circuit.set_parameters(parameters)
        return (forward - backward) / (2 * step_size)


def _flat_parameters(circuit):
    out = []
    for g in circuit.trainable_gates:
        out.extend(float(p) for p in g.parameters)
    return out
