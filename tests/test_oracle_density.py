"""The oracle's density-matrix path, pinned on the CPU.

Checks that do not restate the oracle's code:
* a circuit of unitary gates on |0><0| gives |psi><psi| of the state-vector
  oracle (pinned by the reference's golden vectors), controlled_by included;
* the closed-form channels equal the generic Kraus sum of the channel's own
  operators (ResetChannel's reset formula, ThermalRelaxationChannel's
  reset-based branch); the t_1 < t_2 thermal map equals its populations /
  coherences form written per element;
* DepolarizingChannel at lam = 4^k / (4^k - 1) is the uniform Pauli channel
  sum_{P != I} P rho P / (4^k - 1) (the docstring identity of channels.py:478-487),
  at lam = 1 it leaves Tr_T(rho) (x) I / 2^k;
* every channel preserves the trace and hermiticity;
* probabilities and Pauli expectations of |psi><psi| equal the state-vector ones.
"""

import itertools

import numpy as np
import pytest

from oracle import OracleBackend, OracleDensity
from qibo_amd import Circuit, gates


def _mixed_state(n, rng, rank=3):
    vs = rng.normal(size=(rank, 2**n)) + 1j * rng.normal(size=(rank, 2**n))
    w = rng.uniform(0.1, 1.0, rank)
    rho = sum(wk * np.outer(v, v.conj()) / np.vdot(v, v) for wk, v in zip(w, vs))
    return rho / np.trace(rho)


def _unitary_circuit(n, rng, density_matrix=False):
    c = Circuit(n, density_matrix=density_matrix)
    for q in range(n):
        c.add(gates.H(q))
        c.add(gates.RY(q, rng.uniform(0, 6)))
    c.add(gates.CNOT(0, n - 1))
    c.add(gates.CU1(n - 1, 1, rng.uniform(0, 6)))
    c.add(gates.SWAP(1, n - 1))
    c.add(gates.RX(n - 1, rng.uniform(0, 6)).controlled_by(0, 1))
    c.add(gates.U3(1, *rng.uniform(0, 6, 3)).controlled_by(2, 0))
    u = np.linalg.qr(rng.normal(size=(8, 8)) + 1j * rng.normal(size=(8, 8)))[0]
    c.add(gates.Unitary(u, n - 1, 0, 1))
    c.add(gates.TOFFOLI(2, 0, 1))
    return c


@pytest.mark.parametrize("n", [3, 4, 5])
def test_pure_circuit_is_outer_product(n, rng):
    c = _unitary_circuit(n, rng, density_matrix=True)
    psi = OracleBackend().execute(c)
    rho = OracleDensity().execute_dm(c)
    assert np.max(np.abs(rho - np.outer(psi, psi.conj()))) < 1e-13


@pytest.mark.parametrize("n,qubits", [(3, [0, 1, 2]), (4, [2, 0]), (5, [4]), (5, [3, 1, 0])])
def test_probabilities_of_pure_state(n, qubits, rng):
    c = _unitary_circuit(n, rng)
    ob = OracleBackend()
    psi = ob.execute(c)
    p_sv = ob.calculate_probabilities(psi, qubits, n)
    p_dm = OracleDensity().calculate_probabilities_dm(np.outer(psi, psi.conj()), qubits, n)
    assert np.max(np.abs(p_sv - p_dm)) < 1e-14


def test_pauli_expectation_of_pure_state(rng):
    n = 4
    c = _unitary_circuit(n, rng)
    ob = OracleBackend()
    psi = ob.execute(c)
    terms = ["ZZ", "X", "YZ", "XYZ"]
    qubits = [(0, 1), (2,), (3, 0), (1, 2, 3)]
    coeffs = [0.5, -1.25, 0.75, 0.3]
    e_sv = ob.exp_value_observable_symbolic(psi, terms, qubits, coeffs, n)
    e_dm = OracleDensity().expectation_dm(np.outer(psi, psi.conj()), terms, qubits, coeffs, n)
    assert abs(e_sv - e_dm) < 1e-13


@pytest.mark.parametrize("n,qubit,p", [(3, 0, (0.3, 0.2)), (4, 2, (0.0, 0.7)), (4, 3, (1.0, 0.0)), (3, 1, (0.25, 0.75))])
def test_reset_formula_equals_its_kraus_sum(n, qubit, p, rng):
    rho = _mixed_state(n, rng)
    ch = gates.ResetChannel(qubit, list(p))
    od = OracleDensity()
    closed = od.reset_dm(ch, rho.copy(), n)
    kraus = od.apply_channel_dm(ch, rho.copy(), n)
    assert np.max(np.abs(closed - kraus)) < 1e-14


@pytest.mark.parametrize("params", [[1.0, 1.5, 0.3], [2.0, 3.0, 0.7, 0.2]])
def test_thermal_map_closed_form(params, rng):
    # t_1 < t_2: populations relax towards (p0, p1), coherences decay by e_t2,
    # written out per 2x2 block of the target qubit (the channel's Kraus list
    # in the reference, with its p_0 * sigma+ term, is not this map and is
    # never applied)
    n, qubit = 3, 1
    rho = _mixed_state(n, rng)
    ch = gates.ThermalRelaxationChannel(qubit, params)
    p0, p1, e = ch.init_kwargs["p_0"], ch.init_kwargs["p_1"], ch.init_kwargs["e_t2"]
    got = OracleDensity().thermal_dm(ch, rho.copy(), n)
    b = 1 << (n - 1 - qubit)
    want = rho.copy()
    for i in range(2**n):
        for j in range(2**n):
            if i & b or j & b:
                continue
            r00, r01, r10, r11 = rho[i, j], rho[i, j | b], rho[i | b, j], rho[i | b, j | b]
            want[i, j] = (1 - p1) * r00 + p0 * r11
            want[i | b, j | b] = p1 * r00 + (1 - p0) * r11
            want[i, j | b] = e * r01
            want[i | b, j] = e * r10
    assert np.max(np.abs(got - want)) < 1e-15


@pytest.mark.parametrize("params", [[1.0, 0.5, 0.3], [2.0, 1.0, 0.7, 0.4]])
def test_thermal_reset_branch_on_qubit_zero(params, rng):
    # t_1 >= t_2: reset - p_z rho + p_z Z rho Z, with Z on qubit 0 as the
    # reference writes it; on qubit 0 this is the channel's Kraus sum
    n = 3
    rho = _mixed_state(n, rng)
    ch = gates.ThermalRelaxationChannel(0, params)
    od = OracleDensity()
    assert np.max(np.abs(od.thermal_dm(ch, rho.copy(), n) - od.apply_channel_dm(ch, rho.copy(), n))) < 1e-14


@pytest.mark.parametrize("n,targets", [(3, (1,)), (3, (0, 2)), (4, (3, 1)), (4, (2,))])
def test_full_depolarizing_is_uniform_pauli_channel(n, targets, rng):
    rho = _mixed_state(n, rng)
    k = len(targets)
    lam = 4**k / (4**k - 1)
    od = OracleDensity()
    got = od.depolarize_dm(gates.DepolarizingChannel(targets, lam), rho.copy(), n)
    paulis = ["".join(p) for p in itertools.product("IXYZ", repeat=k)][1:]
    pn = gates.PauliNoiseChannel(targets, [(p, 1 / (4**k - 1)) for p in paulis])
    want = od.apply_channel_dm(pn, rho.copy(), n)
    assert np.max(np.abs(got - want)) < 1e-14


@pytest.mark.parametrize("n,targets", [(3, (1,)), (4, (0, 2))])
def test_depolarizing_one_leaves_maximally_mixed_targets(n, targets, rng):
    rho = _mixed_state(n, rng)
    od = OracleDensity()
    got = od.depolarize_dm(gates.DepolarizingChannel(targets, 1.0), rho.copy(), n)
    rest = [q for q in range(n) if q not in targets]
    # the state is the traced-out part times I / 2^k, in the targets' place
    t = od.partial_trace(rho, targets)
    want = np.zeros_like(rho)
    for i in range(2**n):
        for j in range(2**n):
            bi = [(i >> (n - 1 - q)) & 1 for q in range(n)]
            bj = [(j >> (n - 1 - q)) & 1 for q in range(n)]
            if all(bi[q] == bj[q] for q in targets):
                ri = int("".join(str(bi[q]) for q in rest), 2)
                rj = int("".join(str(bj[q]) for q in rest), 2)
                want[i, j] = t[ri, rj] / 2 ** len(targets)
    assert np.max(np.abs(got - want)) < 1e-15


def _channels(n):
    return [
        gates.AmplitudeDampingChannel(1, 0.3),
        gates.PhaseDampingChannel(n - 1, 0.45),
        gates.DepolarizingChannel((0, 2), 0.2),
        gates.PauliNoiseChannel(1, [("X", 0.1), ("Y", 0.05), ("Z", 0.2)]),
        gates.ResetChannel(2, [0.3, 0.1]),
        gates.ThermalRelaxationChannel(0, [1.0, 1.5, 0.4, 0.1]),
        gates.ThermalRelaxationChannel(0, [1.0, 0.5, 0.4]),
        gates.ReadoutErrorChannel(1, np.array([[0.9, 0.1], [0.2, 0.8]])),
        gates.UnitaryChannel([(0,), (1, 2)], [(0.2, np.array([[0, 1], [1, 0]])),
                                              (0.3, np.eye(4)[[0, 1, 3, 2]])]),
        gates.KrausChannel([(0,), (0,)], [np.sqrt(0.6) * np.eye(2), np.sqrt(0.4) * np.diag([1, -1])]),
    ]


@pytest.mark.parametrize("k", range(10))
def test_channels_preserve_trace_and_hermiticity(k, rng):
    n = 3
    rho = _mixed_state(n, rng)
    ch = _channels(n)[k]
    out = OracleDensity().apply_dm(ch, rho.copy(), n)
    assert abs(np.trace(out) - 1) < 1e-14
    assert np.max(np.abs(out - out.conj().T)) < 1e-15
    assert np.min(np.linalg.eigvalsh(out)) > -1e-14


@pytest.mark.parametrize("k", range(10))
def test_channel_transfer_matrices(k, rng):
    """The 4x4 maps that density-matrix circuits run for one-qubit channels
    (GatePlanner.channel_transfer_matrix) equal the oracle's channel
    application: the map applied by the state-vector oracle to the 2n-qubit
    vectorized matrix on (row qubit q, column qubit q)."""
    from oracle.density import _Matrix
    from qibo_amd.backend import GatePlanner

    n = 3
    ch = _channels(n)[k]
    sop = GatePlanner().channel_transfer_matrix(ch)
    t = tuple(ch.target_qubits)
    if len(t) > 2:  # three-qubit channels run term by term
        assert sop is None
        return
    assert sop.shape == (4 ** len(t), 4 ** len(t))
    rho = _mixed_state(n, rng)
    od = OracleDensity()
    got = od.apply_gate(_Matrix(sop, t + tuple(q + n for q in t)), rho.ravel().copy(), 2 * n).reshape(rho.shape)
    assert np.max(np.abs(got - od.apply_dm(ch, rho.copy(), n))) < 1e-15


def test_channels_without_transfer_matrix():
    from qibo_amd.backend import GatePlanner

    gp = GatePlanner()
    assert gp.channel_transfer_matrix(gates.ThermalRelaxationChannel(1, [1.0, 0.5, 0.4])) is None
    assert gp.channel_transfer_matrix(gates.DepolarizingChannel((0, 1, 2), 0.1)) is None


@pytest.mark.parametrize("targets", [(2, 0), (1, 3)])
def test_two_qubit_pauli_channel_transfer_matrix(targets, rng):
    from oracle.density import _Matrix
    from qibo_amd.backend import GatePlanner

    n = 4
    ch = gates.PauliNoiseChannel(targets, [("XZ", 0.1), ("YY", 0.05), ("IX", 0.2), ("ZI", 0.15)])
    sop = GatePlanner().channel_transfer_matrix(ch)
    rho = _mixed_state(n, rng)
    od = OracleDensity()
    t = tuple(ch.target_qubits)
    got = od.apply_gate(_Matrix(sop, t + tuple(q + n for q in t)), rho.ravel().copy(), 2 * n).reshape(rho.shape)
    assert np.max(np.abs(got - od.apply_dm(ch, rho.copy(), n))) < 1e-15
