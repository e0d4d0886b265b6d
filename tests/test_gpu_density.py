"""Density matrices and noise channels on the GPU (needs an MI355X).

Checked against the oracle's restatement of the reference's density-matrix
path (oracle/density.py, pinned in test_oracle_density.py):
* noisy circuits (every channel class, controlled gates, fused gates) from
  |0><0| and from a mixed initial state, complex128 to 1e-12 and complex64 to
  2e-5, on 3..7 qubits (7 qubits = a 14-qubit vectorized matrix, which runs
  as blocked passes), with one-qubit channels as transfer matrices inside the
  passes and term by term;
* probabilities of the same matrix: all qubits bit for bit, marginals to 1e-14;
* Pauli-sum expectations to 1e-12;
* the collapse of a matrix is the normalised (outcome, outcome) block;
* repeated execution with a collapsing measurement: the mean final matrix is
  diagonal with the outcome frequencies;
* unitary channels on state vectors draw their gates from numpy's global
  generator in the reference's order: the samples equal the oracle's, shot by
  shot, for the same seed.
"""

import numpy as np
import pytest
import torch

from oracle import OracleBackend, OracleDensity
from qibo_amd import Circuit, gates

pytestmark = pytest.mark.gpu


def _np(x):
    return x.detach().cpu().numpy() if isinstance(x, torch.Tensor) else np.asarray(x)


def _mixed_state(n, rng, rank=3):
    vs = rng.normal(size=(rank, 2**n)) + 1j * rng.normal(size=(rank, 2**n))
    w = rng.uniform(0.1, 1.0, rank)
    rho = sum(wk * np.outer(v, v.conj()) / np.vdot(v, v) for wk, v in zip(w, vs))
    return rho / np.trace(rho)


def _noisy_circuit(n, rng):
    c = Circuit(n, density_matrix=True)
    for q in range(n):
        c.add(gates.H(q))
        c.add(gates.RY(q, rng.uniform(0, 6)))
    c.add(gates.CNOT(0, n - 1))
    c.add(gates.AmplitudeDampingChannel(1, 0.3))
    c.add(gates.CU1(n - 1, 1, rng.uniform(0, 6)))
    c.add(gates.DepolarizingChannel((0, 2), 0.2))
    c.add(gates.SWAP(1, n - 1))
    c.add(gates.PauliNoiseChannel(1, [("X", 0.1), ("Y", 0.05), ("Z", 0.2)]))
    c.add(gates.RX(n - 1, rng.uniform(0, 6)).controlled_by(0, 1))
    c.add(gates.ResetChannel(2, [0.3, 0.1]))
    c.add(gates.ThermalRelaxationChannel(n - 1, [1.0, 1.5, 0.4, 0.1]))
    c.add(gates.ThermalRelaxationChannel(0, [1.0, 0.5, 0.4]))
    u = np.linalg.qr(rng.normal(size=(8, 8)) + 1j * rng.normal(size=(8, 8)))[0]
    c.add(gates.Unitary(u, n - 1, 0, 1))
    c.add(gates.PhaseDampingChannel(0, 0.45))
    c.add(gates.UnitaryChannel([(0,), (1, 2)], [(0.2, np.array([[0, 1], [1, 0]])), (0.3, np.eye(4)[[0, 1, 3, 2]])]))
    c.add(gates.ReadoutErrorChannel(1, np.array([[0.9, 0.1], [0.2, 0.8]])))
    c.add(gates.TOFFOLI(2, 0, 1))
    for q in range(n):
        c.add(gates.RZ(q, rng.uniform(0, 6)))
    c.add(gates.DepolarizingChannel((n - 1,), 0.1))
    c.add(gates.PauliNoiseChannel((n - 1, 1), [("XZ", 0.1), ("YY", 0.05), ("IX", 0.2)]))
    c.add(gates.DepolarizingChannel((0, 1, 2), 0.05))
    return c


@pytest.mark.parametrize("transfer", [True, False])
@pytest.mark.parametrize("n", [3, 5, 6, 7])
def test_noisy_circuit_matches_oracle(backend, n, transfer, rng):
    """One-qubit channels as 4x4 transfer matrices inside the blocked passes
    (the default) and term by term in the reference's order."""
    c = _noisy_circuit(n, rng)
    want = OracleDensity().execute_dm(c)
    backend.dm_superoperators = transfer
    try:
        got = _np(backend.execute_circuit(c).state())
    finally:
        backend.dm_superoperators = True
    assert got.shape == (2**n, 2**n)
    assert np.max(np.abs(got - want)) < 1e-12


@pytest.mark.parametrize("n", [4, 7])
def test_noisy_circuit_complex64(backend64, n, rng):
    c = _noisy_circuit(n, rng)
    want = OracleDensity().execute_dm(c)
    got = _np(backend64.execute_circuit(c).state())
    assert got.dtype == np.complex64
    assert np.max(np.abs(got - want)) < 2e-5


def test_mixed_initial_state(backend, rng):
    n = 6
    c = _noisy_circuit(n, rng)
    rho0 = _mixed_state(n, rng)
    want = OracleDensity().execute_dm(c, initial_state=rho0)
    got = _np(backend.execute_circuit(c, initial_state=rho0.copy()).state())
    assert np.max(np.abs(got - want)) < 1e-12
    with pytest.raises(ValueError):
        backend.execute_circuit(c, initial_state=np.ones(2**n) / 2 ** (n / 2))


def test_fused_circuit(backend, rng):
    n = 6
    c = Circuit(n, density_matrix=True)
    for q in range(n):
        c.add(gates.U3(q, *rng.uniform(0, 6, 3)))
    for q in range(0, n - 1, 2):
        c.add(gates.CZ(q, q + 1))
    for q in range(n):
        c.add(gates.RX(q, rng.uniform(0, 6)))
    fused = c.fuse(max_qubits=3)
    want = OracleDensity().execute_dm(c)
    got = _np(backend.execute_circuit(fused).state())
    assert np.max(np.abs(got - want)) < 1e-12


@pytest.mark.parametrize("n,qubits", [(3, [0, 1, 2]), (5, [4, 0, 2, 1, 3]), (6, [2, 0]), (7, [6]), (7, [3, 1, 5])])
def test_probabilities(backend, n, qubits, rng):
    rho = _mixed_state(n, rng)
    want = OracleDensity().calculate_probabilities_dm(rho, qubits, n)
    got = _np(backend.calculate_probabilities(backend.cast(rho), qubits, n, density_matrix=True))
    if len(qubits) == n:
        assert np.array_equal(got, want)
    else:
        assert np.max(np.abs(got - want)) < 1e-14


def test_expectation(backend, rng):
    from qibo_amd import hamiltonians

    n = 5
    rho = _mixed_state(n, rng)
    ham = hamiltonians.TFIM(n, h=0.7, closed_boundary=False, backend=backend)
    coeffs, strings, qubits = ham.simple_terms
    want = OracleDensity().expectation_dm(rho, strings, qubits, coeffs, n) + ham.constant
    got = ham.expectation_from_state(backend.cast(rho))
    assert abs(got - want) < 1e-12


@pytest.mark.parametrize("n,qubits,shot", [(3, [1], 1), (5, [0, 3], 2), (6, [4, 1, 2], 5)])
def test_collapse(backend, n, qubits, shot, rng):
    rho = _mixed_state(n, rng)
    got = _np(backend.collapse_state(backend.cast(rho), qubits, np.array([shot]), n, density_matrix=True))
    keep = [i for i in range(2**n)
            if int("".join(str((i >> (n - 1 - q)) & 1) for q in qubits), 2) == shot]
    want = np.zeros_like(rho)
    block = rho[np.ix_(keep, keep)]
    want[np.ix_(keep, keep)] = block / np.real(np.trace(block))
    assert np.max(np.abs(got - want)) < 1e-14


def test_repeated_execution_with_collapse(backend):
    n, nshots = 3, 60
    c = Circuit(n, density_matrix=True)
    c.add(gates.H(0))
    c.add(gates.RY(1, 0.7))
    c.add(gates.M(0, collapse=True))
    c.add(gates.M(0, 1))
    np.random.seed(11)
    result = backend.execute_circuit(c, nshots=nshots)
    rho = _np(result.state())
    samples = _np(result.samples())
    assert samples.shape == (nshots, 2)
    assert abs(np.trace(rho) - 1) < 1e-14
    # qubit 0 was collapsed every shot: no coherence between its two values
    b = 1 << (n - 1)
    for i in range(2**n):
        for j in range(2**n):
            if (i ^ j) & b:
                assert abs(rho[i, j]) < 1e-15
    # the weight of |1> on qubit 0 is the frequency of that collapse outcome
    p1 = np.real(sum(rho[i, i] for i in range(2**n) if i & b))
    assert abs(p1 - samples[:, 0].mean()) < 1e-12


def test_unitary_channels_on_state_vectors_follow_the_reference_draws(backend):
    n, nshots = 3, 40
    c = Circuit(n)
    for q in range(n):
        c.add(gates.H(q))
    c.add(gates.PauliNoiseChannel(0, [("X", 0.3), ("Z", 0.2)]))
    c.add(gates.DepolarizingChannel((1,), 0.5))
    c.add(gates.RY(2, 0.4))
    c.add(gates.UnitaryChannel([(2,)], [(0.4, np.array([[0, 1], [1, 0]]))]))
    c.add(gates.M(0, 1, 2))
    assert c.repeated_execution
    np.random.seed(123)
    got = _np(backend.execute_circuit(c, nshots=nshots).samples())
    # the reference's loop: each channel draws one index, then the sample
    ob = OracleBackend()
    np.random.seed(123)
    want = []
    for _ in range(nshots):
        psi = ob.zero_state(n)
        for g in c.queue:
            if g.__class__.__name__ == "M":
                continue
            if hasattr(g, "coefficients"):
                p = tuple(g.coefficients) + (1 - sum(g.coefficients),)
                idx = int(ob.sample_shots(np.asarray(p), 1)[0])
                if idx != len(g.gates):
                    psi = ob.apply_gate(g.gates[idx], psi, n)
                continue
            psi = ob.apply_gate(g, psi, n)
        probs = ob.calculate_probabilities(psi, [0, 1, 2], n)
        s = int(ob.sample_shots(probs, 1)[0])
        want.append([(s >> 2) & 1, (s >> 1) & 1, s & 1])
    assert np.array_equal(got, np.array(want))


def test_kraus_channel_rejects_state_vectors(backend):
    c = Circuit(2)
    c.add(gates.AmplitudeDampingChannel(0, 0.2))
    with pytest.raises(NotImplementedError):
        backend.execute_circuit(c)


def test_half_density_matrix_and_direct_channel_calls(backend, rng):
    """apply_gate_half_density_matrix (U rho) and the backend's channel
    entry points called directly, against numpy on the same matrix."""
    n = 4
    rho = _mixed_state(n, rng)
    g = gates.U3(2, *rng.uniform(0, 6, 3))
    u = OracleBackend().matrix(g)
    full = np.kron(np.kron(np.eye(4), u), np.eye(2))  # qubit 2 of 4
    got = _np(backend.apply_gate_half_density_matrix(g, backend.cast(rho), n))
    assert np.max(np.abs(got - full @ rho)) < 1e-14
    od = OracleDensity()
    for ch in (gates.DepolarizingChannel((1, 3), 0.3), gates.ResetChannel(0, [0.2, 0.5]),
               gates.AmplitudeDampingChannel(2, 0.4), gates.ThermalRelaxationChannel(3, [1.0, 1.5, 0.3])):
        want = od.apply_dm(ch, rho.copy(), n)
        got = _np(ch.apply(backend, backend.cast(rho), n))
        assert np.max(np.abs(got - want)) < 1e-14, ch.__class__.__name__
    with pytest.raises(ValueError):
        backend.depolarizing_error_density_matrix(gates.DepolarizingChannel((0,), 0.1), backend.cast(rho[0]), n)


def test_program_cache_replays_density_circuits(backend, rng):
    """A second execution of the same noisy circuit replays the compiled
    passes (one cache hit) and gives the same matrix bit for bit."""
    n = 7
    c = Circuit(n, density_matrix=True)
    for q in range(n):
        c.add(gates.RY(q, float(rng.uniform(0, 6))))
        c.add(gates.DepolarizingChannel((q,), 0.02))
    c.add(gates.CZ(0, 6))
    c.add(gates.PauliNoiseChannel((1, 4), [("XZ", 0.1), ("YY", 0.05)]))
    first = _np(backend.execute_circuit(c).state())
    hits = backend.program_cache_hits
    c._final_state = None
    second = _np(backend.execute_circuit(c).state())
    assert backend.program_cache_hits == hits + 1
    assert np.array_equal(first, second)
    assert np.max(np.abs(first - OracleDensity().execute_dm(c))) < 1e-12


@pytest.mark.parametrize("n", [1, 2])
def test_smallest_density_matrices(backend, n, rng):
    """One- and two-qubit density matrices (a 2- or 4-qubit vectorized matrix,
    run inside a padded 12-qubit pass), empty circuits and measurements."""
    empty = Circuit(n, density_matrix=True)
    want = OracleDensity().execute_dm(empty)
    assert np.array_equal(_np(backend.execute_circuit(empty).state()), want)
    c = Circuit(n, density_matrix=True)
    c.add(gates.H(0))
    c.add(gates.AmplitudeDampingChannel(0, 0.3))
    c.add(gates.DepolarizingChannel((n - 1,), 0.2))
    if n == 2:
        c.add(gates.CNOT(0, 1))
        c.add(gates.PauliNoiseChannel((0, 1), [("XY", 0.1), ("ZZ", 0.2)]))
    c.add(gates.M(*range(n)))
    want = OracleDensity().execute_dm(c)
    result = backend.execute_circuit(c, nshots=100)
    assert np.max(np.abs(_np(result.state()) - want)) < 1e-14
    p = _np(result.probabilities())
    assert np.max(np.abs(p - np.real(np.diag(want)))) < 1e-14
    assert _np(result.samples()).shape == (100, n)
