"""Collapsing measurements and repeated execution on the GPU (needs an MI355X).

Checked against the oracle's restatement of execute_circuit_repeated with the
same numpy seed: every shot's collapse outcome and final sample are the
reference's (the draws consume numpy's global generator in the same order);
the collapsed state matches the oracle's to 1e-14.
"""

import numpy as np
import pytest
import torch

from oracle import OracleBackend
from qibo_amd import Circuit, gates

pytestmark = pytest.mark.gpu


def _random_state(n, rng):
    psi = rng.normal(size=2**n) + 1j * rng.normal(size=2**n)
    return psi / np.linalg.norm(psi)


def _np(x):
    return x.detach().cpu().numpy() if isinstance(x, torch.Tensor) else np.asarray(x)


@pytest.mark.parametrize("n,qubits", [(3, [1]), (5, [0, 3, 4]), (13, [2, 7]), (14, [0, 13])])
def test_collapse_state_matches_oracle(backend, n, qubits, rng):
    psi = _random_state(n, rng)
    state = backend.cast(psi, copy=True)
    for shot in (0, 2 ** len(qubits) - 1):
        got = _np(backend.collapse_state(state.clone(), qubits, np.array([shot]), n))
        want = OracleBackend().collapse_state(psi.copy(), qubits, np.array([shot]), n)
        assert np.max(np.abs(got - want)) < 1e-14


def _circuit(n, kind):
    c = Circuit(n)
    if kind == 0:  # Bell pair, collapse the first qubit, rotate, measure all
        c.add(gates.H(0))
        c.add(gates.CNOT(0, 1))
        c.add(gates.M(0, collapse=True))
        c.add(gates.RY(0, 0.7))
        c.add(gates.RX(n - 1, 1.1))
        c.add(gates.M(*range(n)))
    else:  # layers with two mid-circuit collapses on several qubits
        for q in range(n):
            c.add(gates.RY(q, 0.3 + 0.2 * q))
        for q in range(0, n - 1, 2):
            c.add(gates.CZ(q, q + 1))
        c.add(gates.M(1, n - 2, collapse=True))
        for q in range(n):
            c.add(gates.RX(q, 0.5 - 0.1 * q))
        c.add(gates.M(0, n - 1, collapse=True))
        for q in range(1, n - 1, 2):
            c.add(gates.CNOT(q, q + 1))
        c.add(gates.M(*range(0, n, 3)))
    return c


@pytest.mark.parametrize("n", [4, 13])
@pytest.mark.parametrize("kind", [0, 1])
def test_repeated_execution_matches_oracle(backend, n, kind):
    nshots = 12
    c = _circuit(n, kind)
    collapsing = [g for g in c.queue if g.__class__.__name__ == "M" and g.collapse]
    np.random.seed(77)
    result = backend.execute_circuit(c, nshots=nshots)
    got = _np(result.samples())
    got_collapse = [_np(g.result.samples()) for g in collapsing]
    c2 = _circuit(n, kind)
    np.random.seed(77)
    want, want_collapse = OracleBackend().execute_repeated(c2, nshots)
    assert np.array_equal(got, want)
    for a, b in zip(got_collapse, want_collapse):
        assert np.array_equal(a, b)
    # frequencies come from the per-shot bitstrings
    freqs = result.frequencies()
    assert sum(freqs.values()) == nshots
    strings = ["".join(str(int(x)) for x in row) for row in want]
    assert dict(freqs) == {k: strings.count(k) for k in set(strings)}


@pytest.mark.parametrize("n,targets", [(2, [1]), (3, [1]), (4, [1, 3]), (5, [0, 3, 4]), (6, [1, 3]),
                                       (4, [0, 2]), (13, [3, 9])])
def test_measurement_collapse(backend, n, targets, rng):
    """tests/test_measurements_collapse.py:10-25 of the reference: the collapse
    outcome and a following measurement of the same qubits agree."""
    c = Circuit(n)
    for q in range(0, n, 2):
        c.add(gates.H(q))
    r = c.add(gates.M(*targets, collapse=True))
    c.add(gates.M(*targets))
    outcome = backend.execute_circuit(c, backend.cast(_random_state(n, rng), copy=True), nshots=1)
    assert np.array_equal(_np(r.samples())[0], _np(outcome.samples())[0])


@pytest.mark.parametrize("n", [3, 13])
def test_measurement_conditioned_gates_match_oracle(backend, n):
    """Gates whose angles depend on a collapse outcome (r.symbols): every shot
    equals the oracle's for the same seed, on small (one-pass) and blocked
    registers."""
    def circuit():
        c = Circuit(n)
        for q in range(3):
            c.add(gates.H(q))
        r = c.add(gates.M(0, 2, collapse=True))
        c.add(gates.RX(1, theta=np.pi * r.symbols[0] / 3))
        c.add(gates.RY(n - 1, theta=np.pi * (r.symbols[0] + 2 * r.symbols[1]) / 5))
        c.add(gates.CNOT(1, n - 1))
        c.add(gates.M(0, 1, n - 1))
        return c

    np.random.seed(31)
    want, _ = OracleBackend().execute_repeated(circuit(), 60)
    np.random.seed(31)
    got = backend.execute_circuit(circuit(), nshots=60).samples()
    got = got.cpu().numpy() if isinstance(got, torch.Tensor) else np.asarray(got)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("effect", [False, True])
def test_measurement_result_parameters_density(backend, effect):
    """The reference's test_measurement_result_parameters (density matrices):
    X(0) then a collapse of qubit 0 conditions RX(pi/4) on qubit 1."""
    c = Circuit(4, density_matrix=True)
    if effect:
        c.add(gates.X(0))
    r = c.add(gates.M(0, collapse=True))
    c.add(gates.RX(1, theta=np.pi * r.symbols[0] / 4))
    t = Circuit(4, density_matrix=True)
    if effect:
        t.add(gates.X(0))
        t.add(gates.RX(1, theta=np.pi / 4))
    got = backend.execute_circuit(c, nshots=1).state().cpu().numpy()
    want = backend.execute_circuit(t).state().cpu().numpy()
    assert np.max(np.abs(got - want)) < 1e-14
