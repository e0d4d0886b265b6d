"""Blocked multi-gate passes on the GPU (needs an MI355X).

With circuit order kept (``blocking_ordered``) a blocked pass must give
bitwise the same state as the one-gate-per-pass kernels (same fused
multiply-adds in the same order). With commuting gates reordered (the
default) the rounding changes: tolerance 1e-14 absolute (complex128) / 1e-6
(complex64) against the per-gate state, on normalised random states. Both
match the oracle within 1e-12 (complex128) / 5e-6 (complex64).
"""

import numpy as np
import pytest
import torch

from oracle import OracleBackend
from qibo_amd import Circuit, gates, models

pytestmark = pytest.mark.gpu


def _random_circuit(n, ngates, seed):
    rng = np.random.default_rng(seed)
    c = Circuit(n)
    for _ in range(ngates):
        kind = rng.integers(0, 9)
        q = [int(x) for x in rng.choice(n, size=3, replace=False)]
        th = float(rng.uniform(0, 6.3))
        if kind == 0:
            c.add(gates.RX(q[0], th))
        elif kind == 1:
            c.add(gates.CNOT(q[0], q[1]))
        elif kind == 2:
            c.add(gates.SWAP(q[0], q[1]))
        elif kind == 3:
            c.add(gates.fSim(q[0], q[1], th, 0.7))
        elif kind == 4:
            c.add(gates.TOFFOLI(q[0], q[1], q[2]))
        elif kind == 5:
            c.add(gates.RY(q[0], th).controlled_by(q[1], q[2]))
        elif kind == 6:
            c.add(gates.CU1(q[0], q[1], th))
        elif kind == 7:
            c.add(gates.RZZ(q[0], q[1], th))
        else:
            c.add(gates.U3(q[0], th, 0.2, -0.5))
    return c


def _run(backend, circuit, psi, blocking, ordered=False):
    backend.blocking = blocking
    backend.blocking_ordered = ordered
    try:
        out = backend.execute_circuit(circuit, initial_state=psi).state().detach().cpu().numpy()
    finally:
        backend.blocking = True
        backend.blocking_ordered = False
    return out


CIRCUITS = [
    ("layered", lambda n: models.random_layered_circuit(n, 8, seed=n, measure=False)),
    ("qft", lambda n: models.QFT(n)),
    ("random", lambda n: _random_circuit(n, 300, seed=n)),
    ("hea", lambda n: models.hardware_efficient_ansatz(n, 4, params=np.linspace(0, 3, 8 * n))),
]


@pytest.mark.parametrize("n", [12, 15, 20])
@pytest.mark.parametrize("name,make", CIRCUITS)
def test_blocked_bitwise_equals_per_gate_c128(backend, name, make, n, rng):
    c = make(n)
    psi = rng.normal(size=2**n) + 1j * rng.normal(size=2**n)
    psi /= np.linalg.norm(psi)
    per_gate = _run(backend, c, psi, False)
    ordered = _run(backend, c, psi, True, ordered=True)
    np.testing.assert_array_equal(ordered, per_gate)
    blocked = _run(backend, c, psi, True)
    np.testing.assert_allclose(blocked, per_gate, rtol=0, atol=1e-14)
    if n <= 15:
        want = OracleBackend().execute(c, psi)
        assert np.max(np.abs(blocked - want)) <= 1e-12


@pytest.mark.parametrize("name,make", CIRCUITS)
def test_blocked_bitwise_equals_per_gate_c64(backend64, name, make, rng):
    n = 16
    c = make(n)
    psi = (rng.normal(size=2**n) + 1j * rng.normal(size=2**n))
    psi = (psi / np.linalg.norm(psi)).astype(np.complex64)
    per_gate = _run(backend64, c, psi, False)
    ordered = _run(backend64, c, psi, True, ordered=True)
    np.testing.assert_array_equal(ordered, per_gate)
    blocked = _run(backend64, c, psi, True)
    np.testing.assert_allclose(blocked, per_gate, rtol=0, atol=1e-6)
    want = OracleBackend().execute(c, psi.astype(np.complex128))
    assert np.max(np.abs(blocked - want)) <= 5e-6


def test_blocked_pass_count_and_norm(backend):
    n = 24
    c = models.random_layered_circuit(n, 20, seed=2, measure=False)
    backend.block_stats.clear()
    state = backend.execute_circuit(c).state()
    assert 0 < len(backend.block_stats) <= 20
    assert sum(backend.block_stats) == c.ngates
    p = backend.calculate_probabilities(state, list(range(n)), n)
    assert abs(float(p.sum()) - 1.0) < 1e-12
    del state, p
    torch.cuda.empty_cache()


@pytest.mark.parametrize("first", ["blocked", "single", "empty"])
def test_zero_state_produced_by_first_pass(backend, first):
    """From |0...0> the first blocked pass writes the state without reading it;
    a one-gate first launch, or no gate at all, fills the zero state first."""
    n = 14
    c = Circuit(n)
    if first == "single":
        u = np.linalg.qr(np.arange(256).reshape(16, 16) * 1j + np.eye(16))[0]
        c.add(gates.Unitary(u, 0, 5, 9, 13))
    if first != "empty":
        for q in range(n):
            c.add(gates.RY(q, 0.1 * (q + 1)))
        c.add(gates.CZ(0, 1))
    c.add(gates.M(*range(n)))
    got = backend.execute_circuit(c, nshots=10).state().detach().cpu().numpy()
    want = OracleBackend().execute(c)
    assert np.max(np.abs(got - want)) <= 1e-12


@pytest.mark.parametrize("first", ["blocked", "single"])
def test_program_cache_replay(backend, first, rng):
    """A second execution of an identical gate queue replays the compiled
    passes (bitwise the same state); a changed parameter compiles anew."""
    n = 16
    c = models.random_layered_circuit(n, 6, seed=5, measure=False)
    if first == "single":
        u = np.linalg.qr(rng.normal(size=(32, 32)) + 1j * rng.normal(size=(32, 32)))[0]
        c0 = Circuit(n)
        c0.add(gates.Unitary(u, 0, 3, 6, 9, 12))
        c = c0 + c
    backend._programs.clear()
    hits = backend.program_cache_hits
    a = backend.execute_circuit(c).state().detach().cpu().numpy()
    b = backend.execute_circuit(c).state().detach().cpu().numpy()
    assert backend.program_cache_hits == hits + 1
    assert np.array_equal(a, b)
    want = OracleBackend().execute(c)
    assert np.max(np.abs(a - want)) <= 1e-12
    # with an initial state the key differs (no from-zero first pass)
    psi = rng.normal(size=2**n) + 1j * rng.normal(size=2**n)
    psi /= np.linalg.norm(psi)
    s1 = backend.execute_circuit(c, initial_state=psi).state().detach().cpu().numpy()
    s2 = backend.execute_circuit(c, initial_state=psi).state().detach().cpu().numpy()
    assert backend.program_cache_hits == hits + 2
    assert np.array_equal(s1, s2)
    assert np.max(np.abs(s1 - OracleBackend().execute(c, psi))) <= 1e-12
    # a changed parameter is a different program
    params = [g for g in c.queue if isinstance(g, gates.RY)]
    params[3].parameters = params[3].parameters[0] + 0.25
    d = backend.execute_circuit(c).state().detach().cpu().numpy()
    assert backend.program_cache_hits == hits + 2
    assert np.max(np.abs(d - OracleBackend().execute(c))) <= 1e-12
    assert np.max(np.abs(d - a)) > 1e-3
