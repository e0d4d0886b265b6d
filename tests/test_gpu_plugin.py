"""The backend driven the way qibo drives a plugin (needs an MI355X).

* Circuits built from stand-ins of qibo's own classes (``tests/standins.py``,
  none a ``qibo_amd`` subclass), including the headline form: a layered
  circuit followed by ``M(all)`` and shots. Tolerances: amplitudes 1e-12 vs
  the oracle; sampled shots bitwise equal to numpy's ``np.random.choice``
  under the same seed.
* The program cache under threads (qibo's parallel.py calls execute_circuit
  from several threads): more distinct circuits than the cache holds, from
  four threads, each state within 1e-12 of the oracle.
* Blocked passes with two low tile bits (``blocking_low_bits = 2``) against the
  oracle (1e-12), and the cache keyed by that setting.
"""

import threading

import numpy as np
import pytest
import torch

from oracle import OracleBackend
from qibo_amd import models
from tests import standins as si

pytestmark = pytest.mark.gpu

TOL128 = 1e-12


def to_np(t):
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("n,layers", [(10, 3), (14, 4)])
def test_standin_layered_measured_bit_exact(backend, n, layers):
    c = si.layered(n, layers, seed=n)
    own = models.random_layered_circuit(n, layers, seed=n, measure=False)
    want = OracleBackend().execute(own)
    backend.set_seed(123)
    result = backend.execute_circuit(c, nshots=3000)
    assert c._final_state is result
    got = to_np(result.state())
    assert np.max(np.abs(got - want)) <= TOL128
    samples = to_np(result.samples())
    assert samples.shape == (3000, n)
    probs = OracleBackend().calculate_probabilities(want, list(range(n)), n)
    np.random.seed(123)
    idx = OracleBackend().sample_shots(probs, 3000)
    bits = (idx[:, None] >> np.arange(n - 1, -1, -1)) & 1
    np.testing.assert_array_equal(samples, bits)
    # the stand-in measurement gate got its register's samples
    np.testing.assert_array_equal(to_np(c.measurements[0].result.samples()), bits)


def test_standin_fused_groups(backend, rng):
    """A stand-in FusedGate (3 and 5 qubits) through matrix_fused and the
    block kernels, against the oracle applying its gates one by one."""
    n = 13
    psi = rng.normal(size=2**n) + 1j * rng.normal(size=2**n)
    psi /= np.linalg.norm(psi)
    c = si.Circuit(n)
    own = models.Circuit(n)
    from qibo_amd import gates as own_gates

    for qubits, (si_gates, own_list) in [
        ((1, 4, 9), ([si.H(1), si.CZ(1, 4), si.RY(9, 0.4), si.CNOT(9, 4)],
                     [own_gates.H(1), own_gates.CZ(1, 4), own_gates.RY(9, 0.4), own_gates.CNOT(9, 4)])),
        ((0, 3, 6, 11, 12), ([si.H(0), si.RY(3, 1.1), si.CZ(0, 6), si.RX(11, 0.3), si.CNOT(12, 3), si.RZ(6, 2.0)],
                             [own_gates.H(0), own_gates.RY(3, 1.1), own_gates.CZ(0, 6), own_gates.RX(11, 0.3),
                              own_gates.CNOT(12, 3), own_gates.RZ(6, 2.0)])),
    ]:
        fg = si.FusedGate(*qubits)
        for g in si_gates:
            fg.append(g)
        c.add(fg)
        for g in own_list:
            own.add(g)
    got = to_np(backend.execute_circuit(c, initial_state=psi.copy()).state())
    want = OracleBackend().execute(own, psi.copy())
    assert np.max(np.abs(got - want)) <= TOL128


def test_program_cache_under_threads(backend):
    """20 distinct circuits (more than the 16 cached queues) run twice each from
    4 threads: hits, misses and evictions interleave; every state matches."""
    circuits = [models.random_layered_circuit(12 + (s % 3), 3, seed=100 + s, measure=False) for s in range(20)]
    want = [OracleBackend().execute(c) for c in circuits]
    errors = []

    def work(tid):
        try:
            with torch.cuda.stream(torch.cuda.Stream(device=backend.torch_device)):
                for rep in range(2):
                    for i in range(tid, 20, 4):
                        got = to_np(backend.execute_circuit(circuits[i]).state())
                        err = float(np.max(np.abs(got - want[i])))
                        if err > TOL128:
                            errors.append((tid, rep, i, err))
        except Exception as exc:  # surfaced below
            errors.append((tid, repr(exc)))

    threads = [threading.Thread(target=work, args=(t,)) for t in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors


@pytest.mark.parametrize("dtype", ["complex128", "complex64"])
def test_low_bits_two(backend, backend64, dtype):
    be = backend if dtype == "complex128" else backend64
    tol = TOL128 if dtype == "complex128" else 5e-6
    n = 14
    c = models.random_layered_circuit(n, 6, seed=21, measure=False)
    want = OracleBackend().execute(c)
    old = be.blocking_low_bits
    try:
        be.blocking_low_bits = 2
        got2 = to_np(be.execute_circuit(c).state())
        again = to_np(be.execute_circuit(c).state())  # replay of the low = 2 programs
        be.blocking_low_bits = old
        default = to_np(be.execute_circuit(c).state())  # not the low = 2 programs
    finally:
        be.blocking_low_bits = old
    assert np.max(np.abs(got2 - want)) <= tol
    np.testing.assert_array_equal(again, got2)
    assert np.max(np.abs(default - want)) <= tol
