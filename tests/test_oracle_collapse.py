"""The oracle's collapsing measurement and repeated execution, pinned on the CPU.

Checks that do not restate the oracle's code: the collapsed state is the
normalised projection computed index by index; a collapse followed by a
measurement of the same qubits repeats the collapse outcome
(tests/test_measurements_collapse.py:10-25 of the reference); and the draws
consume numpy's global generator in the reference's order (one uniform per
collapsing measurement, then one for the final measurement, per shot).
"""

import numpy as np
import pytest

from oracle import OracleBackend
from qibo_amd import Circuit, gates


def _random_state(n, rng):
    psi = rng.normal(size=2**n) + 1j * rng.normal(size=2**n)
    return psi / np.linalg.norm(psi)


@pytest.mark.parametrize("n,qubits,shot", [(3, [1], 1), (4, [0, 2], 2), (5, [0, 3, 4], 5), (6, [1, 3], 0)])
def test_collapse_is_normalised_projection(n, qubits, shot, rng):
    psi = _random_state(n, rng)
    got = OracleBackend().collapse_state(psi.copy(), qubits, np.array([shot]), n)
    want = np.zeros_like(psi)
    keep = []
    for i in range(2**n):
        bits = [(i >> (n - 1 - q)) & 1 for q in qubits]
        if int("".join(map(str, bits)), 2) == shot:
            keep.append(i)
    p = sum(abs(psi[i]) ** 2 for i in keep)
    for i in keep:
        want[i] = psi[i] / np.sqrt(p)
    assert np.max(np.abs(got - want)) < 1e-14
    assert abs(np.linalg.norm(got) - 1) < 1e-14


@pytest.mark.parametrize("n,targets", [(2, [1]), (3, [1]), (4, [1, 3]), (5, [0, 3, 4]), (6, [1, 3]), (4, [0, 2])])
def test_collapse_then_measure_repeats_outcome(n, targets, rng):
    c = Circuit(n)
    for q in range(0, n, 2):
        c.add(gates.H(q))
    c.add(gates.M(*targets, collapse=True))
    c.add(gates.M(*targets))
    np.random.seed(3)
    final, (collapsed,) = OracleBackend().execute_repeated(c, 20, _random_state(n, rng))
    assert final.shape == (20, len(targets)) and collapsed.shape == (20, len(targets))
    order = np.argsort(targets)  # the collapse records sorted qubits, the measurement its own order
    assert np.array_equal(final[:, order], collapsed)


def test_repeated_execution_draw_order():
    """H(0), M(0, collapse), M(0): per shot, outcome = (u >= 1/2) for the first of
    two uniforms; the second is consumed by the final measurement."""
    c = Circuit(2)
    c.add(gates.H(0))
    c.add(gates.M(0, collapse=True))
    c.add(gates.M(0, 1))
    np.random.seed(11)
    final, (collapsed,) = OracleBackend().execute_repeated(c, 12)
    np.random.seed(11)
    u = np.random.random_sample(24)
    expect = (u[0::2] >= 0.5).astype(np.int64)
    assert np.array_equal(collapsed[:, 0], expect)
    assert np.array_equal(final[:, 0], expect) and not final[:, 1].any()


def test_measurement_conditioned_flip():
    """RX(pi * outcome) on qubit 1 after collapsing qubit 0 of |+0>: qubit 1
    copies qubit 0 in every shot (the feedback of
    tests/test_measurements_collapse.py:100-123 of the reference)."""
    c = Circuit(3)
    c.add(gates.H(0))
    r = c.add(gates.M(0, collapse=True))
    c.add(gates.RX(1, theta=np.pi * r.symbols[0]))
    c.add(gates.M(0, 1))
    np.random.seed(5)
    samples, records = OracleBackend().execute_repeated(c, 40)
    assert samples.shape == (40, 2)
    assert np.array_equal(samples[:, 0], samples[:, 1])
    assert np.array_equal(records[0][:, 0], samples[:, 0])
    assert 0 < samples[:, 0].sum() < 40
