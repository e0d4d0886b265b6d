"""Derivative matrices of the adjoint sweep (host only, no GPU).

Every parametrized gate outside the generator families is differentiated by
the shift rule on its matrix, dU/dx = k (U(x + pi/2k) - U(x - pi/2k)) / 2,
exact when x enters the entries with the single frequency k. Checked against
a three-level Richardson extrapolation of central differences (error below
1e-11 for these trigonometric entries), against -i/2 G U for the rotation
families, and for the fallback (a gate class not in the frequency table)."""

import numpy as np
import pytest

from qibo_amd import adjoint, gates
from qibo_amd.backend import GatePlanner

B = GatePlanner("complex128")


def _richardson(gate, k, h=1e-3):
    p = list(gate.parameters)

    def u(x):
        q = list(p)
        q[k] = x
        gate.parameters = q
        return np.asarray(gate.matrix(B), dtype=complex)

    def d(step):
        return (u(p[k] + step) - u(p[k] - step)) / (2 * step)

    d1 = (4 * d(h / 2) - d(h)) / 3
    d2 = (4 * d(h / 4) - d(h / 2)) / 3
    gate.parameters = p
    return (16 * d2 - d1) / 15


CASES = [
    lambda: gates.PRX(0, 0.3, 1.1), lambda: gates.GPI(0, 0.7), lambda: gates.GPI2(0, -0.4),
    lambda: gates.U2(0, 0.3, -1.2), lambda: gates.U3(0, 0.5, 0.2, -0.7), lambda: gates.U1q(0, 0.9, 0.4),
    lambda: gates.CU2(0, 1, 0.3, 0.8), lambda: gates.CU3(0, 1, 0.5, 1.2, -0.3), lambda: gates.fSim(0, 1, 0.6, 1.3),
    lambda: gates.RXXYY(0, 1, 0.7), lambda: gates.MS(0, 1, 0.2, 0.9, 1.4), lambda: gates.GIVENS(0, 1, 0.8),
    lambda: gates.RBS(0, 1, -0.5), lambda: gates.DEUTSCH(0, 1, 2, 0.45),
    lambda: gates.RX(0, 0.3), lambda: gates.CRY(0, 1, -1.7), lambda: gates.RZX(0, 1, 2.2), lambda: gates.CU1(0, 1, 0.9),
]


@pytest.mark.parametrize("make", CASES)
def test_derivative_matrices_match_richardson(make):
    g = make()
    ds = adjoint.derivative_matrices(B, g)
    assert len(ds) == g.nparams
    for k, d in enumerate(ds):
        assert np.max(np.abs(d - _richardson(g, k))) < 1e-11


def test_shift_rule_exact_at_large_angles():
    # far from 0 the shift rule stays exact (no step-size error to grow)
    g = gates.U3(0, 7.9, -11.3, 25.1)
    for k, d in enumerate(adjoint.derivative_matrices(B, g)):
        assert np.max(np.abs(d - _richardson(g, k))) < 1e-11


def test_unlisted_gate_falls_back_to_central_difference(monkeypatch):
    g = gates.U3(0, 0.5, 0.2, -0.7)
    exact = adjoint.derivative_matrices(B, g)
    monkeypatch.delitem(adjoint._FREQUENCIES, "U3")
    approx = adjoint.derivative_matrices(B, g)
    for a, b in zip(approx, exact):
        assert np.max(np.abs(a - b)) < 1e-7
    assert list(g.parameters) == [0.5, 0.2, -0.7]  # parameters restored
