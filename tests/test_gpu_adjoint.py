"""Adjoint gradients on the GPU (needs an MI355X).

Checked against: the reference's gradient goldens (tests/test_derivative.py:31
of qibo), the oracle's parameter-shift rule (exact for the rotation gates
used, tolerance 1e-10) and central finite differences with eps = 1e-4 in
complex128 (truncation error O(eps^2), tolerance 1e-6).
"""

import numpy as np
import pytest
import torch

from oracle import OracleBackend
from qibo_amd import Circuit, gates, hamiltonians, models
from qibo_amd.adjoint import AdjointExpectation, expectation_and_gradient

pytestmark = pytest.mark.gpu


def _dummy(params):
    c = Circuit(1)
    c.add(gates.RX(0, params[0]))
    c.add(gates.RY(0, params[1]))
    c.add(gates.RZ(0, params[2]))
    c.add(gates.M(0))
    return c


def test_golden_gradients(backend, golden):
    g = golden["derivative_dummy_circuit"]
    ham = hamiltonians.SymbolicHamiltonian([(1.0, "Z", (0,))], 1, backend=backend)
    e, grad = expectation_and_gradient(backend, _dummy(g["params"]), ham)
    np.testing.assert_allclose(grad.cpu().numpy(), g["grads_scale_1"], atol=1e-8)
    a, b, _ = g["params"]
    assert abs(float(e) - np.cos(a) * np.cos(b)) < 1e-14


@pytest.mark.parametrize("n,layers", [(6, 2), (13, 2)])
def test_adjoint_vs_parameter_shift(backend, n, layers):
    rng = np.random.default_rng(n)
    params = rng.uniform(0, 2 * np.pi, 2 * n * layers)
    c = models.hardware_efficient_ansatz(n, layers, params=params)
    ham = hamiltonians.TFIM(n, h=1.0, closed_boundary=False, backend=backend)
    e, grad = expectation_and_gradient(backend, c, ham)
    terms = ham.simple_terms
    ob = OracleBackend()
    idx = range(len(params)) if n <= 6 else rng.choice(len(params), 8, replace=False)
    for k in idx:
        psr = ob.parameter_shift(c, terms, int(k))
        assert abs(float(grad[k]) - psr) < 1e-10, k
    assert abs(float(e) - ob.expectation(c, terms, n)) < 1e-12


def test_adjoint_vs_central_fd_config_shape(backend):
    """n = 20, 10 layers RY+RZ + CNOT ladder, TFIM(h=1, open): a subset of the
    400 parameters against GPU central differences (eps = 1e-4, complex128)."""
    n, layers = 20, 10
    rng = np.random.default_rng(7)
    params = rng.uniform(0, 2 * np.pi, 2 * n * layers)
    c = models.hardware_efficient_ansatz(n, layers, params=params)
    ham = hamiltonians.TFIM(n, h=1.0, closed_boundary=False, backend=backend)
    _, grad = expectation_and_gradient(backend, c, ham)
    eps = 1e-4
    for k in rng.choice(len(params), 6, replace=False):
        p = params.copy()
        p[k] += eps
        c.set_parameters(p)
        fp = ham.expectation(c)
        p[k] -= 2 * eps
        c.set_parameters(p)
        fm = ham.expectation(c)
        fd = (fp - fm) / (2 * eps)
        assert abs(float(grad[k]) - fd) < 1e-6, (k, float(grad[k]), fd)
    c.set_parameters(params)


def test_autograd_batch(backend):
    n, layers, batch = 8, 2, 3
    rng = np.random.default_rng(11)
    c = models.hardware_efficient_ansatz(n, layers)
    ham = hamiltonians.TFIM(n, h=1.0, closed_boundary=False, backend=backend)
    params = torch.tensor(rng.uniform(0, 2 * np.pi, (batch, 2 * n * layers)), dtype=torch.float64,
                          requires_grad=True)
    energies = AdjointExpectation.apply(params, backend, c, ham)
    weights = torch.tensor([1.0, -2.0, 0.5], dtype=torch.float64)
    (energies * weights).sum().backward()
    ob = OracleBackend()
    terms = ham.simple_terms
    for b in range(batch):
        c.set_parameters(params[b].detach().numpy())
        assert abs(float(energies[b]) - ob.expectation(c, terms, n)) < 1e-12
        for k in (0, 5, 17):
            psr = ob.parameter_shift(c, terms, k)
            assert abs(float(params.grad[b, k]) - weights[b].item() * psr) < 1e-10


def test_adjoint_c64_close_to_c128(backend, backend64):
    n, layers = 12, 3
    rng = np.random.default_rng(3)
    params = rng.uniform(0, 2 * np.pi, 2 * n * layers)
    c = models.hardware_efficient_ansatz(n, layers, params=params)
    ham128 = hamiltonians.TFIM(n, h=1.0, closed_boundary=False, backend=backend)
    ham64 = hamiltonians.TFIM(n, h=1.0, closed_boundary=False, backend=backend64)
    e1, g1 = expectation_and_gradient(backend, c, ham128)
    e2, g2 = expectation_and_gradient(backend64, c, ham64)
    assert abs(float(e1) - float(e2)) < 1e-4
    assert float(torch.max(torch.abs(g1 - g2))) < 1e-3


@pytest.mark.parametrize("n,batch", [(12, 3), (16, 5)])
def test_blocked_batched_matches_per_set_sweep(backend, n, batch):
    """The batched blocked sweep (all sets in one launch per pass) against the
    per-set sweep: energies and gradients within 1e-11 (complex128)."""
    from qibo_amd.adjoint import blocked_expectation_and_gradient

    rng = np.random.default_rng(n)
    c = models.hardware_efficient_ansatz(n, 3)
    c.add(gates.CRX(0, n - 1, 0.3))
    c.add(gates.CU1(n - 2, 1, 0.4))
    c.add(gates.RZZ(2, 5, 0.5))
    c.add(gates.RX(3, 0.6).controlled_by(7))
    c.add(gates.CZ(4, 9))  # a sign flip record on both states
    c.add(gates.RY(9, 0.2))
    c.add(gates.CZ(0, n - 1))
    ham = hamiltonians.TFIM(n, h=0.7, closed_boundary=False, backend=backend)
    params = rng.uniform(0, 2 * np.pi, (batch, len(c.get_parameters())))
    e, g = blocked_expectation_and_gradient(backend, c, ham, params)
    for b in range(batch):
        c.set_parameters(params[b])
        e1, g1 = expectation_and_gradient(backend, c, ham)
        assert abs(float(e[b]) - float(e1)) < 1e-11
        assert float(torch.max(torch.abs(g[b] - g1))) < 1e-11


def test_blocked_batched_c64(backend64):
    from qibo_amd.adjoint import blocked_expectation_and_gradient

    n, batch = 13, 2
    rng = np.random.default_rng(1)
    c = models.hardware_efficient_ansatz(n, 2)
    ham = hamiltonians.TFIM(n, h=1.0, closed_boundary=False, backend=backend64)
    params = rng.uniform(0, 2 * np.pi, (batch, len(c.get_parameters())))
    e, g = blocked_expectation_and_gradient(backend64, c, ham, params)
    ob = OracleBackend()
    terms = ham.simple_terms
    for b in range(batch):
        c.set_parameters(params[b])
        assert abs(float(e[b]) - ob.expectation(c, terms, n)) < 1e-4
        for k in (0, 7, 30):
            assert abs(float(g[b, k]) - ob.parameter_shift(c, terms, k)) < 1e-3
