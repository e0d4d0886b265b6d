"""Adjoint-method gradients of expectation values, and their autograd wrapper.

For E(theta) = <psi|H|psi> with |psi> = U_G ... U_1 |0>, the adjoint sweep
(no stored intermediate states, gates un-applied in reverse):

    psi <- U_G ... U_1 |0>;   lambda <- H psi;   E = Re <psi|lambda>
    for j = G .. 1:
        psi <- U_j^dagger psi                     (psi is now U_{j-1}...U_1|0>)
        for each parameter theta of gate j:
            dE/dtheta += 2 Re <lambda| dU_j/dtheta |psi>
        lambda <- U_j^dagger lambda

Every step is a kernel of libqsim_mi355x.so: gate passes (the forward state
uses blocked passes), ``qsim_apply_pauli_sum`` for H psi, and
``qsim_inner_matrix`` for the derivative terms, all in FP64 accumulation.

The reference differentiates by re-execution: ``parameter_shift`` and
``finite_differences`` (src/qibo/derivative.py:7-226) run two full
simulations per parameter; this sweep costs about three state passes per
gate for the whole gradient.

Derivative matrices: for exp(-i theta G / 2) rotations (RX, RY, RZ, CRX, CRY,
CRZ, RXX, RYY, RZZ, RZX) dU/dtheta = -i/2 G U; for U1 / CU1 it is
diag(0, i e^{i theta}) on the controlled block; other parametrized gates use a
central difference of the host matrix (step 1e-7), documented as such.
"""

from typing import List, Sequence, Tuple

import numpy as np
import torch

from . import gates
from .backend import plan_matrix
from .config import raise_error

_X = np.array([[0, 1], [1, 0]], dtype=complex)
_Y = np.array([[0, -1j], [1j, 0]], dtype=complex)
_Z = np.array([[1, 0], [0, -1]], dtype=complex)
_GENERATORS = {
    "RX": _X, "RY": _Y, "RZ": _Z,
    "RXX": np.kron(_X, _X), "RYY": np.kron(_Y, _Y), "RZZ": np.kron(_Z, _Z), "RZX": np.kron(_Z, _X),
}
_CONTROLLED_GEN = {"CRX": _X, "CRY": _Y, "CRZ": _Z}


def _gate_qubits(gate) -> Tuple[Tuple[int, ...], Tuple[int, ...]]:
    """(targets, controls) that the gate's matrix acts on / is controlled by."""
    if gate.is_controlled_by:
        return tuple(gate.target_qubits), tuple(gate.control_qubits)
    return tuple(gate.qubits), ()


def derivative_matrices(backend, gate) -> List[np.ndarray]:
    """dU/dtheta_k for each parameter of ``gate`` (matrix over _gate_qubits)."""
    name = gate.__class__.__name__
    u = np.asarray(gate.matrix(backend), dtype=np.complex128)
    if name in _GENERATORS:
        return [-0.5j * _GENERATORS[name] @ u]
    if name in _CONTROLLED_GEN:
        d = np.zeros_like(u)
        d[2:, 2:] = -0.5j * _CONTROLLED_GEN[name] @ u[2:, 2:]
        return [d]
    if name == "U1":
        return [np.diag([0, 1j * u[1, 1]])]
    if name == "CU1":
        return [np.diag([0, 0, 0, 1j * u[3, 3]])]
    # other parametrized gates: central difference of the host matrix
    out = []
    params = list(gate.parameters)
    eps = 1e-7
    for k in range(len(params)):
        p_plus = list(params)
        p_minus = list(params)
        p_plus[k] = params[k] + eps
        p_minus[k] = params[k] - eps
        gate.parameters = p_plus
        up = np.asarray(gate.matrix(backend), dtype=np.complex128)
        gate.parameters = p_minus
        um = np.asarray(gate.matrix(backend), dtype=np.complex128)
        out.append((up - um) / (2 * eps))
    gate.parameters = params
    return out


def _dagger_op(backend, gate):
    targets, controls = _gate_qubits(gate)
    m = np.asarray(gate.matrix(backend))
    return plan_matrix(np.conj(m.T), targets, controls)


def expectation_and_gradient(backend, circuit, hamiltonian, initial_state=None):
    """(E, dE/dtheta) for the circuit's trainable parameters (flat order of
    ``circuit.get_parameters()``), by one forward pass and one adjoint sweep."""
    if circuit.density_matrix:
        raise_error(NotImplementedError, "adjoint gradients are for state vectors")
    n = circuit.nqubits
    xm, zm, ny, coeffs = hamiltonian.pauli_tables()
    queue = [g for g in circuit.queue if not isinstance(g, gates.M)]
    if initial_state is None:
        psi = backend.zero_state(n)
    else:
        psi = backend.cast(initial_state, copy=True)
    psi = backend.run_queue(queue, psi, n)
    lam = backend.apply_pauli_sum(psi, n, xm, zm, ny, coeffs)
    energy = backend.expectation_pauli(psi, n, xm, zm, ny, coeffs) + hamiltonian.constant
    slots = {}
    k = 0
    for g in circuit.trainable_gates:
        slots[id(g)] = (k, g.nparams)
        k += g.nparams
    grads = [None] * k
    for gate in reversed(queue):
        dop = _dagger_op(backend, gate)
        backend.apply_op(dop, psi, n)
        if id(gate) in slots and gate.trainable:
            first, count = slots[id(gate)]
            targets, controls = _gate_qubits(gate)
            for j, dm in enumerate(derivative_matrices(backend, gate)[:count]):
                grads[first + j] = 2.0 * backend.inner_matrix(lam, psi, n, targets, controls, dm)
        backend.apply_op(dop, lam, n)
    zero = torch.zeros((), dtype=torch.float64, device=psi.device)
    grad = torch.stack([g if g is not None else zero for g in grads]) if grads else zero.reshape(0)
    return energy, grad


def batched_expectation_and_gradient(backend, circuit, hamiltonian, params) -> Tuple[torch.Tensor, torch.Tensor]:
    """Energies [B] and gradients [B, P] for a batch of parameter sets [B, P]."""
    params = np.asarray(params.detach().cpu().numpy() if isinstance(params, torch.Tensor) else params,
                        dtype=np.float64)
    energies, grads = [], []
    for row in params:
        circuit.set_parameters(row)
        e, g = expectation_and_gradient(backend, circuit, hamiltonian)
        energies.append(e)
        grads.append(g)
    return torch.stack(energies), torch.stack(grads)


class AdjointExpectation(torch.autograd.Function):
    """E(params) = <psi(params)|H|psi(params)> for a batch of parameter sets.

    ``forward`` runs the circuit and the adjoint sweep for each row of
    ``params`` ([B, P] float64) and keeps the gradient; ``backward`` scales
    it by the incoming gradient. Usage::

        energies = AdjointExpectation.apply(params, backend, circuit, hamiltonian)
        energies.sum().backward()
    """

    @staticmethod
    def forward(ctx, params, backend, circuit, hamiltonian):
        energies, grads = batched_expectation_and_gradient(backend, circuit, hamiltonian, params)
        ctx.save_for_backward(grads.to(params.device))
        return energies.to(params.device, dtype=params.dtype)

    @staticmethod
    def backward(ctx, grad_out):
        (grads,) = ctx.saved_tensors
        return grad_out[:, None].to(grads.dtype) * grads, None, None, None
