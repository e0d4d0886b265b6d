"""Dense-block (MFMA) kernels at their real shapes and the benchmark
configurations as written (needs an MI355X).

* Dense 4- and 5-qubit blocks (``gates.Unitary``) at n = 22 and 24, where the
  block kernels walk 1024 and 4096 tiles, more than one per resident
  workgroup, on several target sets (the top qubits, bit 0 of the index,
  scattered): complex64 within 2e-6 and complex128 within 1e-12 of the
  oracle.
* ``QFT(22).fuse(max_qubits=5)``, whose groups include 5-qubit FusedGates, in
  complex64 (2e-5) and complex128 (1e-12) against the oracle and the DFT.
* A 32-qubit complex64 state (32 GiB) through a sequence of dense 5-qubit
  blocks U_i and then their adjoints in reverse: back to |0...0> within 1e-4.
* Configuration 1 as written: QFT(15) from |0...0>, complex128, full state,
  against the oracle (1e-12) and the DFT's first column (1/sqrt(2^15) in
  every amplitude).
* Configuration 4 in complex64 and as a batch: the 20-qubit, 10-layer
  RY/RZ + CNOT-ladder ansatz (400 parameters), TFIM(20, h=1, open), 64
  parameter sets in one batched sweep; energies and a subset of gradients
  against the oracle's parameter-shift rule (exact for these gates):
  complex128 1e-10, complex64 1e-3.
"""

import math

import numpy as np
import pytest
import torch

from oracle import OracleBackend
from qibo_amd import gates, hamiltonians, models

pytestmark = pytest.mark.gpu


def to_np(t):
    return t.detach().cpu().numpy()


def rand_state(n, rng):
    v = rng.normal(size=2**n) + 1j * rng.normal(size=2**n)
    return v / np.linalg.norm(v)


def haar(k, rng):
    a = rng.normal(size=(2**k, 2**k)) + 1j * rng.normal(size=(2**k, 2**k))
    q, r = np.linalg.qr(a)
    return q * (np.diagonal(r) / np.abs(np.diagonal(r)))


TARGETS = {
    4: [(0, 1, 2, 3), (None, 0, 7, None), (5, None, 2, 11)],
    5: [(0, 1, 2, 3, 4), (None, 3, 0, 9, None), (6, None, 1, 12, 4)],
}


def _targets(spec, n):
    # None stands for the last qubit (bit 0 of the amplitude index) or the one before it
    out, low = [], [n - 1, n - 2]
    for t in spec:
        out.append(low.pop(0) if t is None else t)
    return out


@pytest.mark.parametrize("n", [22, 24])
@pytest.mark.parametrize("k", [4, 5])
@pytest.mark.parametrize("dtype", ["complex64", "complex128"])
def test_dense_block_multi_tile(backend, backend64, n, k, dtype):
    be = backend64 if dtype == "complex64" else backend
    tol = 2e-6 if dtype == "complex64" else 1e-12
    rng = np.random.default_rng(100 * n + k)
    psi = rand_state(n, rng)
    ob = OracleBackend()
    for spec in TARGETS[k]:
        targets = _targets(spec, n)
        g = gates.Unitary(haar(k, rng), *targets)
        state = be.cast(psi.astype(np.complex64 if dtype == "complex64" else np.complex128))
        got = to_np(be.apply_gate(g, state, n))
        want = ob.apply_gate(g, psi.copy(), n)
        err = float(np.max(np.abs(got - want)))
        assert err <= tol, (targets, err)
        del state


@pytest.mark.parametrize("dtype", ["complex64", "complex128"])
def test_qft22_fused5(backend, backend64, dtype):
    be = backend64 if dtype == "complex64" else backend
    tol = 2e-5 if dtype == "complex64" else 1e-12
    n = 22
    rng = np.random.default_rng(22)
    psi = rand_state(n, rng)
    c = models.QFT(n)
    fused = c.fuse(max_qubits=5)
    assert max(len(g.target_qubits) for g in fused.queue) == 5
    got = to_np(be.execute_circuit(fused, initial_state=psi.copy()).state())
    # QFT = DFT with e^{+2 pi i jk / N} / sqrt(N): numpy's inverse FFT times sqrt(N)
    dft = np.fft.ifft(psi) * math.sqrt(2**n)
    assert np.max(np.abs(got - dft)) <= tol
    want = OracleBackend().execute(c, psi.copy())
    assert np.max(np.abs(got - want)) <= tol


def test_dense_blocks_round_trip_32(backend64):
    n = 32
    rng = np.random.default_rng(32)
    blocks = [(haar(5, rng), t) for t in [(0, 1, 2, 3, 4), (31, 30, 29, 28, 27), (3, 17, 31, 8, 25),
                                         (12, 13, 14, 15, 16), (30, 0, 16, 5, 21)]]
    state = backend64._alloc_state(n)
    backend64._fill_zero(state, n)
    for u, t in blocks:
        backend64.apply_gate(gates.Unitary(u, *t), state, n)
    for u, t in reversed(blocks):
        backend64.apply_gate(gates.Unitary(np.conj(u.T), *t), state, n)
    amp0 = complex(state[0].item())
    norm = float(torch.linalg.vector_norm(state).item())
    del state
    torch.cuda.empty_cache()
    assert abs(amp0 - 1.0) < 1e-4
    assert abs(norm - 1.0) < 1e-4


def test_config1_qft15_from_zero(backend):
    n = 15
    c = models.QFT(n)
    got = to_np(backend.execute_circuit(c).state())
    want = OracleBackend().execute(c)
    assert np.max(np.abs(got - want)) <= 1e-12
    assert np.max(np.abs(got - 1 / math.sqrt(2**n))) <= 1e-12


def _config4(batch):
    n, layers = 20, 10
    rng = np.random.default_rng(4)
    params = rng.uniform(0, 2 * np.pi, (batch, 2 * n * layers))
    return n, models.hardware_efficient_ansatz(n, layers), params


def test_config4_batch64(backend, backend64):
    from qibo_amd.adjoint import blocked_expectation_and_gradient

    n, c, params = _config4(64)
    ham128 = hamiltonians.TFIM(n, h=1.0, closed_boundary=False, backend=backend)
    ham64 = hamiltonians.TFIM(n, h=1.0, closed_boundary=False, backend=backend64)
    e, g = blocked_expectation_and_gradient(backend, c, ham128, params)
    e64, g64 = blocked_expectation_and_gradient(backend64, c, ham64, params)
    assert tuple(g.shape) == (64, 400) and tuple(g64.shape) == (64, 400)
    ob = OracleBackend()
    terms = ham128.simple_terms
    for b, ks in ((0, (3, 217)), (63, (398,))):
        c.set_parameters(params[b])
        e_ref = ob.expectation(c, terms, n)
        assert abs(float(e[b]) - e_ref) < 1e-10
        assert abs(float(e64[b]) - e_ref) < 1e-3
        for k in ks:
            psr = ob.parameter_shift(c, terms, k)
            assert abs(float(g[b, k]) - psr) < 1e-10, (b, k)
            assert abs(float(g64[b, k]) - psr) < 1e-3, (b, k)
    # complex64 and complex128 gradients of every set agree to single precision
    assert float(torch.max(torch.abs(g.double() - g64.double()))) < 1e-2
