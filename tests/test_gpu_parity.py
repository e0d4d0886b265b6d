"""Parity of the HIP path with the CPU oracle (needs an MI355X).

Tolerances (stated per test):
* complex128 amplitudes: max |GPU - oracle| <= 1e-12;
* complex64 amplitudes vs the complex128 oracle: <= 5e-6 (single precision);
* probabilities of all qubits: bitwise equal to numpy's ``np.abs(psi)**2``;
* cumulative sums: bitwise equal to numpy's sequential ``cumsum``;
* sampled shot indices given the same seed: bitwise equal to ``np.random.choice``;
* marginals / expectations (FP64 reductions in another order): <= 1e-12.
"""

import math
import threading
from fractions import Fraction

import numpy as np
import pytest
import torch

from oracle import OracleBackend
from qibo_amd import Circuit, config, gates, hamiltonians, models

pytestmark = pytest.mark.gpu

TOL128 = 1e-12
TOL64 = 5e-6


def rand_state(n, rng, dtype=np.complex128):
    v = rng.normal(size=2**n) + 1j * rng.normal(size=2**n)
    return (v / np.linalg.norm(v)).astype(dtype)


def to_np(t):
    return t.detach().cpu().numpy()


GATES = [
    lambda n: gates.H(0), lambda n: gates.H(n - 1), lambda n: gates.X(n // 2), lambda n: gates.Y(1),
    lambda n: gates.Z(n - 2), lambda n: gates.S(2), lambda n: gates.T(n - 1), lambda n: gates.SX(3),
    lambda n: gates.RX(n - 1, 0.37), lambda n: gates.RY(0, 1.91), lambda n: gates.RZ(n - 3, -2.2),
    lambda n: gates.U1(1, 0.7), lambda n: gates.U2(n - 1, 0.3, 0.9), lambda n: gates.U3(2, 0.1, -0.4, 2.0),
    lambda n: gates.CNOT(0, n - 1), lambda n: gates.CNOT(n - 1, 0), lambda n: gates.CZ(1, n - 2),
    lambda n: gates.CY(2, 3), lambda n: gates.CRX(n - 1, 1, 0.5), lambda n: gates.CRZ(0, 2, -0.3),
    lambda n: gates.CU1(n - 2, n - 1, 0.8), lambda n: gates.SWAP(0, n - 1), lambda n: gates.SWAP(n - 2, n - 1),
    lambda n: gates.iSWAP(1, 4), lambda n: gates.fSim(n - 1, 2, 0.3, 1.1), lambda n: gates.RXX(0, 3, 0.6),
    lambda n: gates.RZZ(2, n - 1, 0.4), lambda n: gates.TOFFOLI(0, n - 1, 3), lambda n: gates.CCZ(1, 2, n - 1),
    lambda n: gates.RX(3, 0.8).controlled_by(0, n - 1), lambda n: gates.SWAP(1, n - 1).controlled_by(2),
    lambda n: gates.H(n - 1).controlled_by(0, 1, 2), lambda n: gates.U1(0, 0.2).controlled_by(n - 1),
    lambda n: gates.Unitary(np.linalg.qr(np.arange(64).reshape(8, 8) * 1j + np.eye(8))[0], 4, 0, n - 1),
    lambda n: gates.Unitary(np.linalg.qr(np.arange(256).reshape(16, 16) * 1j + np.eye(16))[0], 5, 0, n - 1, 2),
]


@pytest.mark.parametrize("make", GATES)
def test_single_gate_c128(backend, make, rng):
    n = 9
    psi = rand_state(n, rng)
    g = make(n)
    got = to_np(backend.apply_gate(g, backend.cast(psi), n))
    want = OracleBackend().apply_gate(g, psi.copy(), n)
    assert np.max(np.abs(got - want)) <= TOL128


@pytest.mark.parametrize("make", GATES)
def test_single_gate_c64(backend64, make, rng):
    n = 9
    psi = rand_state(n, rng)
    g = make(n)
    got = to_np(backend64.apply_gate(g, backend64.cast(psi.astype(np.complex64)), n))
    want = OracleBackend().apply_gate(g, psi.copy(), n)
    assert np.max(np.abs(got - want)) <= TOL64


def test_zero_state(backend):
    s = to_np(backend.zero_state(6))
    want = np.zeros(64, dtype=complex)
    want[0] = 1
    np.testing.assert_array_equal(s, want)


@pytest.mark.parametrize("n,layers", [(12, 6), (17, 3)])
def test_layered_circuit_c128(backend, n, layers):
    c = models.random_layered_circuit(n, layers, seed=n, measure=False)
    got = to_np(backend.execute_circuit(c).state())
    want = OracleBackend().execute(c)
    assert np.max(np.abs(got - want)) <= TOL128


@pytest.mark.parametrize("n", [10, 13])
def test_qft_c128(backend, n, rng):
    psi = rand_state(n, rng)
    c = models.QFT(n)
    got = to_np(backend.execute_circuit(c, initial_state=psi).state())
    dim = 2**n
    dft = np.exp(2j * np.pi * np.outer(np.arange(dim), np.arange(dim)) / dim) / np.sqrt(dim)
    assert np.max(np.abs(got - dft @ psi)) <= TOL128
    assert np.max(np.abs(got - OracleBackend().execute(c, psi))) <= TOL128


@pytest.mark.parametrize("maxq", [2, 3, 4, 5])
def test_fused_circuit_c128(backend, maxq):
    n = 11
    c = models.random_layered_circuit(n, 4, seed=7, measure=False)
    c.add(gates.CU1(0, n - 1, 0.3))
    c.add(gates.TOFFOLI(2, 5, 9))
    want = OracleBackend().execute(c)
    got = to_np(backend.execute_circuit(c.fuse(max_qubits=maxq)).state())
    assert np.max(np.abs(got - want)) <= TOL128


@pytest.mark.parametrize("n", [11, 14])
@pytest.mark.parametrize("maxq", [4, 5])
def test_fused_circuit_c64(backend64, maxq, n):
    """The layered circuit grouped by fuse(4) / fuse(5): qibo's fusion forms 2-
    and 3-qubit groups here (dense records in blocked passes, complex64)."""
    c = models.random_layered_circuit(n, 4, seed=8, measure=False)
    want = OracleBackend().execute(c)
    got = to_np(backend64.execute_circuit(c.fuse(max_qubits=maxq)).state())
    assert np.max(np.abs(got - want)) <= 2e-5


@pytest.mark.parametrize("k", [4, 5])
def test_mfma_block_random_unitary(backend64, k, rng):
    """A random dense block on scattered targets (including the lowest bit), MFMA path."""
    n = 15
    a = rng.normal(size=(2**k, 2**k)) + 1j * rng.normal(size=(2**k, 2**k))
    u, _ = np.linalg.qr(a)
    targets = [14, 3, 9, 0, 7][:k]
    psi = rand_state(n, rng)
    g = gates.Unitary(u, *targets)
    got = to_np(backend64.apply_gate(g, backend64.cast(psi.astype(np.complex64)), n))
    want = OracleBackend().apply_gate(g, psi.copy(), n)
    assert np.max(np.abs(got - want)) <= 2e-6


def _formula_abs2(z):
    re, im = abs(float(z.real)), abs(float(z.imag))
    lg, sm = max(re, im), min(re, im)
    r = 0.0 if lg == 0 else sm / lg
    h = math.sqrt(float(Fraction(r) * Fraction(r) + 1)) * lg
    return h * h


def test_probabilities_all_bitwise(backend, rng):
    n = 16
    psi = rand_state(n, rng) * 10.0 ** rng.integers(-3, 1, size=2**n)
    psi[:4] = [0, 1j * 1e-200, -3.0, 1e-170]
    p = to_np(backend.calculate_probabilities(backend.cast(psi), list(range(n)), n))
    assert p.dtype == np.float64
    sample = rng.integers(0, 2**n, size=500)
    assert all(p[i] == _formula_abs2(psi[i]) for i in sample)
    np.testing.assert_array_equal(p, np.abs(psi) ** 2)


def test_probabilities_all_bitwise_c64(backend64, rng):
    n = 14
    psi = rand_state(n, rng, np.complex64)
    p = to_np(backend64.calculate_probabilities(backend64.cast(psi, dtype="complex64"), list(range(n)), n))
    assert p.dtype == np.float32
    np.testing.assert_array_equal(p, np.abs(psi) ** 2)


@pytest.mark.parametrize("qubits", [[3, 0], [5], [0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13],
                                    [13, 12, 11, 10, 9, 8, 7, 6, 5, 4, 3, 2, 1, 0], [7, 2, 11, 4, 0, 9],
                                    list(range(1, 14))])
def test_probabilities_marginal(backend, rng, qubits):
    n = 14
    psi = rand_state(n, rng)
    got = to_np(backend.calculate_probabilities(backend.cast(psi), qubits, n))
    want = OracleBackend().calculate_probabilities(psi, qubits, n)
    assert np.max(np.abs(got - want)) <= 1e-12


def _prob_vectors(rng):
    out = {}
    for n in (1, 3, 4095, 4096, 4097, 70001, 2**20, 2**22 + 3):
        p = rng.random(n)
        out[f"uniform-{n}"] = p / p.sum()
    p = rng.exponential(size=2**21)
    p[rng.random(2**21) < 0.3] = 0.0
    out["sparse"] = p / p.sum()
    p = 0.5 ** np.arange(1, 1100, dtype=np.float64)
    out["geometric"] = p
    p = np.full(2**20, 2.0**-20)
    out["exact-uniform"] = p
    p = rng.random(2**20) * 10.0 ** rng.integers(-300, 0, size=2**20)
    out["wide-range"] = p / p.sum()
    p = np.zeros(50000)
    p[-1] = 1.0
    out["last-only"] = p
    return out


def test_exact_cumsum_bitwise(backend, rng):
    for name, p in _prob_vectors(rng).items():
        t = torch.as_tensor(p, device=backend.torch_device)
        got = to_np(backend.exact_cdf(t))
        np.testing.assert_array_equal(got, np.cumsum(p), err_msg=name)


def test_exact_cumsum_float32(backend, rng):
    p = rng.random(2**20 + 17).astype(np.float32)
    p /= p.sum()
    got = to_np(backend.exact_cdf(torch.as_tensor(p, device=backend.torch_device)))
    np.testing.assert_array_equal(got, np.cumsum(p.astype(np.float64)))


@pytest.mark.parametrize("n", [3, 12, 22])
def test_sample_shots_bit_exact(backend, rng, n):
    psi = rand_state(n, rng)
    probs = np.abs(psi) ** 2
    np.random.seed(1000 + n)
    want = OracleBackend().sample_shots(probs, 100000)
    np.random.seed(1000 + n)
    got = to_np(backend.sample_shots(backend.cast(probs, dtype="float64"), 100000))
    np.testing.assert_array_equal(got, want)


def test_sample_shots_from_state_bit_exact(backend, rng):
    """Probabilities computed on the GPU from the same state feed the same draws."""
    n = 20
    psi = rand_state(n, rng)
    probs_gpu = backend.calculate_probabilities(backend.cast(psi), list(range(n)), n)
    probs_cpu = OracleBackend().calculate_probabilities(psi, list(range(n)), n)
    np.random.seed(5)
    want = OracleBackend().sample_shots(probs_cpu, 50000)
    np.random.seed(5)
    got = to_np(backend.sample_shots(probs_gpu, 50000))
    np.testing.assert_array_equal(got, want)


def test_seeded_golden_frequencies(backend, golden):
    g = golden["probabilistic_measurement"]
    want = {int(k): v for k, v in g["frequencies"].items()}
    for use_samples in (True, False):
        c = Circuit(4)
        c.add([gates.H(0), gates.H(1)])
        c.add(gates.M(0, 1))
        result = backend.execute_circuit(c, nshots=g["nshots"])
        backend.set_seed(g["seed"])
        if use_samples:
            _ = result.samples()
        assert sum(result.frequencies().values()) == 1000
        assert dict(result.frequencies(binary=False)) == want
    g = golden["unbalanced_probabilistic_measurement"]
    want = {int(k): v for k, v in g["frequencies"].items()}
    psi = np.array(g["initial_state_unnormalised"]) / g["norm"]
    for use_samples in (True, False):
        c = Circuit(2)
        c.add(gates.M(0, 1))
        result = backend.execute_circuit(c, initial_state=psi.copy(), nshots=g["nshots"])
        backend.set_seed(g["seed"])
        if use_samples:
            _ = result.samples()
        assert dict(result.frequencies(binary=False)) == want


def test_result_api_registers(backend, rng):
    c = Circuit(5)
    c.add(gates.H(i) for i in (0, 2, 4))
    c.add(gates.CNOT(0, 3))  # Bell pair on (0, 3)
    c.add(gates.M(0, 2, register_name="a"))
    c.add(gates.M(3, 4, register_name="b"))
    res = backend.execute_circuit(c, nshots=2000)
    backend.set_seed(3)
    s = to_np(res.samples())
    assert s.shape == (2000, 4)
    np.testing.assert_array_equal(s[:, 0], s[:, 2])  # columns: q0, q2, q3, q4
    regs = res.samples(registers=True)
    np.testing.assert_array_equal(to_np(regs["a"]), s[:, :2])
    fr = res.frequencies(registers=True)
    assert sum(fr["b"].values()) == 2000
    probs = to_np(res.probabilities([0, 3]))
    np.testing.assert_allclose(probs, [0.5, 0, 0, 0.5], atol=1e-15)


def test_expectation_tfim(backend, rng):
    n = 12
    psi = rand_state(n, rng)
    ham = hamiltonians.TFIM(n, h=1.0, closed_boundary=False)
    extra = [(0.37, "XYZ", (1, 3, 5)), (-0.2, "YY", (0, 2)), (0.11, "ZXXZ", (0, 4, 7, 11))]
    terms = ham.terms + extra
    h2 = hamiltonians.SymbolicHamiltonian(terms, n, backend=backend)
    got = h2.expectation_from_state(backend.cast(psi))
    coeffs, strings, qubits = h2.simple_terms
    want = OracleBackend().exp_value_observable_symbolic(psi, strings, qubits, coeffs, n)
    assert abs(got - want) <= 1e-12


def test_expectation_of_circuit(backend):
    n = 10
    c = models.hardware_efficient_ansatz(n, 2, params=np.linspace(0.1, 2.0, 4 * n))
    ham = hamiltonians.TFIM(n, h=1.0, closed_boundary=False, backend=backend)
    got = ham.expectation(c)
    coeffs, strings, qubits = ham.simple_terms
    want = OracleBackend().expectation(c, (coeffs, strings, qubits), n)
    assert abs(got - want) <= 1e-12


def test_large_state_round_trip(backend):
    """n = 26: U then U^dagger returns |0...0>; norm kept (size-independent checks)."""
    n = 26
    c = models.random_layered_circuit(n, 2, seed=3, measure=False)
    state = backend.execute_circuit(c).state()
    p = backend.calculate_probabilities(state, list(range(n)), n)
    assert abs(float(p.sum()) - 1.0) < 1e-10
    back = backend.execute_circuit(c.invert(), initial_state=state).state()
    amp0 = complex(back[0].item())
    assert abs(amp0 - 1.0) < 1e-10
    del state, back, p
    torch.cuda.empty_cache()


def test_concurrent_threads(backend):
    """Independent circuits from several threads (qibo's parallel.py pattern)."""
    circuits = [models.random_layered_circuit(12, 3, seed=s, measure=False) for s in range(4)]
    want = [OracleBackend().execute(c) for c in circuits]
    results = [None] * 4

    def work(i):
        with torch.cuda.stream(torch.cuda.Stream(device=backend.torch_device)):
            results[i] = to_np(backend.execute_circuit(circuits[i]).state())
            torch.cuda.current_stream().synchronize()

    threads = [threading.Thread(target=work, args=(i,)) for i in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    for r, w in zip(results, want):
        assert np.max(np.abs(r - w)) <= TOL128


def test_errors(backend):
    with pytest.raises(ValueError):
        backend.apply_gate(gates.H(9), backend.zero_state(3), 3)
    config.set_hbm_budget(1 << 20)
    try:
        with pytest.raises(RuntimeError, match="does not fit"):
            backend.execute_circuit(models.QFT(20))
    finally:
        config.set_hbm_budget(None)
    with pytest.raises(ValueError, match="sum to 1"):
        backend.sample_shots(np.array([0.5, 0.6]), 10)


@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_probability_checks(backend, dtype):
    """The checks np.random.choice makes before drawing, on device vectors of
    2^20 entries: NaN, a negative entry, a sum off by more than sqrt(eps);
    no uniform is drawn when a check fails."""
    n = 1 << 20
    p = np.full(n, 1.0 / n)
    bad = [(np.nan, "NaN"), (-1e-9, "non-negative")]
    for value, msg in bad:
        q = p.copy()
        q[n // 3] = value
        state = np.random.get_state()
        with pytest.raises(ValueError, match=msg):
            backend.sample_shots(torch.as_tensor(q.astype(dtype), device=backend.torch_device), 5)
        after = np.random.get_state()
        assert state[2] == after[2] and np.array_equal(state[1], after[1])
    with pytest.raises(ValueError, match="sum to 1"):
        backend.sample_shots(torch.as_tensor((p * 1.001).astype(dtype), device=backend.torch_device), 5)
    # a sum within tolerance passes
    got = backend.sample_shots(torch.as_tensor((p * (1 + 1e-9)).astype(dtype), device=backend.torch_device), 5)
    assert got.shape[0] == 5


@pytest.mark.parametrize("m", [0, 3, 7, 8, 13, 14, 20, 22])
def test_np_sum_bitwise(backend, rng, m):
    p = rng.random(2**m) * 10.0 ** rng.integers(-6, 6, size=2**m)
    got = float(backend.np_sum(torch.as_tensor(p, device=backend.torch_device)).item())
    assert got == float(np.sum(p))
    p32 = p.astype(np.float32)
    got32 = backend.np_sum(torch.as_tensor(p32, device=backend.torch_device)).cpu().numpy()[0]
    assert got32 == np.sum(p32)


@pytest.mark.parametrize("m,nshots", [(10, 3 * 2**18 + 17), (20, 2**18 + 5)])
def test_sample_frequencies_bit_exact(backend, rng, m, nshots):
    """Counts equal the oracle's (numpy) for the same seed: normalisation by
    numpy's summation order, the same batching of uniforms, exact cumsum."""
    p = rng.random(2**m) ** 4
    p /= p.sum()
    np.random.seed(77)
    want = OracleBackend().sample_frequencies(p, nshots)
    np.random.seed(77)
    got = backend.sample_frequencies(torch.as_tensor(p, device=backend.torch_device), nshots)
    assert dict(got) == dict(want)
