"""Accuracy reports of results (needs an MI355X).

``qsim_compare_states`` against numpy on random states of every dtype pair
(overlap, norms and max-abs difference in FP64), and ``result.accuracy()`` of
complex64 circuits against the oracle's complex128 state: qibo's infidelity
1 - |<a|b>|^2 (src/qibo/quantum_info/metrics.py:258-291) and the max-abs
amplitude error, recomputed with numpy from the two states.
"""

import numpy as np
import pytest
import torch

from oracle import OracleBackend
from qibo_amd import accuracy, models

pytestmark = pytest.mark.gpu


def rand_state(n, rng):
    v = rng.normal(size=2**n) + 1j * rng.normal(size=2**n)
    return v / np.linalg.norm(v)


@pytest.mark.parametrize("da,db", [(torch.complex128, torch.complex128), (torch.complex64, torch.complex128),
                                   (torch.complex128, torch.complex64), (torch.complex64, torch.complex64)])
@pytest.mark.parametrize("n", [0, 5, 13, 21])
def test_compare_states_matches_numpy(backend, rng, da, db, n):
    a = rand_state(n, rng)
    b = a + 1e-3 * rand_state(n, rng)
    ta = torch.from_numpy(a).to(da).to(backend.torch_device)
    tb = torch.from_numpy(b).to(db).to(backend.torch_device)
    rep = accuracy.compare_states(backend, ta, tb)
    an, bn = ta.cpu().numpy().astype(np.complex128), tb.cpu().numpy().astype(np.complex128)
    ov = np.vdot(an, bn)
    assert abs(rep["infidelity"] - (1 - abs(ov) ** 2)) <= 1e-13
    want_norm = 1 - abs(ov) ** 2 / (np.vdot(an, an).real * np.vdot(bn, bn).real)
    assert abs(rep["infidelity_normalized"] - want_norm) <= 1e-13
    assert rep["max_abs_error"] == pytest.approx(np.max(np.abs(an - bn)), rel=1e-14, abs=0)
    assert abs(rep["norm_error"] - (np.vdot(an, an).real - 1)) <= 1e-13
    # a fixed reduction order: a second call is bitwise equal
    assert accuracy.compare_states(backend, ta, tb) == rep


def test_compare_states_rejects_mismatch(backend):
    a = torch.zeros(8, dtype=torch.complex128, device=backend.torch_device)
    with pytest.raises(ValueError):
        accuracy.compare_states(backend, a, a[:4])
    with pytest.raises(TypeError):
        accuracy.compare_states(backend, a.real, a.real)


@pytest.mark.parametrize("n,fuse", [(14, False), (14, True), (12, False)])
def test_complex64_result_accuracy_against_oracle(backend64, n, fuse):
    c = models.random_layered_circuit(n, 8, seed=5, measure=False)
    if fuse:
        c = c.fuse(max_qubits=5)
    result = backend64.execute_circuit(c)
    rep = result.accuracy()
    got = result.state(numpy=True).astype(np.complex128)
    want = OracleBackend().execute(models.random_layered_circuit(n, 8, seed=5, measure=False))
    ov = np.vdot(got, want)
    # the reference state is the GPU complex128 run (1e-12 of the oracle)
    assert abs(rep["infidelity"] - (1 - abs(ov) ** 2)) <= 1e-11
    assert abs(rep["max_abs_error"] - np.max(np.abs(got - want))) <= 1e-11
    assert 0 < rep["max_abs_error"] < 1e-5
    assert abs(rep["infidelity_normalized"]) < 1e-6
    # the circuit's own result is untouched by the reference run
    assert c._final_state is result


def test_report_accuracy_attaches_reports(backend64, backend):
    backend64.report_accuracy = True
    try:
        c = models.random_layered_circuit(12, 4, seed=9)
        res = backend64.execute_circuit(c, nshots=100)
        assert res.accuracy_report["max_abs_error"] < 1e-5
        assert res.accuracy_report["reference"].startswith("complex128")
        # measurements still sample through the complex64 backend's result
        assert res.samples().shape == (100, 12)
        assert c.measurements[0].result.backend is backend64
    finally:
        backend64.report_accuracy = False
    r128 = backend.execute_circuit(models.random_layered_circuit(12, 2, seed=9, measure=False))
    assert r128.accuracy()["max_abs_error"] == 0.0


def test_accuracy_with_initial_state(backend64, rng):
    n = 12
    psi = rand_state(n, rng)
    c = models.QFT(n)
    res = backend64.execute_circuit(c, initial_state=psi.astype(np.complex64))
    rep = res.accuracy()
    want = OracleBackend().execute(models.QFT(n), initial_state=psi.astype(np.complex64).astype(np.complex128))
    got = res.state(numpy=True).astype(np.complex128)
    assert abs(rep["max_abs_error"] - np.max(np.abs(got - want))) <= 1e-11
    assert rep["max_abs_error"] < 1e-5
