"""qsim_expval_pauli_noalloc through the C ABI (needs an MI355X): the entry
point that allocates nothing and keeps its partial sums in the library's own
device array gives bitwise the result of qsim_expval_pauli with a caller
workspace, on two calls in a row with the second on a smaller state and
fewer terms (the array keeps the first call's contents), in complex128 and
complex64; and it rejects a term table with more groups than its array
holds before any device work."""

import numpy as np
import pytest
import torch

from oracle import OracleBackend
from qibo_amd import _lib, hamiltonians

pytestmark = pytest.mark.gpu


def _rand(n, rng):
    v = rng.normal(size=2**n) + 1j * rng.normal(size=2**n)
    return v / np.linalg.norm(v)


def _both(be, state, n, xm, zm, ny, c):
    lib = _lib.load()
    code = be._dtype_code(state)
    stream = be._stream(state)
    xm = np.ascontiguousarray(xm, dtype=np.uint64)
    zm = np.ascontiguousarray(zm, dtype=np.uint64)
    ny = np.ascontiguousarray(ny, dtype=np.int32)
    c = np.ascontiguousarray(c, dtype=np.float64)
    ref = torch.full((1,), float("nan"), dtype=torch.float64, device=state.device)
    got = torch.full((1,), float("nan"), dtype=torch.float64, device=state.device)
    wsb = lib.qsim_expval_workspace(n, len(c))
    ws = torch.empty(wsb, dtype=torch.uint8, device=state.device)
    _lib.check(lib.qsim_expval_pauli(state.data_ptr(), n, code, len(c), xm.ctypes.data, zm.ctypes.data,
                                     ny.ctypes.data, c.ctypes.data, ref.data_ptr(), ws.data_ptr(), wsb, stream), lib)
    _lib.check(lib.qsim_expval_pauli_noalloc(state.data_ptr(), n, code, len(c), xm.ctypes.data, zm.ctypes.data,
                                             ny.ctypes.data, c.ctypes.data, got.data_ptr(), stream), lib)
    torch.cuda.synchronize(state.device)
    return float(ref.item()), float(got.item())


@pytest.mark.parametrize("dtype", ["complex128", "complex64"])
def test_noalloc_expval_matches_workspace_entry(backend, backend64, dtype):
    be = backend if dtype == "complex128" else backend64
    npdt = np.complex128 if dtype == "complex128" else np.complex64
    rng = np.random.default_rng(5)
    ob = OracleBackend()
    # call 1: 20 qubits, TFIM (19 ZZ terms in one group + 20 X groups) plus X/Y strings
    n1 = 20
    ham = hamiltonians.TFIM(n1, h=0.8, closed_boundary=False, backend=be)
    coeffs, strings, qubits = ham.simple_terms
    strings = list(strings) + ["XY", "YZX"]
    qubits = list(qubits) + [(0, 19), (3, 7, 11)]
    coeffs = list(coeffs) + [0.25, -0.6]
    xm, zm, ny = be.pauli_tables(strings, qubits, n1)
    psi1 = _rand(n1, rng)
    s1 = be.cast(psi1.astype(npdt))
    ref1, got1 = _both(be, s1, n1, xm, zm, ny, coeffs)
    assert got1 == ref1
    want1 = ob.exp_value_observable_symbolic(psi1.astype(npdt).astype(np.complex128), strings, qubits, coeffs, n1)
    assert abs(got1 - want1) <= (1e-12 if dtype == "complex128" else 1e-4)
    # call 2 in a row: a smaller state and fewer groups
    n2 = 13
    strings2, qubits2, coeffs2 = ["Z", "ZZ", "X", "Y"], [(0,), (2, 12), (5,), (12,)], [1.0, -0.5, 0.3, 0.7]
    xm2, zm2, ny2 = be.pauli_tables(strings2, qubits2, n2)
    psi2 = _rand(n2, rng)
    s2 = be.cast(psi2.astype(npdt))
    ref2, got2 = _both(be, s2, n2, xm2, zm2, ny2, coeffs2)
    assert got2 == ref2
    want2 = ob.exp_value_observable_symbolic(psi2.astype(npdt).astype(np.complex128), strings2, qubits2, coeffs2, n2)
    assert abs(got2 - want2) <= (1e-12 if dtype == "complex128" else 1e-4)


def test_noalloc_expval_rejects_too_many_groups(backend):
    lib = _lib.load()
    n = 10
    limit = lib.qsim_expval_noalloc_max_groups()
    assert limit == 64
    nterms = limit + 1  # distinct X masks: one group each
    xm = np.arange(1, nterms + 1, dtype=np.uint64)
    zm = np.zeros(nterms, dtype=np.uint64)
    ny = np.zeros(nterms, dtype=np.int32)
    c = np.ones(nterms, dtype=np.float64)
    state = backend.zero_state(n)
    out = torch.zeros(1, dtype=torch.float64, device=state.device)
    rc = lib.qsim_expval_pauli_noalloc(state.data_ptr(), n, _lib.C128, nterms, xm.ctypes.data, zm.ctypes.data,
                                       ny.ctypes.data, c.ctypes.data, out.data_ptr(), backend._stream(state))
    assert rc == _lib.EINVAL
    assert b"exceed the maximum of 64" in lib.qsim_last_error()
    assert float(out.item()) == 0.0  # nothing ran
    # exactly the maximum is accepted
    rc = lib.qsim_expval_pauli_noalloc(state.data_ptr(), n, _lib.C128, limit, xm.ctypes.data, zm.ctypes.data,
                                       ny.ctypes.data, c.ctypes.data, out.data_ptr(), backend._stream(state))
    assert rc == _lib.OK
    torch.cuda.synchronize(state.device)
    # <0...0| X^x |0...0> = 0 for x != 0
    assert float(out.item()) == 0.0
