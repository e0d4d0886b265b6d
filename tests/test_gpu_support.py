"""Blocked passes that run only the tiles inside a support (needs an MI355X).

A state built from |0...0> by passes vanishes wherever an index bit outside
the union of their tile bits is 1, and a pass keeps it so (its non-diagonal
work stays on its own tile bits). ``qsim_apply_blocked_support`` runs only the
tiles inside the support; the from-zero pass writes zeros and runs tile 0
alone. Checked: equal amplitudes to the pass over every tile on states with
such a support (every record kind, both precisions), the backend's circuits
from |0...0> against the same backend without the skipping and against the
oracle, and that the skipping happened.
"""

import numpy as np
import pytest
import torch

from oracle import OracleBackend
from qibo_amd import _lib, blocking, models
from qibo_amd.backend import MI355XBackend

from tests.specialize_programs import programs

pytestmark = pytest.mark.gpu


def _upload(prog):
    return torch.from_numpy(np.frombuffer(prog + bytes(blocking.PAD_BYTES), dtype=np.uint8).copy()).cuda()


@pytest.mark.parametrize("cdtype,code,tdt", [(np.complex128, _lib.C128, torch.complex128),
                                             (np.complex64, _lib.C64, torch.complex64)])
def test_support_pass_equals_full_pass(backend, rng, cdtype, code, tdt):
    lib = backend.lib
    n = 16
    for tb, prog, count in programs(n, cdtype, seed=9, depth=2)[:3]:
        tmask = sum(1 << b for b in tb)
        outer = [b for b in range(n) if b not in tb]
        for nsup in (0, 1, 3):
            support = int(sum(1 << int(b) for b in rng.choice(outer, nsup, replace=False))) if nsup else 0
            # a random state that vanishes outside support | tile bits
            idx = np.arange(2**n)
            keep = (idx & ~(support | tmask)) == 0
            v = (rng.normal(size=2**n) + 1j * rng.normal(size=2**n)) * keep
            psi = torch.as_tensor((v / np.linalg.norm(v)).astype(cdtype), device="cuda")
            dev = _upload(prog)
            a, b = psi.clone(), psi.clone()
            tba = _lib.int_array(tb)
            _lib.check(lib.qsim_apply_blocked(a.data_ptr(), n, code, tba, dev.data_ptr(), len(prog), count, None), lib)
            _lib.check(lib.qsim_apply_blocked_support(b.data_ptr(), n, code, tba, support, dev.data_ptr(), len(prog),
                                                      count, None), lib)
            torch.cuda.synchronize()
            assert torch.equal(a, b), (tb, nsup)
            assert not b.cpu().numpy()[~keep].any()


@pytest.mark.parametrize("dtype", ["complex128", "complex64"])
@pytest.mark.parametrize("n", [16, 24])
def test_circuits_from_zero_skip_zero_tiles(dtype, n):
    skip = MI355XBackend(dtype=dtype)
    skip.specialize = False
    full = MI355XBackend(dtype=dtype)
    full.specialize, full.skip_zero_tiles = False, False
    c = models.random_layered_circuit(n, 8, seed=21, measure=False)
    a = skip.execute_circuit(c).state()
    # at 24 qubits the second pass runs 2^9 of the 2^12 tiles (at 16 the
    # first two passes' tile bits already cover every qubit)
    assert skip.sparse_launches == (1 if n == 24 else 0)
    b = full.execute_circuit(models.random_layered_circuit(n, 8, seed=21, measure=False)).state()
    assert full.sparse_launches == 0
    assert torch.equal(a, b)
    if n <= 16:
        want = OracleBackend().execute(models.random_layered_circuit(n, 8, seed=21, measure=False))
        tol = 1e-12 if dtype == "complex128" else 5e-6
        assert np.max(np.abs(a.cpu().numpy().astype(np.complex128) - want)) <= tol


def test_initial_state_runs_every_tile(backend, rng):
    """A circuit from a given state has no known support: every pass runs
    every tile."""
    n = 16
    v = rng.normal(size=2**n) + 1j * rng.normal(size=2**n)
    v /= np.linalg.norm(v)
    before = backend.sparse_launches
    got = backend.execute_circuit(models.random_layered_circuit(n, 4, seed=3, measure=False),
                                  initial_state=v.copy()).state(numpy=True)
    assert backend.sparse_launches == before
    want = OracleBackend().execute(models.random_layered_circuit(n, 4, seed=3, measure=False), initial_state=v)
    assert np.max(np.abs(got - want)) <= 1e-12
