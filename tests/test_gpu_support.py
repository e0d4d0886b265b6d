"""Blocked passes that run only the tiles inside a support (needs an MI355X).

A state built from |0...0> by passes vanishes wherever an index bit outside
the union of their tile bits is 1, and a pass keeps it so (its non-diagonal
work stays on its own tile bits). ``qsim_apply_blocked_support`` runs only the
tiles inside the support and reads only the amplitudes inside it, so the
rest of the state need not be written until ``qsim_zero_outside_support``
completes it: from |0...0> the backend writes amplitude 0, runs the first pass
on tile 0 alone and completes the state only if a pass over every tile has not
written it by the time anything else reads it. Checked: equal amplitudes to
the pass over every tile on states with such a support (every record kind,
both precisions, with NaN outside the support), the backend's circuits from
|0...0> against the same backend without the skipping and against the oracle
(the support filling up, and not: completed at the end of the circuit, before
a one-gate launch, on a replay of the compiled passes), and that the skipping
happened.
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
            # and some of the tile bits (all of them once)
            support |= tmask if nsup == 1 else int(sum(1 << int(b) for b in rng.choice(tb, 5, replace=False)))
            # a random state that vanishes outside the support; the pass's
            # output vanishes outside support | tile bits
            idx = np.arange(2**n)
            keep_in = (idx & ~support) == 0
            keep = (idx & ~(support | tmask)) == 0
            v = (rng.normal(size=2**n) + 1j * rng.normal(size=2**n)) * keep_in
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
            # the amplitudes outside the support are not read: NaN there, then
            # completed by qsim_zero_outside_support
            c = psi.clone()
            c[torch.as_tensor(~keep_in, device="cuda")] = float("nan")
            _lib.check(lib.qsim_apply_blocked_support(c.data_ptr(), n, code, tba, support, dev.data_ptr(), len(prog),
                                                      count, None), lib)
            _lib.check(lib.qsim_zero_outside_support(c.data_ptr(), n, code, support | tmask, None), lib)
            torch.cuda.synchronize()
            assert torch.equal(a, c), (tb, nsup)


@pytest.mark.parametrize("code,tdt", [(_lib.C128, torch.complex128), (_lib.C64, torch.complex64)])
def test_zero_outside_support(backend, rng, code, tdt):
    lib = backend.lib
    n = 18
    v = torch.as_tensor(rng.normal(size=2**n) + 1j * rng.normal(size=2**n), dtype=tdt, device="cuda")
    for support in (0, 0b101, (1 << n) - 1 - (1 << 7), (1 << n) - 1, (1 << 64) - 1):
        a = v.clone()
        _lib.check(lib.qsim_zero_outside_support(a.data_ptr(), n, code, support, None), lib)
        torch.cuda.synchronize()
        idx = np.arange(2**n)
        inside = torch.as_tensor((idx & (((1 << n) - 1) & ~support)) == 0, device="cuda")
        assert torch.equal(a[inside], v[inside]), support
        assert not a[~inside].any(), support


def _poison(n, dtype):
    """Leave NaN in the block that torch's caching allocator hands to the
    next state of 2^n amplitudes: unwritten amplitudes that a pass read would
    show up."""
    junk = torch.full((2**n,), complex(float("nan"), float("nan")), dtype=getattr(torch, dtype), device="cuda")
    torch.cuda.synchronize()
    del junk


@pytest.mark.parametrize("dtype", ["complex128", "complex64"])
@pytest.mark.parametrize("n,spec", [(16, False), (24, False), (24, True)])
def test_circuits_from_zero_skip_zero_tiles(dtype, n, spec):
    skip = MI355XBackend(dtype=dtype)
    skip.specialize, skip.specialize_after = spec, 1
    full = MI355XBackend(dtype=dtype)
    full.specialize, full.skip_zero_tiles = False, False
    c = models.random_layered_circuit(n, 8, seed=21, measure=False)
    _poison(n, dtype)
    a = skip.execute_circuit(c).state()
    assert not torch.isnan(a).any()
    # at 24 qubits the first pass runs tile 0 and the second 2^9 of the 2^12
    # tiles (at 16 the first pass runs every tile from zeros in registers and
    # the first two passes' tile bits already cover every qubit)
    assert skip.sparse_launches == (2 if n == 24 else 0)
    b = full.execute_circuit(models.random_layered_circuit(n, 8, seed=21, measure=False)).state()
    assert full.sparse_launches == 0
    assert torch.equal(a, b)
    if spec:  # the same circuit again through the specialised passes
        skip.wait_specializations()
        _poison(n, dtype)
        a = skip.execute_circuit(models.random_layered_circuit(n, 8, seed=21, measure=False)).state()
        assert skip.specialized_launches > 0
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


def _partial_circuit(n, rng, single):
    """Gates on qubits 0-13 of n only (the support never covers the state),
    with a 4-qubit unitary (a one-gate launch) in the middle if ``single``."""
    from qibo_amd import Circuit, gates
    c = Circuit(n)
    for layer in range(4):
        for q in range(14):
            c.add(gates.RY(q, float(rng.uniform(-3, 3))))
        for q in range(layer % 2, 13, 2):
            c.add(gates.CZ(q, q + 1))
        if single and layer == 1:
            m, _ = np.linalg.qr(rng.normal(size=(16, 16)) + 1j * rng.normal(size=(16, 16)))
            c.add(gates.Unitary(m, 2, 5, 9, 12))
    return c


@pytest.mark.parametrize("dtype", ["complex128", "complex64"])
@pytest.mark.parametrize("single", [False, True])
@pytest.mark.parametrize("spec", [False, True])
def test_partial_support_completed(dtype, single, spec):
    n = 24
    skip = MI355XBackend(dtype=dtype)
    skip.specialize, skip.specialize_after = spec, 1
    full = MI355XBackend(dtype=dtype)
    full.specialize, full.skip_zero_tiles = False, False
    for rep in range(2):  # the second run replays the compiled passes (specialised with spec)
        _poison(n, dtype)
        a = skip.execute_circuit(_partial_circuit(n, np.random.default_rng(5), single)).state()
        b = full.execute_circuit(_partial_circuit(n, np.random.default_rng(5), single)).state()
        assert torch.equal(a, b), rep
        assert not torch.isnan(a).any()
        if spec:
            skip.wait_specializations()
    assert skip.sparse_launches > 0
    assert (skip.specialized_launches > 0) == spec
    small = OracleBackend().execute(_partial_circuit(14, np.random.default_rng(5), single))
    # qubit q of the 24 is index bit 23 - q: the 14-qubit state sits on the
    # indices whose 10 low bits are zero
    got = a.cpu().numpy().astype(np.complex128)
    tol = 1e-12 if dtype == "complex128" else 5e-6
    assert np.max(np.abs(got[:: 1 << 10] - small)) <= tol
    assert not np.delete(got, np.arange(0, 2**n, 1 << 10)).any()
