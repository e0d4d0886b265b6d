"""Run-time specialised blocked passes on the GPU (needs an MI355X).

A kernel specialised to a program's record sequence must give bitwise the
amplitudes of the interpreted pass (k_blocked) on the same program: checked
for passes holding every record kind (fast and ordered per-gate forms), in
complex128 and complex64, from a random state and from |0...0>, with tile
conditions set on some records, and for a program with other values (angles,
phases) but the same sequence run through the same kernel. End to end, a
backend that specialises after one sighting returns the same final states
and samples as one that interprets every pass.
"""

import ctypes
import struct

import numpy as np
import pytest
import torch

from oracle import OracleBackend
from qibo_amd import _lib, blocking, models
from qibo_amd.backend import MI355XBackend

from tests.specialize_programs import programs, record_ops

pytestmark = pytest.mark.gpu

ALL = (1 << 64) - 1  # support of a pass over every tile


def _upload(prog):
    return torch.from_numpy(np.frombuffer(prog + bytes(blocking.PAD_BYTES), dtype=np.uint8).copy()).cuda()


def _specialize(lib, prog, count, code, from_zero):
    handle = ctypes.c_uint64(0)
    rc = lib.qsim_blocked_specialize(prog, len(prog), count, code, int(from_zero), ctypes.byref(handle))
    assert rc == 0, lib.qsim_last_error().decode()[-2000:]
    return handle


def _specialize_all(lib, jobs):
    """Compile the kernels of (prog, count, code, from_zero) jobs on 8 host
    threads (hiprtc runs on the host; the module loads on the current device)."""
    from concurrent.futures import ThreadPoolExecutor

    dev = torch.cuda.current_device()

    def one(job):
        with torch.cuda.device(dev):
            return _specialize(lib, *job)

    with ThreadPoolExecutor(8) as ex:
        return list(ex.map(one, jobs))


def _both(lib, n, code, tb, prog, count, state, from_zero):
    """(interpreted, specialised) results of one pass on copies of state."""
    handle = _specialize(lib, prog, count, code, from_zero)
    dev = _upload(prog)
    tba = _lib.int_array(tb)
    a, b = state.clone(), state.clone()
    interp = lib.qsim_apply_blocked_from_zero if from_zero else lib.qsim_apply_blocked
    _lib.check(interp(a.data_ptr(), n, code, tba, dev.data_ptr(), len(prog), count, None), lib)
    _lib.check(lib.qsim_apply_blocked_specialized(handle.value, b.data_ptr(), 1, n, code, tba, ALL, dev.data_ptr(), 0,
                                                  len(prog), count, None), lib)
    torch.cuda.synchronize()
    return a, b


def _with_tile_conditions(prog, cdtype, n, tb, rng):
    """The program with a tile condition (a bit outside the tile) on some
    records other than mapping changes."""
    outside = [b for b in range(n) if b not in tb]
    esize = 16 if cdtype == np.complex128 else 8
    buf = bytearray(prog)
    off = 0
    for op, _, _ in record_ops(prog, cdtype):
        if op != 2 and rng.random() < 0.3:
            bits = rng.choice(outside, size=int(rng.integers(1, 3)), replace=False)
            struct.pack_into("<Q", buf, off + 24, int(sum(1 << int(x) for x in bits)))
        (nent,) = struct.unpack_from("<I", buf, off + 4)
        off += 32 + ((min(nent, 64) * esize + 15) & ~15)
    return bytes(buf)


@pytest.mark.parametrize("cdtype,code", [(np.complex128, _lib.C128), (np.complex64, _lib.C64)])
@pytest.mark.parametrize("ordered", [False, True])
def test_specialised_pass_bitwise_equals_interpreter(backend, rng, cdtype, code, ordered):
    lib = backend.lib
    n = 16
    tdt = torch.complex128 if code == _lib.C128 else torch.complex64
    psi = torch.randn(2**n, dtype=tdt, device="cuda")
    psi /= psi.norm()
    cases = []
    for i, (tb, prog, count) in enumerate(programs(n, cdtype, seed=5, depth=2, ordered=ordered)[:4]):
        cases += [(i, tb, prog, count, False), (i, tb, prog, count, True)]
        cases.append((i, tb, _with_tile_conditions(prog, cdtype, n, tb, rng), count, False))
    _specialize_all(lib, [(prog, count, code, fz) for _, _, prog, count, fz in cases])
    kinds = set()
    for i, tb, prog, count, from_zero in cases:
        kinds |= {op for op, _, _ in record_ops(prog, cdtype)}
        a, b = _both(lib, n, code, tb, prog, count, psi, from_zero)
        assert torch.equal(a, b), (i, from_zero)
    assert ({0, 1, 2, 3} if ordered else {1, 2, 3, 4, 5, 7, 10, 11, 12}) <= kinds


def test_one_kernel_for_programs_with_other_values(backend):
    """Two layered circuits with other angles compile to programs of one
    record sequence: the second reuses the first's kernel and still matches
    its own interpreted pass bitwise."""
    lib = backend.lib
    n = 16
    from qibo_amd.backend import GatePlanner

    def first_pass(seed):
        ops = []
        for g in models.random_layered_circuit(n, 4, seed=seed, measure=False).queue:
            ops.extend(GatePlanner().plan_ops(g))
        kind, (tb, pops) = next(iter(blocking.partition(ops, n, 3, False)))
        prog, count, _ = blocking.compile_pass(tb, pops, n, np.complex128, False)
        return list(tb), bytes(prog), count

    tb1, p1, c1 = first_pass(1)
    tb2, p2, c2 = first_pass(2)
    assert [op for op, _, _ in record_ops(p1, np.complex128)] == [op for op, _, _ in record_ops(p2, np.complex128)]
    assert p1 != p2
    h1, h2 = ctypes.c_uint64(0), ctypes.c_uint64(0)
    assert lib.qsim_blocked_specialize(p1, len(p1), c1, _lib.C128, 0, ctypes.byref(h1)) == 0
    assert lib.qsim_blocked_specialized_lookup(p2, len(p2), c2, _lib.C128, 0, ctypes.byref(h2), None) == 0
    if tb1 == tb2:
        assert h2.value == h1.value
    psi = torch.randn(2**n, dtype=torch.complex128, device="cuda")
    a, b = _both(lib, n, _lib.C128, tb2, p2, c2, psi, False)
    assert torch.equal(a, b)


def test_specialised_launch_errors(backend):
    lib = backend.lib
    n = 14
    tb, prog, count = programs(n, np.complex128, seed=6, depth=1)[0]
    h = ctypes.c_uint64(0)
    assert lib.qsim_blocked_specialize(prog, len(prog), count, _lib.C128, 0, ctypes.byref(h)) == 0
    dev = _upload(prog)
    psi = torch.zeros(2**n, dtype=torch.complex128, device="cuda")
    tba = _lib.int_array(tb)
    bad = [
        (h.value + 10**6, psi, _lib.C128, len(prog), count),  # unknown handle
        (h.value, psi, _lib.C64, len(prog), count),  # other dtype
        (h.value, psi, _lib.C128, 64, count),  # program shorter than the records
        (h.value, psi, _lib.C128, len(prog), count - 1),
    ]
    for handle, st, code, nbytes, cnt in bad:
        assert lib.qsim_apply_blocked_specialized(handle, st.data_ptr(), 1, n, code, tba, ALL, dev.data_ptr(), 0,
                                                  nbytes, cnt, None) == _lib.EINVAL


@pytest.mark.parametrize("dtype", ["complex128", "complex64"])
def test_backend_specialises_repeated_circuits(dtype):
    """A backend that specialises after one sighting: the second and third
    executions run specialised kernels and every result equals the
    interpreting backend's bitwise (states and seeded samples)."""
    spec = MI355XBackend(dtype=dtype)
    spec.specialize, spec.specialize_after = True, 1
    plain = MI355XBackend(dtype=dtype)
    plain.specialize = False
    n = 18
    for rep in range(3):
        c = models.random_layered_circuit(n, 6, seed=11, measure=True)
        got = spec.execute_circuit(c, nshots=2000)
        if rep == 0:
            assert spec.wait_specializations(timeout=300) >= 1
        np.random.seed(7 + rep)
        gs = got.samples(binary=False).cpu().numpy()
        c2 = models.random_layered_circuit(n, 6, seed=11, measure=True)
        want = plain.execute_circuit(c2, nshots=2000)
        np.random.seed(7 + rep)
        ws = want.samples(binary=False).cpu().numpy()
        assert torch.equal(got.state(), want.state())
        assert np.array_equal(gs, ws)
    assert spec.specialized_launches >= 2
    # and the oracle agrees
    ref = OracleBackend().execute(models.random_layered_circuit(n, 6, seed=11, measure=False))
    tol = 1e-12 if dtype == "complex128" else 5e-6
    assert np.max(np.abs(got.state(numpy=True).astype(np.complex128) - ref)) <= tol


@pytest.mark.parametrize("dtype", ["complex128", "complex64"])
def test_adjoint_sweeps_specialised_bitwise(dtype):
    """Energies and gradients of a batch of parameter sets from sweeps whose
    forward and reverse passes run their specialised kernels equal those of
    the interpreted sweeps bitwise."""
    from qibo_amd import hamiltonians
    from qibo_amd.adjoint import batched_expectation_and_gradient

    from qibo_amd import Circuit, gates

    # small enough that every sweep pass fits the library's compile budget
    n, nsets = 13, 8
    c = Circuit(n)
    for q in range(8):
        c.add(gates.RY(q, 0.1))
    for q in range(n - 1):
        c.add(gates.CNOT(q, q + 1))
    for q in range(4):
        c.add(gates.RZ(q, 0.2))
        c.add(gates.RX(q + 5, 0.3))
    params = np.random.default_rng(3).uniform(0, 2 * np.pi, (nsets, 16))
    spec = MI355XBackend(dtype=dtype)
    spec.specialize, spec.specialize_after, spec.specialize_sweeps = True, 1, True
    plain = MI355XBackend(dtype=dtype)
    plain.specialize = False
    ham_s = hamiltonians.TFIM(n, h=1.0, closed_boundary=False, backend=spec)
    ham_p = hamiltonians.TFIM(n, h=1.0, closed_boundary=False, backend=plain)
    batched_expectation_and_gradient(spec, c, ham_s, params)  # compiles start here
    assert spec.wait_specializations(timeout=600) >= 2
    from qibo_amd.adjoint import sweep_template

    tmpl = sweep_template(spec, c)
    code = _lib.C128 if dtype == "complex128" else _lib.C64
    for p, variant in [(p, _lib.SPEC_PASS) for p in tmpl.fwd] + [(p, _lib.SPEC_ADJOINT) for p in tmpl.rev]:
        h = ctypes.c_uint64(0)
        row = p.row.tobytes()
        assert spec.lib.qsim_blocked_specialized_lookup(row, len(row), p.count, code, variant, ctypes.byref(h),
                                                        None) == 0
        assert h.value, "every sweep pass has its kernel"
    before = spec.specialized_launches
    e1, g1 = batched_expectation_and_gradient(spec, c, ham_s, params)
    assert spec.specialized_launches > before
    e2, g2 = batched_expectation_and_gradient(plain, c, ham_p, params)
    assert torch.equal(e1, e2)
    assert torch.equal(g1, g2)
