"""The native record compiler (qsim_compile_pass) against its Python
restatement (blocking.compile_pass(native=False)): the same records, header
for header, on layered, QFT, hardware-efficient and random-gate circuits
(controlled_by gates, SWAPs, TOFFOLIs, CU1 phases), complex128 and complex64,
ordered and not. Entries are equal too: where the Python side folds phase
tables with numpy's complex multiply the native side forms the same fused
multiply-adds, which numpy uses on AVX-512 hosts; elsewhere a last-bit
difference of those products is allowed (2 ulp). Host code only: needs the
built library, no GPU."""

import struct

import numpy as np
import pytest

from qibo_amd import Circuit, blocking, gates, models
from qibo_amd.backend import GatePlanner


def _random_circuit(n, ngates, rng):
    c = Circuit(n)
    for _ in range(ngates):
        k = int(rng.integers(0, 11))
        q = [int(x) for x in rng.choice(n, 4, replace=False)]
        th = float(rng.uniform(-6, 6))
        if k == 0:
            c.add(gates.RY(q[0], th))
        elif k == 1:
            c.add(gates.RZ(q[0], th))
        elif k == 2:
            c.add(gates.CZ(q[0], q[1]))
        elif k == 3:
            c.add(gates.CNOT(q[0], q[1]))
        elif k == 4:
            c.add(gates.TOFFOLI(q[0], q[1], q[2]))
        elif k == 5:
            c.add(gates.H(q[0]))
        elif k == 6:
            c.add(gates.CU1(q[0], q[1], th))
        elif k == 7:
            c.add(gates.SWAP(q[0], q[1]))
        elif k == 8:
            c.add(gates.RX(q[0], th).controlled_by(q[1], q[2]))
        elif k == 9:
            c.add(gates.U3(q[0], th, 0.3, -0.7))
        else:
            c.add(gates.fSim(q[0], q[1], th, 0.4))
    return c


def _circuits():
    rng = np.random.default_rng(17)
    yield models.random_layered_circuit(20, 8, seed=3, measure=False)
    yield models.random_layered_circuit(30, 20, seed=1234, measure=False)
    yield models.QFT(16)
    yield models.hardware_efficient_ansatz(14, 3, params=rng.uniform(0, 6, 2 * 14 * 3))
    for n in (12, 15, 18):
        yield _random_circuit(n, 300, rng)


def _records(program, count, itemsize):
    off, out = 0, []
    for _ in range(count):
        head = program[off: off + 32]
        nent = struct.unpack_from("<I", head, 4)[0]
        size = blocking.record_bytes(nent, itemsize)
        ent = np.frombuffer(program, dtype=np.complex128 if itemsize == 16 else np.complex64, count=nent,
                            offset=off + 32)
        out.append((head, ent))
        off += size
    assert off == len(program)
    return out


@pytest.mark.parametrize("cdtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("ordered", [False, True])
def test_native_compiler_matches_python(cdtype, ordered):
    gp = GatePlanner("complex128" if cdtype == np.complex128 else "complex64")
    itemsize = np.dtype(cdtype).itemsize
    npasses = nrec = 0
    for c in _circuits():
        n = c.nqubits
        ops = [gp.plan_gate(g) for g in c.queue]
        for kind, payload in blocking.partition(ops, n, 3, ordered):
            if kind != "block":
                continue
            bits, pops = payload
            want, wc, ws = blocking.compile_pass(bits, pops, n, cdtype, ordered, native=False)
            got, gc, gs = blocking.compile_pass(bits, pops, n, cdtype, ordered, native=True)
            assert (gc, gs, len(got)) == (wc, ws, len(want))
            for (h1, e1), (h2, e2) in zip(_records(got, gc, itemsize), _records(want, wc, itemsize)):
                assert h1 == h2
                tol = 2 * np.finfo(cdtype).eps * np.maximum(np.abs(e2), 1e-300)
                assert np.all(np.abs(e1 - e2) <= tol)
            npasses += 1
            nrec += gc
    assert npasses >= 20 and nrec >= 500


def test_native_compiler_rejects_bad_input():
    from qibo_amd import _lib

    lib = _lib.load()
    gp = GatePlanner("complex128")
    op = gp.plan_gate(gates.H(20))  # bit 0 of a 21-qubit register: outside these tile bits
    with pytest.raises(ValueError):
        blocking.compile_pass(list(range(1, 13)), [op], 21, np.complex128)
    assert lib.qsim_compile_pass_capacity(10, _lib.C128) > 10 * 32
