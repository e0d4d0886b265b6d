"""Blocked-pass programs that cover every record kind, for the tests of the
run-time specialised passes (test_specialize_source.py on the CPU,
test_gpu_specialize.py on the GPU)."""

import struct

import numpy as np

from qibo_amd import Circuit, blocking, gates
from qibo_amd.backend import GatePlanner

OP_NAMES = {0: "DIAG", 1: "DENSE1", 2: "MAP", 3: "DENSE_LDS", 4: "DENSE1_REAL", 5: "REG_PHASE",
            6: "LANE_PHASE", 7: "NEG", 10: "ROT", 11: "LANE_REGPHASE", 12: "QFLUSH", 0xFF: "NOP"}


def _unitary(k, rng):
    q, _ = np.linalg.qr(rng.normal(size=(2**k, 2**k)) + 1j * rng.normal(size=(2**k, 2**k)))
    return q


def mixed_circuit(n, rng, depth=6):
    """Every gate family the planner has a record form for, on all qubits,
    with controls inside and outside a pass's tile."""
    c = Circuit(n)
    for layer in range(depth):
        for q in range(n):
            c.add(gates.RY(q, float(rng.uniform(-3, 3))))
            c.add(gates.RZ(q, float(rng.uniform(-3, 3))))
        for q in range(layer % 2, n - 1, 2):
            c.add(gates.CZ(q, q + 1))
        a, b = (int(x) for x in rng.choice(n, 2, replace=False))
        c.add(gates.H(a))
        c.add(gates.RX(b, float(rng.uniform(-3, 3))))
        c.add(gates.U3(a, 0.3, -0.7, 1.1))
        c.add(gates.CU1(a, b, float(rng.uniform(-3, 3))))
        c.add(gates.CNOT(b, a))
        c.add(gates.SWAP(a, (a + 3) % n))
        c.add(gates.fSim(b, (b + 5) % n, 0.4, 1.3))
        trio = [int(x) for x in rng.choice(np.arange(1, n - 2), 3, replace=False)]  # clear of the controls
        c.add(gates.Unitary(_unitary(2, rng), trio[0], trio[1]))
        c.add(gates.Unitary(_unitary(3, rng), *trio))
        c.add(gates.TOFFOLI(*trio))
        c.add(gates.CCZ(trio[2], trio[0], trio[1]))
        c.add(gates.RX(trio[1], 0.8).controlled_by(n - 1, 0))
        c.add(gates.U1(trio[2], 0.2).controlled_by(n - 2))
        c.add(gates.Z(trio[0]))
        c.add(gates.S(trio[1]))
        c.add(gates.T(trio[2]).controlled_by(n - 1))
    return c


def programs(n, cdtype, seed=0, depth=6, ordered=False):
    """(tile_bits, program bytes, record count) of every blocked pass of the
    mixed circuit."""
    rng = np.random.default_rng(seed)
    planner = GatePlanner()
    ops = []
    for g in mixed_circuit(n, rng, depth).queue:
        ops.extend(planner.plan_ops(g))
    low = 3 if cdtype == np.complex128 else 4
    out = []
    for kind, payload in blocking.partition(ops, n, low, ordered):
        if kind != "block":
            continue
        tile_bits, pops = payload
        prog, count, _ = blocking.compile_pass(tile_bits, pops, n, cdtype, ordered)
        out.append((list(tile_bits), bytes(prog), count))
    return out


def record_ops(prog, cdtype):
    """(op, arg, has tile condition) of every record of a program."""
    esize = 16 if cdtype == np.complex128 else 8
    off, out = 0, []
    while off + 32 <= len(prog):
        op, arg, _, _, nent = struct.unpack_from("<BBBBI", prog, off)
        (tmask,) = struct.unpack_from("<Q", prog, off + 24)
        out.append((op, arg, tmask != 0))
        off += 32 + ((min(nent, 64) * esize + 15) & ~15)
    return out
