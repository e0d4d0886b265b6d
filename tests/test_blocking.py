"""The blocked-pass scheduler and its instruction encoding, checked on the CPU.

An independent numpy interpreter of the instruction records (the semantics
k_blocked implements, see csrc/blocked.hip) runs the compiled passes and must
reproduce the oracle's gate-by-gate state; the partition must keep every pair
of non-commuting gates in circuit order.
"""

import numpy as np
import pytest

from oracle import OracleBackend
from qibo_amd import Circuit, gates, models
from qibo_amd import blocking
from qibo_amd.backend import GatePlanner

GROUPS = {0: (8, 9, 10, 11), 1: (4, 5, 6, 7), 2: (0, 1, 2, 3)}


def decode(program: bytes, first: int, count: int, itemsize: int):
    rec = blocking.HEADER_BYTES + blocking.NMAT * itemsize
    cdt = np.complex128 if itemsize == 16 else np.complex64
    out = []
    for k in range(first, first + count):
        raw = program[k * rec:(k + 1) * rec]
        op, arg = np.frombuffer(raw[0:8], dtype="<u4")
        cmask = int(np.frombuffer(raw[8:16], dtype="<u8")[0])
        nt = int(np.frombuffer(raw[16:20], dtype="<u4")[0])
        bits = [int(b) for b in np.frombuffer(raw[20:28], dtype=np.uint8)]
        mat = np.frombuffer(raw[32:], dtype=cdt)
        out.append((int(op), int(arg), cmask, nt, bits, mat))
    return out


def run_pass(state, n, tile_bits, instrs):
    """Interpret one blocked pass on a numpy state (all tiles at once)."""
    others = [b for b in range(n) if b not in tile_bits]
    ntiles = 2 ** len(others)
    t = np.arange(ntiles)
    base = np.zeros(ntiles, dtype=np.int64)
    for i, b in enumerate(others):
        base |= ((t >> i) & 1) << b
    u = np.arange(4096)
    loc = np.zeros(4096, dtype=np.int64)
    for i, b in enumerate(tile_bits):
        loc |= ((u >> i) & 1) << b
    gidx = base[:, None] | loc[None, :]          # (ntiles, 4096) global indices
    tile = state[gidx].copy()
    cur = 0
    for op, arg, cmask, nt, bits, mat in instrs:
        active = (gidx & cmask) == cmask
        if op == blocking.OP_DIAG:
            di = np.zeros_like(gidx)
            for tb in bits[:nt]:
                di = (di << 1) | ((gidx >> tb) & 1)
            tile = np.where(active, tile * mat[di], tile)
        elif op == blocking.OP_MAP:
            cur = arg
        elif op == blocking.OP_DENSE1:
            lb = GROUPS[cur][arg]
            m = mat[:4].reshape(2, 2)
            i0 = u[(u >> lb) & 1 == 0]
            i1 = i0 | (1 << lb)
            x0, x1 = tile[:, i0], tile[:, i1]
            act = active[:, i0]
            tile[:, i0] = np.where(act, m[0, 0] * x0 + m[0, 1] * x1, x0)
            tile[:, i1] = np.where(act, m[1, 0] * x0 + m[1, 1] * x1, x1)
        else:
            k = arg
            lbits = bits[:k]
            d = 2**k
            m = mat[: d * d].reshape(d, d)
            tmask = sum(1 << b for b in lbits)
            bases = u[(u & tmask) == 0]
            offs = []
            for r in range(d):
                o = 0
                for j, b in enumerate(lbits):
                    if (r >> (k - 1 - j)) & 1:
                        o |= 1 << b
                offs.append(o)
            xs = np.stack([tile[:, bases | o] for o in offs], axis=-1)   # (ntiles, nb, d)
            ys = xs @ m.T
            act = active[:, bases]
            for r, o in enumerate(offs):
                tile[:, bases | o] = np.where(act, ys[..., r], xs[..., r])
    out = state.copy()
    out[gidx] = tile
    return out


def run_plan(ops, n, psi, cdtype=np.complex128):
    plan, program = blocking.compile_circuit(ops, n, cdtype, 3 if cdtype == np.complex128 else 4)
    itemsize = np.dtype(cdtype).itemsize
    ob = OracleBackend()
    state = psi.copy()
    for entry in plan:
        if entry[0] == "single":
            op = entry[1]
            # apply through the dense embedding of the op (oracle arithmetic)
            state = _apply_op_dense(op, state, n)
            continue
        _, tile_bits, first, count, _ = entry
        state = run_pass(state, n, list(tile_bits), decode(program, first, count, itemsize))
    return state, plan


def _apply_op_dense(op, state, n):
    from tests.test_host import emulate

    return emulate(op, state, n)


def _circuits(n):
    c1 = models.random_layered_circuit(n, 6, seed=4, measure=False)
    c2 = models.QFT(n)
    c3 = Circuit(n)
    rng = np.random.default_rng(9)
    for _ in range(120):
        kind = rng.integers(0, 7)
        q = [int(x) for x in rng.choice(n, size=3, replace=False)]
        if kind == 0:
            c3.add(gates.RX(q[0], float(rng.uniform(0, 6))))
        elif kind == 1:
            c3.add(gates.CNOT(q[0], q[1]))
        elif kind == 2:
            c3.add(gates.SWAP(q[0], q[1]))
        elif kind == 3:
            c3.add(gates.fSim(q[0], q[1], 0.3, 0.7))
        elif kind == 4:
            c3.add(gates.TOFFOLI(q[0], q[1], q[2]))
        elif kind == 5:
            c3.add(gates.RY(q[0], 0.4).controlled_by(q[1], q[2]))
        else:
            c3.add(gates.CU1(q[0], q[1], 0.9))
    return [c1, c2, c3]


@pytest.mark.parametrize("which", [0, 1, 2])
def test_compiled_passes_reproduce_oracle(which, rng):
    n = 14
    c = _circuits(n)[which]
    psi = rng.normal(size=2**n) + 1j * rng.normal(size=2**n)
    psi /= np.linalg.norm(psi)
    ops = [GatePlanner().plan_gate(g) for g in c.queue]
    got, plan = run_plan(ops, n, psi)
    want = OracleBackend().execute(c, psi)
    assert np.max(np.abs(got - want)) < 1e-12
    assert any(e[0] == "block" for e in plan)


def test_partition_respects_order():
    n = 16
    c = _circuits(n)[2]
    ops = [GatePlanner().plan_gate(g) for g in c.queue]
    index = {id(op): i for i, op in enumerate(ops)}
    order = []
    for kind, payload in blocking.partition(ops, n, 3):
        order.extend([payload] if kind == "single" else payload[1])
    assert sorted(index[id(o)] for o in order) == list(range(len(ops)))
    pos = {id(o): k for k, o in enumerate(order)}
    items = [blocking._Item(op, n, i) for i, op in enumerate(ops)]
    for i in range(len(ops)):
        for j in range(i + 1, len(ops)):
            if not blocking._commute(items[i], items[j]):
                assert pos[id(ops[i])] < pos[id(ops[j])]


def test_layered_30q_pass_count():
    c = models.random_layered_circuit(30, 20, seed=1, measure=False)
    ops = [GatePlanner().plan_gate(g) for g in c.queue]
    parts = blocking.partition(ops, 30, 3)
    assert len(parts) <= 24  # 1490 gates in a couple of dozen state traversals
    for kind, payload in parts:
        if kind == "block":
            bits, pops = payload
            assert len(bits) == blocking.TILE_BITS and bits[:3] == [0, 1, 2]
