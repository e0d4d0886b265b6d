"""The blocked-pass scheduler and its instruction encoding, checked on the CPU.

An independent numpy interpreter of the instruction records (the semantics
k_blocked implements, see csrc/blocked.hip) runs the compiled passes and must
reproduce the oracle's gate-by-gate state; the partition must keep every pair
of non-commuting gates in circuit order.
"""

import copy
import numpy as np
import pytest

from oracle import OracleBackend
from qibo_amd import Circuit, gates, models
from qibo_amd import blocking
from qibo_amd.backend import GatePlanner

GROUPS = {0: (8, 9, 10, 11), 1: (4, 5, 6, 7), 2: (0, 1, 2, 3)}


def decode(program: bytes, first: int, count: int, itemsize: int):
    """Parse ``count`` variable-length records starting at byte ``first``."""
    cdt = np.complex128 if itemsize == 16 else np.complex64
    out = []
    off = first
    for _ in range(count):
        op, arg, nt, rreq = (int(x) for x in program[off: off + 4])
        nent = int(np.frombuffer(program[off + 4: off + 8], dtype="<u4")[0])
        cmask = int(np.frombuffer(program[off + 8: off + 16], dtype="<u8")[0])
        bits = [int(b) for b in program[off + 16: off + 24]]
        assert program[off + 24: off + 32] == bytes(8)
        size = blocking.record_bytes(nent, itemsize)
        mat = np.frombuffer(program[off + 32: off + 32 + nent * itemsize], dtype=cdt)
        if op == blocking.OP_ROT:
            # (t, s) in the header: float64 (complex128) or float32 at bytes 8 and 16
            assert nent == 0
            rdt = "<f8" if itemsize == 16 else "<f4"
            t = float(np.frombuffer(program[off + 8: off + 16], dtype=rdt)[0])
            s = float(np.frombuffer(program[off + 16: off + 24], dtype=rdt)[0])
            mat = np.array([complex(t, s)], dtype=cdt)
            cmask, bits = 0, [0] * 8
        out.append((op, arg, cmask, nt, bits, rreq, mat))
        off += size
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
    pend, qbit = np.ones_like(tile), None  # pending LANE_REGPHASE factor and its register bit
    for op, arg, cmask, nt, bits, rreq, mat in instrs:
        # nothing that fails to commute with a pending factor may run
        if qbit is not None:
            assert op not in (blocking.OP_MAP, blocking.OP_DENSE_LDS)
            if op in (blocking.OP_DENSE1, blocking.OP_DENSE1_REAL, blocking.OP_ROT):
                assert (arg & 3) != qbit
        rb = [tile_bits[lb] for lb in GROUPS[cur]]     # global bit of register bit j
        regmask = sum(1 << b for b in rb)
        if op != blocking.OP_DENSE_LDS:
            # the kernel's split: lane-fixed condition bits in cmask, register bits in rreq
            assert cmask & regmask == 0
            active = (gidx & cmask) == cmask
            for j in range(4):
                if (rreq >> j) & 1:
                    active &= ((gidx >> rb[j]) & 1) == 1
        else:
            assert rreq == 0
            active = (gidx & cmask) == cmask
        if op in (blocking.OP_DIAG, blocking.OP_LANE_PHASE):
            di = np.zeros_like(gidx)
            for tb in bits[:nt]:
                if tb & blocking.REG_FLAG:
                    assert op == blocking.OP_DIAG
                    tb = rb[tb & 3]
                elif op == blocking.OP_LANE_PHASE:
                    assert tb not in rb  # per-lane factor: no register bit may be involved
                di = (di << 1) | ((gidx >> tb) & 1)
            tile = np.where(active, tile * mat[di], tile)
        elif op == blocking.OP_NEG:
            tile = np.where(active, -tile, tile)
        elif op == blocking.OP_LANE_REGPHASE:
            # phase on register bit `arg` under the lane-fixed condition bits,
            # pending until a QFLUSH of that bit
            assert rreq == 0 and nt == 0 and cmask & regmask == 0
            assert qbit in (None, arg)
            qbit = arg
            on = active & (((gidx >> rb[arg]) & 1) == 1)
            pend = np.where(on, pend * mat[0], pend)
        elif op == blocking.OP_QFLUSH:
            assert qbit == arg
            tile = tile * pend
            pend = np.ones_like(tile)
            qbit = None
        elif op == blocking.OP_REG_PHASE:
            assert cmask == 0 and rreq == 0
            dep = arg & 15  # the kernel skips the registers with r & dep == 0
            if dep:
                assert all(mat[r] == 1 for r in range(16) if not r & dep)
            r = np.zeros_like(gidx)
            for j in range(4):
                r |= ((gidx >> rb[j]) & 1) << j
            tile = tile * mat[r]
        elif op == blocking.OP_MAP:
            cur = arg
        elif op == blocking.OP_ROT:
            # three shears per real pair (the kernel's order of steps), c >= 0
            assert cmask == 0 and rreq == 0
            lb = GROUPS[cur][arg & 3]
            mode = arg >> 2
            t, s = mat[0].real, mat[0].imag
            i0 = u[(u >> lb) & 1 == 0]
            i1 = i0 | (1 << lb)
            x0r, x0i = tile[:, i0].real.copy(), tile[:, i0].imag.copy()
            x1r, x1i = tile[:, i1].real.copy(), tile[:, i1].imag.copy()

            def shears(a, b, t, s):
                a = a - t * b
                b = b + s * a
                return a - t * b, b

            if mode == blocking.ROT_RY:
                x0r, x1r = shears(x0r, x1r, t, s)
                x0i, x1i = shears(x0i, x1i, t, s)
            elif mode == blocking.ROT_H:
                for a, b in ((x0r, x1r), (x0i, x1i)):
                    a += t * b
                    b[:] = s * a - b
                    a -= t * b
            else:
                assert mode == blocking.ROT_RX
                x0i, x1r = shears(x0i, x1r, t, s)
                x0r, x1i = shears(x0r, x1i, -t, -s)
            tile[:, i0] = x0r + 1j * x0i
            tile[:, i1] = x1r + 1j * x1i
        elif op in (blocking.OP_DENSE1, blocking.OP_DENSE1_REAL):
            lb = GROUPS[cur][arg]
            m = mat[:4].reshape(2, 2)
            if op == blocking.OP_DENSE1_REAL:
                assert not np.any(np.imag(m))
            i0 = u[(u >> lb) & 1 == 0]
            i1 = i0 | (1 << lb)
            x0, x1 = tile[:, i0], tile[:, i1]
            act = active[:, i0]
            tile[:, i0] = np.where(act, m[0, 0] * x0 + m[0, 1] * x1, x0)
            tile[:, i1] = np.where(act, m[1, 0] * x0 + m[1, 1] * x1, x1)
        else:
            assert op == blocking.OP_DENSE_LDS
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
    assert qbit is None  # flushed before the end of the pass
    out = state.copy()
    out[gidx] = tile
    return out


def run_plan(ops, n, psi, cdtype=np.complex128, ordered=False):
    plan, program = blocking.compile_circuit(ops, n, cdtype, 3 if cdtype == np.complex128 else 4, ordered)
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


def _distinct(ops):
    """Copies of the planned ops, one object per gate (parameterless gates on
    the same qubits share one plan object), so identities name positions."""
    return [copy.copy(op) for op in ops]


def test_partition_respects_order():
    n = 16
    c = _circuits(n)[2]
    ops = _distinct([GatePlanner().plan_gate(g) for g in c.queue])
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
    parts = list(blocking.partition(ops, 30, 3))
    assert len(parts) <= 24  # 1490 gates in a couple of dozen state traversals
    for kind, payload in parts:
        if kind == "block":
            bits, pops = payload
            assert len(bits) == blocking.TILE_BITS and bits[:3] == [0, 1, 2]


def test_ordered_partition_keeps_circuit_order():
    n = 16
    c = _circuits(n)[2]
    ops = _distinct([GatePlanner().plan_gate(g) for g in c.queue])
    order = []
    for kind, payload in blocking.partition(ops, n, 3, ordered=True):
        order.extend([payload] if kind == "single" else payload[1])
    assert [id(o) for o in order] == [id(o) for o in ops if o.kind != "noop"]
    for kind, payload in blocking.partition(ops, n, 3, ordered=True):
        if kind != "block":
            continue
        bits, pops = payload
        code, count, _ = blocking.compile_pass(bits, pops, n, np.complex128, ordered=True)
        kinds = [r[0] for r in decode(code, 0, count, 16)]
        assert set(kinds) <= {blocking.OP_DIAG, blocking.OP_DENSE1, blocking.OP_MAP, blocking.OP_DENSE_LDS}
        gates_emitted = [k for k in kinds if k != blocking.OP_MAP]
        assert len(gates_emitted) == len(pops)
        for got, o in zip(gates_emitted, pops):  # same sequence of gate kinds
            assert (got == blocking.OP_DIAG) == (o.kind == "diag")


def test_ordered_mode_reproduces_oracle(rng):
    n = 14
    c = _circuits(n)[2]
    psi = rng.normal(size=2**n) + 1j * rng.normal(size=2**n)
    psi /= np.linalg.norm(psi)
    ops = [GatePlanner().plan_gate(g) for g in c.queue]
    got, _ = run_plan(ops, n, psi, ordered=True)
    want = OracleBackend().execute(c, psi)
    assert np.max(np.abs(got - want)) < 1e-12


def test_fast_forms_used_on_layered_circuit():
    n = 30
    c = models.random_layered_circuit(n, 20, seed=1, measure=False)
    ops = [GatePlanner().plan_gate(g) for g in c.queue]
    plan, program = blocking.compile_circuit(ops, n, np.complex128, 3)
    kinds = {}
    for entry in plan:
        if entry[0] != "block":
            continue
        for rec in decode(program, entry[2], entry[3], 16):
            kinds[rec[0]] = kinds.get(rec[0], 0) + 1
    # RY is a rotation: every register one-qubit gate uses the shear form
    assert kinds.get(blocking.OP_DENSE1, 0) == 0 and kinds.get(blocking.OP_DENSE1_REAL, 0) == 0
    assert kinds[blocking.OP_ROT] > 0
    # RZ / CZ fold into register tables (several gates per record), sign flips
    # or per-lane factors; no general per-amplitude diagonal is left
    assert kinds[blocking.OP_REG_PHASE] > 0 and kinds.get(blocking.OP_DIAG, 0) == 0
    ngates = sum(e[4] for e in plan if e[0] == "block")
    folded = ngates - kinds[blocking.OP_ROT] - kinds.get(blocking.OP_NEG, 0) \
        - kinds.get(blocking.OP_LANE_PHASE, 0)
    assert kinds[blocking.OP_REG_PHASE] < folded / 2


@pytest.mark.parametrize("which", [0, 1, 2])
def test_compiled_passes_reproduce_oracle_c64(which, rng):
    n = 13
    c = _circuits(n)[which]
    psi = rng.normal(size=2**n) + 1j * rng.normal(size=2**n)
    psi = (psi / np.linalg.norm(psi)).astype(np.complex64)
    ops = [GatePlanner("complex64").plan_gate(g) for g in c.queue]
    got, _ = run_plan(ops, n, psi, cdtype=np.complex64)
    want = OracleBackend().execute(c, psi.astype(np.complex128))
    assert np.max(np.abs(got - want)) < 5e-6


def test_padding_contract_matches_header():
    import os
    import re

    hdr = open(os.path.join(os.path.dirname(__file__), "..", "include", "qsim_mi355x.h")).read()
    pad = int(re.search(r"#define QSIM_BLOCKED_PAD (\d+)", hdr).group(1))
    assert pad == blocking.PAD_BYTES
    # the largest record's entries, read from a record that starts inside the stream
    assert pad >= blocking.record_bytes(blocking.NMAT, 16)


@pytest.mark.parametrize("theta", [0.3, 2.0, 3.1, 4.0, 6.2, -1.7, np.pi])
def test_rotation_forms(theta):
    from qibo_amd import matrices

    for gate, mode in ((gates.RY(0, theta), blocking.ROT_RY), (gates.RX(0, theta), blocking.ROT_RX)):
        m = GatePlanner().matrix(gate)
        mode_got, t, s, neg = blocking.rotation_form(m)
        assert mode_got == mode
        c, sn = np.cos(theta / 2), np.sin(theta / 2)
        assert neg == (c < 0) and abs(t) <= 1 + 1e-15
        assert abs(s - (-sn if neg else sn)) < 1e-16
        # the shear parameters reproduce the (sign-adjusted) matrix: 1 - t s = |c|, t (2 - t s) = s
        assert abs((1 - t * s) - abs(c)) < 1e-15 and abs(t * (2 - t * s) - s) < 1e-15
    assert blocking.rotation_form(GatePlanner().matrix(gates.H(0)))[0] == blocking.ROT_H
    for g in (gates.RZ(0, theta), gates.U3(0, theta, 0.2, 0.3), gates.X(0) if theta > 10 else gates.S(0)):
        assert blocking.rotation_form(GatePlanner().matrix(g)) is None


def _rotation_only_circuit(n, angles):
    c = Circuit(n)
    for k, a in enumerate(angles):
        c.add(gates.RY(k % 3, a))
    return c


@pytest.mark.parametrize("angles,expect_neg", [([4.0, 0.5, 1.0], True), ([4.0, 4.5, 1.0], False)])
def test_rotation_sign_fold(angles, expect_neg, rng):
    """A pass with only rotations and an odd number of c < 0 ends with a sign flip;
    with an even number no record is added. Either way the state is exact."""
    n = 13
    c = _rotation_only_circuit(n, angles)
    psi = rng.normal(size=2**n) + 1j * rng.normal(size=2**n)
    psi /= np.linalg.norm(psi)
    ops = [GatePlanner().plan_gate(g) for g in c.queue]
    got, plan = run_plan(ops, n, psi)
    want = OracleBackend().execute(c, psi)
    assert np.max(np.abs(got - want)) < 1e-14
    _, program = blocking.compile_circuit(ops, n, np.complex128, 3)
    recs = decode(program, 0, plan[0][3], 16)
    assert [r[0] for r in recs].count(blocking.OP_NEG) == (1 if expect_neg else 0)


def test_rotation_sign_absorbed(rng):
    """With an unconditional linear record in the pass the sign goes into its entries."""
    n = 13
    c = _rotation_only_circuit(n, [4.0, 0.5])
    c.add(gates.RZ(1, 0.7))
    c.add(gates.RY(1, 0.2))
    psi = rng.normal(size=2**n) + 1j * rng.normal(size=2**n)
    psi /= np.linalg.norm(psi)
    ops = [GatePlanner().plan_gate(g) for g in c.queue]
    got, plan = run_plan(ops, n, psi)
    want = OracleBackend().execute(c, psi)
    assert np.max(np.abs(got - want)) < 1e-14
    _, program = blocking.compile_circuit(ops, n, np.complex128, 3)
    kinds = [r[0] for r in decode(program, 0, plan[0][3], 16)]
    assert blocking.OP_NEG not in kinds and blocking.OP_REG_PHASE in kinds


def test_small_fused_groups_run_as_their_gates(rng):
    """A fused group of at most 3 qubits is planned as its gates (GatePlanner
    .plan_ops); the passes built from them reproduce the oracle's state of the
    unfused circuit, and groups of 4-5 qubits keep one dense op."""
    n = 14
    c = models.random_layered_circuit(n, 6, seed=4, measure=False)
    fused = c.fuse(max_qubits=5)
    gp = GatePlanner()
    ops = [op for g in fused.queue for op in gp.plan_ops(g)]
    assert len(ops) == c.ngates  # every 2-3-qubit group unfolded
    psi = rng.normal(size=2**n) + 1j * rng.normal(size=2**n)
    psi /= np.linalg.norm(psi)
    got, _ = run_plan(ops, n, psi)
    want = OracleBackend().execute(c, psi)
    assert np.max(np.abs(got - want)) < 1e-12
    big = models.QFT(12).fuse(max_qubits=5)
    five = [g for g in big.queue if len(g.qubits) == 5]
    assert five and all(len(gp.plan_ops(g)) == 1 for g in five)


@pytest.mark.parametrize("theta", [0.3, 2.0, 3.1, -1.7, np.pi])
def test_rotation_forms_complex64(theta):
    """complex64 matrices hold float32 roundings of cos and sin: a ROT form at
    float32 tolerance, not at complex128's; the shears' rotation matches the
    rounded matrix to float32 rounding."""
    m = GatePlanner("complex64").matrix(gates.RY(0, theta))
    assert m.dtype == np.complex64
    tol = blocking.ROT_TOL[np.dtype(np.complex64)]
    form = blocking.rotation_form(m, tol)
    assert form is not None and form[0] == blocking.ROT_RY
    _, t, s, neg = form
    rot = np.array([[1 - t * s, -t * (2 - t * s)], [t * (2 - t * s), 1 - t * s]]) * (-1 if neg else 1)
    assert np.max(np.abs(rot - m.real)) < 2e-7
    # a matrix further from a rotation than float32 rounding is not one
    assert blocking.rotation_form(m * np.complex64(1 + 1e-5), tol) is None
