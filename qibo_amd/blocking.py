"""Grouping of a circuit's gates into blocked passes over the state.

Each gate pass over an HBM-resident state is bandwidth-bound, so the backend
groups consecutive gates into passes that each read and write the state once
(``k_blocked`` in csrc/blocked.hip). A pass owns 12 "tile bits" of the
amplitude index: the lowest bits (whole 128-byte lines, so loads stay
coalesced) plus the target bits of its non-diagonal gates. Diagonal gates and
controls may sit on any bit.

Gates are moved ahead of earlier gates only when they commute with them. The
rule used is exact for the gates of this backend: two planned operations
commute when neither's non-diagonal targets touch any qubit of the other
(diagonal operators -- including control projectors -- commute with each
other).

Inside a pass the instruction list is ordered by list scheduling over the
pass's dependency graph, preferring instructions that run in the current
register mapping (see blocked.hip), so that mapping changes (LDS transposes)
are few.
"""

from typing import List, Sequence, Tuple

import numpy as np

TILE_BITS = 12
HEADER_BYTES = 32
NMAT = 64  # complex entries per record

OP_DIAG, OP_DENSE1, OP_MAP, OP_DENSE_LDS = 0, 1, 2, 3
_REG_GROUPS = {0: (8, 9, 10, 11), 1: (4, 5, 6, 7), 2: (0, 1, 2, 3)}
_SWAP = np.array([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]], dtype=complex)


class _Item:
    """A planned op with its qubit sets as amplitude-index bit positions."""

    __slots__ = ("op", "nd_targets", "all_bits", "idx")

    def __init__(self, op, nqubits, idx):
        self.op = op
        self.idx = idx
        tb = {nqubits - 1 - q for q in op.targets}
        cb = {nqubits - 1 - q for q in op.controls}
        self.nd_targets = frozenset() if op.kind == "diag" else frozenset(tb)
        self.all_bits = frozenset(tb | cb)


def _commute(x: _Item, y: _Item) -> bool:
    return not (x.nd_targets & y.all_bits) and not (y.nd_targets & x.all_bits)


def blockable(op) -> bool:
    return op.kind in ("diag", "dense", "swap") and len(op.targets) <= 3 and (
        op.kind != "diag" or len(op.targets) <= 5)


def partition(ops: Sequence, nqubits: int, low_bits: int) -> List[Tuple[str, object]]:
    """Split ``ops`` (planned gates, circuit order) into passes.

    Returns a list of ("single", op) and ("block", (tile_bits, [ops])) items
    whose sequential execution equals the circuit.
    """
    items = [_Item(op, nqubits, i) for i, op in enumerate(ops) if op.kind != "noop"]
    out = []
    while items:
        first = items[0]
        if not blockable(first.op) or nqubits < TILE_BITS:
            out.append(("single", first.op))
            items = items[1:]
            continue
        tile = set(range(low_bits))
        chosen, deferred = [], []
        td, ad = set(), set()  # non-diagonal targets / all bits of deferred items
        for it in items:
            ok = blockable(it.op) and not (it.nd_targets & ad) and not (td & it.all_bits)
            if ok and it.nd_targets:
                ok = len(tile | it.nd_targets) <= TILE_BITS
            if ok:
                chosen.append(it)
                tile |= it.nd_targets
            else:
                deferred.append(it)
                td |= it.nd_targets
                ad |= it.all_bits
        items = deferred
        if len(chosen) == 1:
            out.append(("single", chosen[0].op))
            continue
        bits = sorted(tile)
        extra = [b for b in range(nqubits) if b not in tile]
        bits = sorted(bits + extra[: TILE_BITS - len(bits)])
        out.append(("block", (bits, [c.op for c in chosen])))
    return out


def _record(itemsize: int) -> int:
    return HEADER_BYTES + NMAT * itemsize


def _pack(buf: bytearray, off: int, op: int, arg: int, cmask: int, nt: int, bits: Sequence[int],
          data: np.ndarray, cdtype):
    hdr = np.zeros(1, dtype=np.dtype([("op", "<u4"), ("arg", "<u4"), ("cmask", "<u8"), ("nt", "<u4"),
                                      ("bits", "u1", (8,)), ("pad", "<u4")]))
    hdr["op"], hdr["arg"], hdr["cmask"], hdr["nt"] = op, arg, cmask, nt
    b = np.zeros(8, dtype=np.uint8)
    b[: len(bits)] = bits
    hdr["bits"] = b
    buf[off: off + HEADER_BYTES] = hdr.tobytes()
    mat = np.zeros(NMAT, dtype=cdtype)
    if data is not None:
        flat = np.asarray(data, dtype=cdtype).ravel()
        mat[: flat.size] = flat
    isz = np.dtype(cdtype).itemsize
    buf[off + HEADER_BYTES: off + HEADER_BYTES + NMAT * isz] = mat.tobytes()


def compile_pass(tile_bits: Sequence[int], ops: Sequence, nqubits: int, cdtype) -> Tuple[bytes, int, int]:
    """Instruction records of one blocked pass; returns (bytes, count, map switches)."""
    local = {b: i for i, b in enumerate(tile_bits)}
    items = [_Item(op, nqubits, i) for i, op in enumerate(ops)]
    npred = [0] * len(items)
    succ = [[] for _ in items]
    for j, y in enumerate(items):
        for i in range(j):
            if not _commute(items[i], y):
                npred[j] += 1
                succ[i].append(j)
    ready = [i for i in range(len(items)) if npred[i] == 0]
    isz = np.dtype(cdtype).itemsize
    rec = _record(isz)
    buf = bytearray()
    count = 0
    switches = 0
    cur = 0

    def emit(op, arg, cmask, nt, bits, data):
        nonlocal count
        off = len(buf)
        buf.extend(bytes(rec))
        _pack(buf, off, op, arg, cmask, nt, bits, data, cdtype)
        count += 1

    def cmask_of(op):
        m = 0
        for q in op.controls:
            m |= 1 << (nqubits - 1 - q)
        return m

    def runs_here(i, mapping):
        op = items[i].op
        if op.kind == "diag":
            return True
        if op.kind == "dense" and len(op.targets) == 1:
            return local[nqubits - 1 - op.targets[0]] in _REG_GROUPS[mapping]
        return False

    while ready:
        ready.sort()
        pick = next((i for i in ready if runs_here(i, cur)), None)
        if pick is None:
            dense1 = [i for i in ready if items[i].op.kind == "dense" and len(items[i].op.targets) == 1]
            if dense1:
                counts = {m: sum(runs_here(i, m) for i in dense1) for m in (0, 1, 2)}
                best = max(counts, key=lambda m: (counts[m], -m))
                emit(OP_MAP, best, 0, 0, (), None)
                switches += 1
                cur = best
                pick = next(i for i in ready if runs_here(i, cur))
            else:
                pick = ready[0]
        ready.remove(pick)
        op = items[pick].op
        cm = cmask_of(op)
        if op.kind == "diag":
            bits = [nqubits - 1 - q for q in op.targets]
            emit(OP_DIAG, 0, cm, len(bits), bits, op.data)
        elif op.kind == "dense" and len(op.targets) == 1 and runs_here(pick, cur):
            lb = local[nqubits - 1 - op.targets[0]]
            emit(OP_DENSE1, _REG_GROUPS[cur].index(lb), cm, 0, (), op.data)
        else:
            mat = _SWAP if op.kind == "swap" else op.data
            lbits = [local[nqubits - 1 - q] for q in op.targets]
            k = len(lbits)
            emit(OP_DENSE_LDS, k, cm, 0, lbits + [0] * (4 - k) + sorted(lbits), mat)
        for j in succ[pick]:
            npred[j] -= 1
            if npred[j] == 0:
                ready.append(j)
    return bytes(buf), count, switches


def compile_circuit(ops: Sequence, nqubits: int, cdtype, low_bits: int):
    """Partition and compile: a list of ("single", op) / ("block", tile_bits, offset, count)
    plus the concatenated instruction bytes of all blocked passes."""
    plan = []
    program = bytearray()
    isz = np.dtype(cdtype).itemsize
    for kind, payload in partition(ops, nqubits, low_bits):
        if kind == "single":
            plan.append(("single", payload))
            continue
        tile_bits, pops = payload
        code, count, _ = compile_pass(tile_bits, pops, nqubits, cdtype)
        plan.append(("block", tuple(tile_bits), len(program) // _record(isz), count, len(pops)))
        program.extend(code)
    return plan, bytes(program)
