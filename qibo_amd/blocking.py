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
other). Commuting is exact in exact arithmetic only: a reordered pair rounds
differently (the products are formed in another order), so results differ
from the per-gate path by a few units in the last place. ``ordered=True``
keeps circuit order everywhere; its passes are then bitwise equal to the
one-gate-per-pass kernels.

Inside a pass the instruction list is ordered by list scheduling over the
pass's dependency graph, preferring instructions that run in the current
register mapping (see blocked.hip), so that mapping changes (LDS transposes)
are few.

Instruction forms (the kernel's register mapping decides which bits are
"register bits": four local tile bits whose values index a lane's 16
amplitudes; the other tile bits and the bits outside the tile are fixed per
lane):

* DIAG: diagonal on up to 5 bits, per amplitude; DENSE1: complex 2x2 on a
  register bit; DENSE_LDS: 2^k x 2^k (k <= 3) on the tile in LDS; MAP: change
  of register mapping. These are the exact per-gate arithmetic (``ordered``).
* DENSE1_REAL: a 2x2 with zero imaginary parts (RY, H, X, ...), half the
  multiplies of DENSE1.
* REG_PHASE: the product of consecutive diagonal gates that involve only
  register bits, folded on the host into one 16-entry table (one complex
  multiply per amplitude for all of them).
* LANE_PHASE: a diagonal gate on bits that are all fixed per lane: it scales
  all 16 amplitudes of a lane by the same factor, so the kernel multiplies it
  into a per-lane pending factor and applies that once, before the next
  mapping change or LDS gate (it commutes with every register-local
  operation in between).
* NEG: the -1 phase of CZ-like gates under register and lane conditions.

Controls are split: ``cmask`` holds the condition bits that are fixed per
lane, ``rreq`` the register bits (as a 4-bit mask over register positions);
DENSE_LDS keeps all condition bits in ``cmask``.
"""

from typing import Iterable, Iterator, Sequence, Tuple

import numpy as np

TILE_BITS = 12
HEADER_BYTES = 32
NMAT = 64  # largest number of complex entries of a record (8x8 block)
PAD_BYTES = 1056  # readable bytes required past a record stream (QSIM_BLOCKED_PAD)

OP_DIAG, OP_DENSE1, OP_MAP, OP_DENSE_LDS = 0, 1, 2, 3
OP_DENSE1_REAL, OP_REG_PHASE, OP_LANE_PHASE, OP_NEG = 4, 5, 6, 7
REG_FLAG = 0x80  # Diag target byte: register bit j of the current mapping is REG_FLAG | j
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


WINDOW = 768  # planned gates a pass may look ahead over


def partition(ops: Iterable, nqubits: int, low_bits: int, ordered: bool = False,
              window: int = WINDOW) -> Iterator[Tuple[str, object]]:
    """Split ``ops`` (planned gates in circuit order, any iterable) into passes.

    Yields ("single", op) and ("block", (tile_bits, [ops])) items whose
    sequential execution equals the circuit, one at a time (the caller can
    launch a pass before the next is formed). A pass is chosen among the
    first ``window`` pending gates, so ``ops`` is consumed lazily: gates
    further on are planned while earlier passes run. With ``ordered`` a pass
    ends at the first gate that does not fit (no gate moves ahead).
    """
    source = iter(ops)
    items = []
    counter = 0

    def fill():
        nonlocal counter
        while len(items) < window:
            op = next(source, None)
            if op is None:
                return
            if op.kind != "noop":
                items.append(_Item(op, nqubits, counter))
                counter += 1

    fill()
    while items:
        first = items[0]
        if not blockable(first.op) or nqubits < TILE_BITS:
            yield ("single", first.op)
            del items[0]
            fill()
            continue
        tile = set(range(low_bits))
        chosen, deferred = [], []
        td, ad = set(), set()  # non-diagonal targets / all bits of deferred items
        for it in items:
            ok = blockable(it.op) and not (it.nd_targets & ad) and not (td & it.all_bits)
            if ok and it.nd_targets:
                ok = len(tile | it.nd_targets) <= TILE_BITS
            if ok and not (ordered and deferred):
                chosen.append(it)
                tile |= it.nd_targets
            else:
                deferred.append(it)
                td |= it.nd_targets
                ad |= it.all_bits
        items[:] = deferred
        fill()
        if len(chosen) == 1:
            yield ("single", chosen[0].op)
            continue
        bits = sorted(tile)
        extra = [b for b in range(nqubits) if b not in tile]
        bits = sorted(bits + extra[: TILE_BITS - len(bits)])
        yield ("block", (bits, [c.op for c in chosen]))


def record_bytes(nent: int, itemsize: int) -> int:
    """Bytes of one record: 32-byte header + nent complex entries, padded to 16."""
    return HEADER_BYTES + (nent * itemsize + 15) // 16 * 16


_HDR = np.dtype([("op", "u1"), ("arg", "u1"), ("nt", "u1"), ("rreq", "u1"), ("nent", "<u4"),
                 ("cmask", "<u8"), ("bits", "u1", (8,)), ("reserved", "<u8")])
assert _HDR.itemsize == HEADER_BYTES


def _pack(buf: bytearray, op: int, arg: int, cmask: int, nt: int, bits: Sequence[int],
          data, cdtype, rreq: int = 0):
    """Append one variable-length record: header (u8 op, u8 arg, u8 nt, u8 rreq,
    u32 nent, u64 cmask, u8 bits[8], u64 0) + nent complex entries."""
    mat = np.zeros(0, dtype=cdtype) if data is None else np.asarray(data, dtype=cdtype).ravel()
    hdr = np.zeros(1, dtype=_HDR)
    hdr["op"], hdr["arg"], hdr["nt"], hdr["rreq"], hdr["nent"], hdr["cmask"] = op, arg, nt, rreq, mat.size, cmask
    b = np.zeros(8, dtype=np.uint8)
    b[: len(bits)] = bits
    hdr["bits"] = b
    size = record_bytes(mat.size, np.dtype(cdtype).itemsize)
    rec = bytearray(size)
    rec[:HEADER_BYTES] = hdr.tobytes()
    rec[HEADER_BYTES: HEADER_BYTES + mat.nbytes] = mat.tobytes()
    buf.extend(rec)


def _dependencies(items, ordered):
    """Dependency graph of a pass: (predecessor counts, successor lists).

    Two items conflict when some bit is a non-diagonal target of one and
    touched by the other. Per bit it suffices to order a non-diagonal user
    after the last non-diagonal user and after every diagonal user since, and
    a diagonal user after the last non-diagonal user: the transitive closure
    is the full conflict relation, so the topological orders are the same.
    ``ordered`` chains the items in their given order."""
    n = len(items)
    succ = [[] for _ in range(n)]
    npred = [0] * n
    if ordered:
        for i in range(n - 1):
            succ[i].append(i + 1)
            npred[i + 1] = 1
        return npred, succ
    last_nd = {}   # bit -> index of the last non-diagonal user
    diag_since = {}  # bit -> diagonal users after it
    for j, it in enumerate(items):
        preds = set()
        for b in it.all_bits:
            if b in it.nd_targets:
                preds.update(diag_since.get(b, ()))
                if b in last_nd:
                    preds.add(last_nd[b])
            elif b in last_nd:
                preds.add(last_nd[b])
        for b in it.all_bits:
            if b in it.nd_targets:
                last_nd[b] = j
                diag_since[b] = []
            else:
                diag_since.setdefault(b, []).append(j)
        for i in preds:
            succ[i].append(j)
        npred[j] = len(preds)
    return npred, succ


_RVALS = np.array([[(r >> j) & 1 for r in range(16)] for j in range(4)], dtype=np.int64)  # [j, r]


def compile_pass(tile_bits: Sequence[int], ops: Sequence, nqubits: int, cdtype,
                 ordered: bool = False) -> Tuple[bytes, int, int]:
    """Instruction records of one blocked pass; returns (bytes, count, map switches).

    With ``ordered`` the instructions keep the order of ``ops`` and use only
    the exact per-gate forms (DIAG, DENSE1, DENSE_LDS, MAP)."""
    local = {b: i for i, b in enumerate(tile_bits)}
    items = [_Item(op, nqubits, i) for i, op in enumerate(ops)]
    npred, succ = _dependencies(items, ordered)
    ready = [i for i in range(len(items)) if npred[i] == 0]
    buf = bytearray()
    count = 0
    switches = 0
    cur = 0
    pending = None  # REG_PHASE table being folded (complex128, 16 entries)

    def emit(op, arg, cmask, nt, bits, data, rreq=0):
        nonlocal count
        _pack(buf, op, arg, cmask, nt, bits, data, cdtype, rreq)
        count += 1

    def flush_reg():
        nonlocal pending
        if pending is not None:
            emit(OP_REG_PHASE, 0, 0, 0, (), pending)
            pending = None

    def reg_bits(mapping):
        return [tile_bits[lb] for lb in _REG_GROUPS[mapping]]  # global bit of register bit j

    def split_cond(op, mapping):
        rb = reg_bits(mapping)
        cm, rreq = 0, 0
        for q in op.controls:
            b = nqubits - 1 - q
            if b in rb:
                rreq |= 1 << rb.index(b)
            else:
                cm |= 1 << b
        return cm, rreq

    def runs_here(i, mapping):
        op = items[i].op
        if op.kind == "diag":
            return True
        if op.kind == "dense" and len(op.targets) == 1:
            return local[nqubits - 1 - op.targets[0]] in _REG_GROUPS[mapping]
        return False

    def emit_diag(op):
        nonlocal pending
        rb = reg_bits(cur)
        tb = [nqubits - 1 - q for q in op.targets]
        cb = [nqubits - 1 - q for q in op.controls]
        d = np.asarray(op.data, dtype=np.complex128).ravel()
        nt = len(tb)
        if not ordered:
            if all(b not in rb for b in tb + cb):
                cm = sum(1 << b for b in cb)
                emit(OP_LANE_PHASE, 0, cm, nt, tb, d)
                return
            if all(b in rb for b in tb + cb):
                table = np.ones(16, dtype=np.complex128) if pending is None else pending
                ok = np.ones(16, dtype=bool)
                for c in cb:
                    ok &= _RVALS[rb.index(c)] == 1
                di = np.zeros(16, dtype=np.int64)
                for b in tb:
                    di = (di << 1) | _RVALS[rb.index(b)]
                pending = np.where(ok, table * d[di], table)
                return
        cm, rreq = split_cond(op, cur)
        enc = [(REG_FLAG | rb.index(b)) if b in rb else b for b in tb]
        if not ordered and nt == 0 and d[0] == -1:
            emit(OP_NEG, 0, cm, 0, (), None, rreq)
        else:
            emit(OP_DIAG, 0, cm, nt, enc, d, rreq)

    def reach(mapping):
        """Register one-qubit gates that could run in ``mapping`` from the
        current state before another mapping change is needed (the mapping
        choice looks ahead instead of counting only the ready gates)."""
        left = list(npred)
        frontier = [i for i in ready if runs_here(i, mapping)]
        done = 0
        while frontier:
            i = frontier.pop()
            if items[i].op.kind != "diag":
                done += 1
            for j in succ[i]:
                left[j] -= 1
                if left[j] == 0 and runs_here(j, mapping):
                    frontier.append(j)
        return done

    while ready:
        ready.sort()
        pick = next((i for i in ready if runs_here(i, cur)), None)
        if pick is None:
            dense1 = [i for i in ready if items[i].op.kind == "dense" and len(items[i].op.targets) == 1]
            if dense1:
                counts = {m: reach(m) for m in (0, 1, 2) if m != cur}
                best = max(counts, key=lambda m: (counts[m], -m))
                flush_reg()
                emit(OP_MAP, best, 0, 0, (), None)
                switches += 1
                cur = best
                pick = next(i for i in ready if runs_here(i, cur))
            else:
                pick = ready[0]
        ready.remove(pick)
        op = items[pick].op
        if op.kind == "diag":
            emit_diag(op)
        elif op.kind == "dense" and len(op.targets) == 1 and runs_here(pick, cur):
            flush_reg()
            lb = local[nqubits - 1 - op.targets[0]]
            cm, rreq = split_cond(op, cur)
            m = np.asarray(op.data)
            real = (not ordered) and not np.any(np.imag(m))
            emit(OP_DENSE1_REAL if real else OP_DENSE1, _REG_GROUPS[cur].index(lb), cm, 0, (), m.ravel()[:4], rreq)
        else:
            flush_reg()
            mat = _SWAP if op.kind == "swap" else op.data
            lbits = [local[nqubits - 1 - q] for q in op.targets]
            k = len(lbits)
            cm = sum(1 << (nqubits - 1 - q) for q in op.controls)
            emit(OP_DENSE_LDS, k, cm, 0, lbits + [0] * (4 - k) + sorted(lbits), mat)
        for j in succ[pick]:
            npred[j] -= 1
            if npred[j] == 0:
                ready.append(j)
    flush_reg()
    return bytes(buf), count, switches


def compile_circuit(ops: Sequence, nqubits: int, cdtype, low_bits: int, ordered: bool = False):
    """Partition and compile: a list of ("single", op) / ("block", tile_bits, byte_offset,
    count, ngates) plus the concatenated records of all blocked passes."""
    plan = []
    program = bytearray()
    for kind, payload in partition(ops, nqubits, low_bits, ordered):
        if kind == "single":
            plan.append(("single", payload))
            continue
        tile_bits, pops = payload
        code, count, _ = compile_pass(tile_bits, pops, nqubits, cdtype, ordered)
        plan.append(("block", tuple(tile_bits), len(program), count, len(pops)))
        program.extend(code)
    return plan, bytes(program)
