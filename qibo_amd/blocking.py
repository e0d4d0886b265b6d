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
  multiply per amplitude for all of them). After a pass's first table, a
  table's value at register 0 is taken out as a global factor (folded into
  the first table): the rest is 1 wherever the bits the table depends on are
  0, and the record multiplies only the other registers (arg = those bits).
* LANE_PHASE: a diagonal gate on bits that are all fixed per lane: it scales
  all 16 amplitudes of a lane by the same factor, so the kernel multiplies it
  into a per-lane pending factor and applies that once, before the next
  mapping change or LDS gate (it commutes with every register-local
  operation in between).
* NEG: the -1 phase of CZ-like gates under register and lane conditions.
* LANE_REGPHASE: a controlled phase (one entry) whose condition holds one
  register bit and otherwise bits fixed per lane (CU1-like phases with one
  qubit in the registers): multiplied into a per-lane pending factor of the
  registers with that bit set; QFLUSH applies it, placed by the compiler
  before the next non-diagonal record on that bit, a LANE_REGPHASE on
  another bit, a mapping change, a tile gather and the end of the pass.
* ROT: an unconditional one-qubit rotation [[c, -s], [s, c]] (RY), reflection
  [[c, s], [s, -c]] (H) or [[c, -is], [-is, c]] (RX) on a register bit as
  three shears x0 -= t x1; x1 += s x0; x0 -= t x1 with t = s / (1 + c)
  (Paeth): three multiply-adds per real pair instead of four operations.
  (t, s) travel in the record header (no entries: no scalar load of its own). The
  shears need c >= 0; a gate with c < 0 is applied as -(its negation), and
  the pass's net sign is folded into the data of one unconditional linear
  record of the pass (or, if it has none, a final sign flip).

Controls are split: ``cmask`` holds the condition bits that are fixed per
lane, ``rreq`` the register bits (as a 4-bit mask over register positions);
DENSE_LDS keeps all condition bits in ``cmask``.
"""

import ctypes
import functools
import struct
from typing import Iterable, Iterator, Sequence, Tuple

import numpy as np

TILE_BITS = 12
HEADER_BYTES = 32
NMAT = 64  # largest number of complex entries of a record (8x8 block)
PAD_BYTES = 1056  # readable bytes required past a record stream (QSIM_BLOCKED_PAD)

OP_DIAG, OP_DENSE1, OP_MAP, OP_DENSE_LDS = 0, 1, 2, 3
OP_DENSE1_REAL, OP_REG_PHASE, OP_LANE_PHASE, OP_NEG = 4, 5, 6, 7
OP_ROT = 10
ROT_RY, ROT_H, ROT_RX = 0, 1, 2  # ROT modes (arg = register bit | mode << 2)
OP_LANE_REGPHASE, OP_QFLUSH = 11, 12
OP_NOP = 0xFF  # header only, no effect: pads programs of a batched launch to one record count
REG_FLAG = 0x80  # Diag target byte: register bit j of the current mapping is REG_FLAG | j
_REG_GROUPS = {0: (8, 9, 10, 11), 1: (4, 5, 6, 7), 2: (0, 1, 2, 3)}
_SWAP = np.array([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]], dtype=complex)


class _Item:
    """A planned op with its qubit sets as amplitude-index bit positions."""

    __slots__ = ("op", "nd_targets", "all_bits", "idx", "nd_mask", "all_mask", "ok")

    def __init__(self, op, nqubits, idx):
        self.op = op
        self.idx = idx
        tb = {nqubits - 1 - q for q in op.targets}
        cb = {nqubits - 1 - q for q in op.controls}
        self.nd_targets = frozenset() if op.kind in ("diag", "innerd") else frozenset(tb)
        self.all_bits = frozenset(tb | cb)
        self.nd_mask = sum(1 << b for b in self.nd_targets)  # the same sets as bit masks
        self.all_mask = sum(1 << b for b in self.all_bits)
        self.ok = blockable(op)


def _commute(x: _Item, y: _Item) -> bool:
    return not (x.nd_targets & y.all_bits) and not (y.nd_targets & x.all_bits)


def blockable(op) -> bool:
    if op.kind in ("diag", "innerd"):
        return len(op.targets) <= 5
    if op.kind == "inner1":
        return len(op.targets) == 1
    return op.kind in ("dense", "swap") and len(op.targets) <= 3


WINDOW = 768  # planned gates a pass may look ahead over


def _scan(items, tile, ordered):
    """Greedy pass over ``items`` starting from the tile bits ``tile``: take
    every gate that commutes with all gates deferred before it and whose
    non-diagonal targets fit in the 12 tile bits. Returns (chosen, deferred, tile)."""
    tmask = sum(1 << b for b in tile)
    chosen, deferred = [], []
    td = ad = 0  # non-diagonal targets / all bits of the deferred items (masks)
    for it in items:
        nd = it.nd_mask
        ok = it.ok and not (nd & ad) and not (td & it.all_mask)
        if ok and nd:
            ok = (tmask | nd).bit_count() <= TILE_BITS
        if ok and not (ordered and deferred):
            chosen.append(it)
            tmask |= nd
        else:
            deferred.append(it)
            td |= nd
            ad |= it.all_mask
    return chosen, deferred, {b for b in range(tmask.bit_length()) if (tmask >> b) & 1}


def _scan_native(items, nqubits, low_bits, ordered):
    """_scan from every seed of ``partition`` in one native call
    (qsim_plan_pass): the same choice as the Python loop, item for item."""
    from . import _lib

    lib = _lib.load()
    nitems = len(items)
    nd = np.fromiter((it.nd_mask for it in items), dtype=np.uint64, count=nitems)
    al = np.fromiter((it.all_mask for it in items), dtype=np.uint64, count=nitems)
    ok = np.fromiter((it.ok for it in items), dtype=np.uint8, count=nitems)
    taken = np.empty(nitems, dtype=np.uint8)
    tile = ctypes.c_uint64(0)
    rc = lib.qsim_plan_pass(nitems, nd.ctypes.data, al.ctypes.data, ok.ctypes.data, nqubits, low_bits, TILE_BITS,
                            1 if ordered else 0, taken.ctypes.data, ctypes.byref(tile))
    if rc < 0:
        _lib.check(rc, lib)
    chosen = [it for it, t in zip(items, taken.tolist()) if t]
    deferred = [it for it, t in zip(items, taken.tolist()) if not t]
    tm = tile.value
    return chosen, deferred, {b for b in range(tm.bit_length()) if (tm >> b) & 1}


def partition(ops: Iterable, nqubits: int, low_bits: int, ordered: bool = False,
              window: int = WINDOW, singles: bool = True, native: bool = True) -> Iterator[Tuple[str, object]]:
    """Split ``ops`` (planned gates in circuit order, any iterable) into passes.

    Yields ("single", op) and ("block", (tile_bits, [ops])) items whose
    sequential execution equals the circuit, one at a time (the caller can
    launch a pass before the next is formed; with ``singles=False`` a
    one-gate pass is still a block). A pass is chosen among the
    first ``window`` pending gates, so ``ops`` is consumed lazily: gates
    further on are planned while earlier passes run. With ``ordered`` a pass
    ends at the first gate that does not fit (no gate moves ahead).
    ``native`` runs the pass selection in the library (qsim_plan_pass);
    False runs the same rule in Python (``_scan``).
    """
    if native:
        yield from _partition_native(ops, nqubits, low_bits, ordered, window, singles)
        return
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
        if native:
            chosen, deferred, tile = _scan_native(items, nqubits, low_bits, ordered)
        else:
            chosen, deferred, tile = _scan(items, set(range(low_bits)), ordered)
        if not ordered and not native:
            # also try tiles of low bits + a window of consecutive bits, and
            # keep the tile that takes the most gates: on circuits of
            # neighbouring gates a contiguous block of qubits can advance
            # several layers before its boundary stops it
            width = TILE_BITS - low_bits
            for start in range(low_bits, nqubits - width + 1):
                seed = set(range(low_bits)) | set(range(start, start + width))
                c2, d2, t2 = _scan(items, seed, False)
                if len(c2) > len(chosen):
                    chosen, deferred, tile = c2, d2, t2
        items[:] = deferred
        fill()
        if len(chosen) == 1 and singles:
            yield ("single", chosen[0].op)
            continue
        bits = sorted(tile)
        extra = [b for b in range(nqubits) if b not in tile]
        bits = sorted(bits + extra[: TILE_BITS - len(bits)])
        yield ("block", (bits, [c.op for c in chosen]))


def _masks(op, nqubits):
    """(non-diagonal target mask, mask of all bits, blockable) of a planned op
    over amplitude bits, kept on the op per register size."""
    cached = op.masks
    if cached is None or cached[0] != nqubits:
        al = 0
        for q in op.targets:
            al |= 1 << (nqubits - 1 - q)
        nd = 0 if op.kind in ("diag", "innerd") else al
        for q in op.controls:
            al |= 1 << (nqubits - 1 - q)
        cached = op.masks = (nqubits, nd, al, blockable(op))
    return cached


def _partition_native(ops, nqubits, low_bits, ordered, window, singles):
    """partition() with the pass selection in the library (qsim_plan_pass):
    the pending gates' bit masks are kept as plain integers beside them."""
    from . import _lib

    lib = _lib.load()
    source = iter(ops)
    pend, nd, al, ok = [], [], [], []

    def fill():
        while len(pend) < window:
            op = next(source, None)
            if op is None:
                return
            if op.kind == "noop":
                continue
            _, m_nd, m_al, m_ok = _masks(op, nqubits)
            pend.append(op)
            nd.append(m_nd)
            al.append(m_al)
            ok.append(m_ok)

    fill()
    tile = ctypes.c_uint64(0)
    while pend:
        if not ok[0] or nqubits < TILE_BITS:
            yield ("single", pend[0])
            del pend[0], nd[0], al[0], ok[0]
            fill()
            continue
        nitems = len(pend)
        nd_a = np.array(nd, dtype=np.uint64)
        al_a = np.array(al, dtype=np.uint64)
        ok_a = np.array(ok, dtype=np.uint8)
        taken = np.empty(nitems, dtype=np.uint8)
        rc = lib.qsim_plan_pass(nitems, nd_a.ctypes.data, al_a.ctypes.data, ok_a.ctypes.data, nqubits, low_bits,
                                TILE_BITS, 1 if ordered else 0, taken.ctypes.data, ctypes.byref(tile))
        if rc < 0:
            _lib.check(rc, lib)
        flags = taken.tolist()
        chosen = [op for op, t in zip(pend, flags) if t]
        keep = [i for i, t in enumerate(flags) if not t]
        pend[:] = [pend[i] for i in keep]
        nd[:] = [nd[i] for i in keep]
        al[:] = [al[i] for i in keep]
        ok[:] = [ok[i] for i in keep]
        fill()
        if len(chosen) == 1 and singles:
            yield ("single", chosen[0])
            continue
        tm = tile.value
        bits = [b for b in range(tm.bit_length()) if (tm >> b) & 1]
        extra = [b for b in range(nqubits) if not (tm >> b) & 1]
        bits = sorted(bits + extra[: TILE_BITS - len(bits)])
        yield ("block", (bits, chosen))


def record_bytes(nent: int, itemsize: int) -> int:
    """Bytes of one record: 32-byte header + nent complex entries, padded to 16."""
    return HEADER_BYTES + (nent * itemsize + 15) // 16 * 16


_HDR = np.dtype([("op", "u1"), ("arg", "u1"), ("nt", "u1"), ("rreq", "u1"), ("nent", "<u4"),
                 ("cmask", "<u8"), ("bits", "u1", (8,)), ("reserved", "<u8")])
assert _HDR.itemsize == HEADER_BYTES
_HDR_STRUCT = struct.Struct("<BBBBIQ8BQ")  # the same layout
assert _HDR_STRUCT.size == HEADER_BYTES


# |c^2 + s^2 - 1| up to which a 2x2 of the ROT forms runs as a rotation: the
# shears reproduce only an exact rotation, so complex128 matrices must be one
# to 1e-14; complex64 matrices hold float32 roundings of cos and sin (|c^2 +
# s^2 - 1| up to ~1.2e-7), and their rotation by atan2(s, c) differs from the
# rounded matrix by less than float32's own rounding
ROT_TOL = {np.dtype(np.complex128): 1e-14, np.dtype(np.complex64): 4e-7}


def rotation_form(m: np.ndarray, tol: float = 1e-14):
    """(mode, t, s, negated) when the 2x2 ``m`` is one of the ROT forms with
    |c^2 + s^2 - 1| <= tol, else None. negated: m = -(the form with c >= 0
    that is encoded)."""
    a, b, cc, d = np.asarray(m, dtype=np.complex128).ravel().tolist()  # Python complex: the same values
    real = a.imag == 0 and b.imag == 0 and cc.imag == 0 and d.imag == 0
    if real and a.real == d.real and b.real == -cc.real:
        mode, c, s = ROT_RY, a.real, cc.real
    elif real and a.real == -d.real and b.real == cc.real:
        mode, c, s = ROT_H, a.real, cc.real
    elif (a.real == d.real and a.imag == 0 and d.imag == 0 and b.real == 0
          and cc.real == 0 and b.imag == cc.imag):
        mode, c, s = ROT_RX, a.real, -b.imag
    else:
        return None
    if abs(c * c + s * s - 1.0) > tol:
        return None
    neg = bool(c < 0)
    if neg:
        c, s = -c, -s
    return mode, s / (1.0 + c), s, neg


def _pack_rot(buf: bytearray, arg: int, t: float, s: float, cdtype):
    """Append a ROT record: no entries; (t, s) ride in the header, in the
    fields a ROT does not use (cmask and bits), so the kernel has them with
    the header it reads one record ahead: complex128 t = bytes 8-15, s =
    bytes 16-23 (float64); complex64 t = bytes 8-11, s = bytes 16-19 (float32)."""
    if np.dtype(cdtype) == np.complex128:
        tb, sb = struct.pack("<d", t), struct.pack("<d", s)
    else:
        tb, sb = struct.pack("<f", t) + bytes(4), struct.pack("<f", s) + bytes(4)
    buf += bytes((OP_ROT, arg, 0, 0)) + bytes(4) + tb + sb + bytes(8)


def _pack(buf: bytearray, op: int, arg: int, cmask: int, nt: int, bits: Sequence[int],
          data, cdtype, rreq: int = 0):
    """Append one variable-length record: header (u8 op, u8 arg, u8 nt, u8 rreq,
    u32 nent, u64 cmask, u8 bits[8], u64 0) + nent complex entries."""
    if data is None:
        nent, raw = 0, b""
    else:
        mat = np.asarray(data, dtype=cdtype).ravel()
        nent, raw = mat.size, mat.tobytes()
    b = list(bits) + [0] * (8 - len(bits))
    buf += _HDR_STRUCT.pack(op, arg, nt, rreq, nent, cmask, *b, 0)
    buf += raw
    pad = (-len(raw)) % 16
    if pad:
        buf += bytes(pad)


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
_ONES16 = np.ones(16, dtype=np.complex128)


@functools.lru_cache(maxsize=None)
def _reg_phase_pattern(tregs: Tuple[int, ...], cregs: Tuple[int, ...]):
    """For a diagonal on register bits ``tregs`` (first = most significant
    entry bit) under register controls ``cregs``: which of the 16 register
    entries it touches and the diagonal entry each one takes."""
    ok = np.ones(16, dtype=bool)
    for c in cregs:
        ok &= _RVALS[c] == 1
    di = np.zeros(16, dtype=np.int64)
    for t in tregs:
        di = (di << 1) | _RVALS[t]
    ok.setflags(write=False)
    di.setflags(write=False)
    return ok, di
_RFLIP = [np.arange(16) ^ (1 << j) for j in range(4)]  # register r -> r with bit j flipped
# registers a REG_PHASE record with dependency mask D multiplies: r & D != 0
_REG_DEP = [np.array([(r & d) != 0 for r in range(16)]) for d in range(16)]


def _schedule(items, npred, succ, ready, runs_here):
    """List scheduling of one pass. Yields ("map", m) when the register
    mapping must change to m and ("op", i) for each item in issue order.
    ``runs_here(i, m)`` says whether item i runs in mapping m without a change;
    items that need a register bit (one-qubit dense gates) pick the mapping,
    looking ahead at how many of them each mapping can run."""
    cur = 0

    def reg_gate(i):
        op = items[i].op
        return (op.kind == "dense" and len(op.targets) == 1) or op.kind == "inner1"

    def reach(mapping):
        left = list(npred)
        frontier = [i for i in ready if runs_here(i, mapping)]
        done = 0
        while frontier:
            i = frontier.pop()
            if reg_gate(i):
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
            if any(reg_gate(i) for i in ready):
                counts = {m: reach(m) for m in (0, 1, 2) if m != cur}
                cur = max(counts, key=lambda m: (counts[m], -m))
                yield "map", cur
                pick = next(i for i in ready if runs_here(i, cur))
            else:
                pick = ready[0]
        ready.remove(pick)
        yield "op", pick
        for j in succ[pick]:
            npred[j] -= 1
            if npred[j] == 0:
                ready.append(j)


_KIND_CODE = {"noop": 0, "diag": 1, "swap": 2, "dense": 3}
_DESC_INTS = 44  # kind, ntargets, ncontrols, -, targets[5], controls[35] (qsim_compile_pass)


def _desc(op):
    """(int32 row, complex128 entries or None) describing a planned op to the
    native compiler; kept on the op (its fields never change after planning)."""
    cached = op.native
    if cached is None:
        row = np.zeros(_DESC_INTS, dtype=np.int32)
        kind = _KIND_CODE[op.kind]
        nt, nc = len(op.targets), len(op.controls)
        if nt > 5 or nc > 35:
            raise ValueError(f"a blocked pass takes at most 5 targets and 35 controls per gate, got {nt} and {nc}")
        row[0], row[1], row[2] = kind, nt, nc
        row[4:4 + nt] = op.targets
        row[9:9 + nc] = op.controls
        data = None
        if op.data is not None and op.kind in ("diag", "dense"):
            data = np.ascontiguousarray(op.data, dtype=np.complex128).ravel()
        cached = op.native = (row, data)
    return cached


def _compile_pass_native(tile_bits, ops, nqubits, cdtype, ordered):
    from . import _lib

    lib = _lib.load()
    rows, datas, offs = [], [], []
    off = 0
    for op in ops:
        row, data = _desc(op)
        rows.append(row)
        if data is None:
            offs.append(-1)
        else:
            offs.append(off)
            datas.append(data)
            off += data.size
    desc = np.concatenate(rows) if rows else np.zeros(_DESC_INTS, dtype=np.int32)
    data = np.concatenate(datas) if datas else np.zeros(1, dtype=np.complex128)
    offs = np.asarray(offs if offs else [-1], dtype=np.int64)
    code = _lib.C64 if np.dtype(cdtype) == np.complex64 else _lib.C128
    cap = lib.qsim_compile_pass_capacity(len(ops), code)
    out = np.empty(cap, dtype=np.uint8)
    nbytes, count, switches = ctypes.c_size_t(0), ctypes.c_int(0), ctypes.c_int(0)
    rc = lib.qsim_compile_pass(nqubits, _lib.int_array(tile_bits), len(ops), desc.ctypes.data, data.ctypes.data,
                               offs.ctypes.data, code, 1 if ordered else 0, out.ctypes.data, cap,
                               ctypes.byref(nbytes), ctypes.byref(count), ctypes.byref(switches))
    _lib.check(rc, lib)
    return out[: nbytes.value].tobytes(), count.value, switches.value


def compile_pass(tile_bits: Sequence[int], ops: Sequence, nqubits: int, cdtype,
                 ordered: bool = False, native: bool = True) -> Tuple[bytes, int, int]:
    """Instruction records of one blocked pass; returns (bytes, count, map switches).

    With ``ordered`` the instructions keep the order of ``ops`` and use only
    the exact per-gate forms (DIAG, DENSE1, DENSE_LDS, MAP). ``native`` runs
    the compiler in the library (qsim_compile_pass, the same records);
    False runs this Python restatement of it."""
    if native:
        return _compile_pass_native(tile_bits, ops, nqubits, cdtype, ordered)
    local = {b: i for i, b in enumerate(tile_bits)}
    items = [_Item(op, nqubits, i) for i, op in enumerate(ops)]
    npred, succ = _dependencies(items, ordered)
    ready = [i for i in range(len(items)) if npred[i] == 0]
    buf = bytearray()
    count = 0
    switches = 0
    cur = 0
    pending = None  # REG_PHASE table being folded (complex128, 16 entries)
    reg_first = None  # (entry offset, table) of the pass's first REG_PHASE record (all registers)
    gfac = np.ones(1, dtype=np.complex128)  # global factor taken out of later tables
    sign = False  # net -1 left by ROT records of gates with c < 0
    absorb = None  # (byte offset, entries) of the first unconditional linear record
    qpend = None  # register bit with a pending LANE_REGPHASE factor

    def emit(op, arg, cmask, nt, bits, data, rreq=0):
        nonlocal count, absorb
        if absorb is None and cmask == 0 and rreq == 0 and data is not None and op in (
                OP_DENSE1, OP_DENSE1_REAL, OP_REG_PHASE, OP_LANE_PHASE, OP_DIAG):
            absorb = (len(buf) + HEADER_BYTES, np.asarray(data).size)
        _pack(buf, op, arg, cmask, nt, bits, data, cdtype, rreq)
        count += 1

    def qflush(j=None):
        """Apply the pending per-lane factor (before anything that does not
        commute with it; with j only before a register gate on its bit)."""
        nonlocal qpend
        if qpend is None or (j is not None and j != qpend):
            return
        emit(OP_QFLUSH, qpend, 0, 0, (), None)
        qpend = None

    def flush_reg(j=None):
        """Emit the folded table; before a register gate on register bit j
        only when the table depends on that bit (otherwise the two commute).

        The pass's first table multiplies all 16 registers. A later table T
        with |T[0]| = 1 is split into the global factor T[0] (collected and
        folded into the first table at the end: a scalar commutes with every
        record) and T / T[0], whose entries are exactly 1 on the registers
        where the bits T depends on are all 0: the record's arg is that bit
        mask D and the kernel multiplies only the registers r with r & D != 0
        (none at all when T is constant)."""
        nonlocal pending, reg_first, gfac
        if pending is None:
            return
        if j is not None and np.array_equal(pending, pending[_RFLIP[j]]):
            return
        if reg_first is None:
            reg_first = (len(buf) + HEADER_BYTES, pending.copy())
            emit(OP_REG_PHASE, 0, 0, 0, (), pending)
        else:
            t0 = complex(pending[0])
            mag = t0.real * t0.real + t0.imag * t0.imag
            if abs(mag - 1.0) <= 1e-12:
                dep = sum(1 << k for k in range(4) if not np.array_equal(pending, pending[_RFLIP[k]]))
                gfac = gfac * np.array([t0])
                if dep:
                    inv = complex(t0.real / mag, -t0.imag / mag)
                    table = np.where(_REG_DEP[dep], pending * inv, _ONES16)
                    emit(OP_REG_PHASE, dep, 0, 0, (), table)
            else:
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
        nonlocal pending, qpend
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
                table = _ONES16 if pending is None else pending
                ok, di = _reg_phase_pattern(tuple(rb.index(b) for b in tb), tuple(rb.index(c) for c in cb))
                pending = np.where(ok, table * d[di], table)
                return
        cm, rreq = split_cond(op, cur)
        enc = [(REG_FLAG | rb.index(b)) if b in rb else b for b in tb]
        if not ordered and nt == 0 and d[0] == -1:
            emit(OP_NEG, 0, cm, 0, (), None, rreq)
        elif not ordered and nt == 0 and rreq in (1, 2, 4, 8):
            j = rreq.bit_length() - 1
            if qpend != j:
                qflush()
                qpend = j
            emit(OP_LANE_REGPHASE, j, cm, 0, (), d[:1])
        else:
            emit(OP_DIAG, 0, cm, nt, enc, d, rreq)

    for event, arg in _schedule(items, npred, succ, ready, runs_here):
        if event == "map":
            flush_reg()
            qflush()
            emit(OP_MAP, arg, 0, 0, (), None)
            switches += 1
            cur = arg
            continue
        pick = arg
        op = items[pick].op
        if op.kind == "diag":
            emit_diag(op)
        elif op.kind == "dense" and len(op.targets) == 1 and runs_here(pick, cur):
            lb = local[nqubits - 1 - op.targets[0]]
            flush_reg(_REG_GROUPS[cur].index(lb))
            qflush(_REG_GROUPS[cur].index(lb))
            cm, rreq = split_cond(op, cur)
            m = np.asarray(op.data)
            j = _REG_GROUPS[cur].index(lb)
            rot = None if (ordered or cm or rreq) else rotation_form(m, ROT_TOL[np.dtype(cdtype)])
            if rot is not None:
                mode, t, sn, neg = rot
                _pack_rot(buf, j | (mode << 2), t, sn, cdtype)
                count += 1
                sign ^= neg
                continue
            real = (not ordered) and not np.any(np.imag(m))
            emit(OP_DENSE1_REAL if real else OP_DENSE1, j, cm, 0, (), m.ravel()[:4], rreq)
        else:
            flush_reg()
            qflush()
            mat = _SWAP if op.kind == "swap" else op.data
            lbits = [local[nqubits - 1 - q] for q in op.targets]
            k = len(lbits)
            cm = sum(1 << (nqubits - 1 - q) for q in op.controls)
            emit(OP_DENSE_LDS, k, cm, 0, lbits + [0] * (4 - k) + sorted(lbits), mat)
    flush_reg()
    qflush()
    if reg_first is not None and gfac[0] != 1:
        off, table = reg_first
        view = np.frombuffer(buf, dtype=cdtype, count=16, offset=off)
        view[:] = (table * gfac[0]).astype(cdtype)
        del view
    if sign:
        # a global -1 commutes with every record: scale one unconditional
        # linear record's entries (all of them), else flip every sign at the end
        if absorb is not None:
            off, nent = absorb
            view = np.frombuffer(buf, dtype=cdtype, count=nent, offset=off)
            view *= -1
            del view
        else:
            emit(OP_NEG, 0, 0, 0, (), None)
    return bytes(buf), count, switches


OP_INNER1, OP_INNER_DIAG = 8, 9


class BatchOp:
    """A planned operation for a batch of B parameter sets: the structure
    (kind, targets, controls) is shared, ``data`` holds one matrix or diagonal
    per set (leading axis B; a broadcast view when all sets share it).
    kind: "dense" (one target), "diag", "inner1" / "innerd" (the adjoint
    sweep's Re <lambda| M |psi> for a 2x2 / diagonal M, accumulated into
    gradient slot ``slot``)."""

    __slots__ = ("kind", "targets", "controls", "data", "slot", "masks")

    def __init__(self, kind, targets, controls, data, slot=-1):
        self.masks = None  # bit masks for the pass selection (_masks)
        self.kind = kind
        self.targets = tuple(targets)
        self.controls = tuple(controls)
        self.data = data
        self.slot = slot


def compile_pass_batched(tile_bits: Sequence[int], ops: Sequence, nqubits: int, cdtype,
                         nbatch: int, entry_map: list = None) -> Tuple[np.ndarray, int, list]:
    """Records of one blocked pass for a batch of parameter sets.

    Returns (programs [B, nbytes] uint8, record count, gradient slot of each
    INNER record in program order). Every set's program has the same records
    (same layout and headers), only the entries differ. One record per
    operation: DENSE1_REAL (when the 2x2 is real for every set), DENSE1,
    DIAG, NEG (a -1 on a controlled subspace for every set), LANE_PHASE,
    INNER1, INNER_DIAG and MAP. With ``entry_map`` (a list)
    it receives (byte offset of the entries in a program, index of the
    operation in ``ops``, entry count) for every record that has entries:
    the entries are the operation's data row, flattened."""
    local = {b: i for i, b in enumerate(tile_bits)}
    items = [_Item(op, nqubits, i) for i, op in enumerate(ops)]
    npred, succ = _dependencies(items, False)
    ready = [i for i in range(len(items)) if npred[i] == 0]
    isz = np.dtype(cdtype).itemsize
    headers, entries = [], []  # per record: header bytes, [B, nent] entries or None
    sources = []  # per record: index of the operation in ``ops``
    slots = []
    cur = 0
    src = -1

    def emit(op, arg, cmask, nt, bits, data, rreq=0):
        nent = 0 if data is None else data.shape[-1]
        b = list(bits) + [0] * (8 - len(bits))
        headers.append(_HDR_STRUCT.pack(op, arg, nt, rreq, nent, cmask, *b, 0))
        sources.append(src)
        if data is None:
            entries.append(None)
        else:
            d = np.broadcast_to(np.asarray(data, dtype=cdtype), (nbatch, nent))
            entries.append(d)

    def reg_bits(mapping):
        return [tile_bits[lb] for lb in _REG_GROUPS[mapping]]

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
        if op.kind in ("diag", "innerd"):
            return True
        if op.kind in ("dense", "inner1") and len(op.targets) == 1:
            return local[nqubits - 1 - op.targets[0]] in _REG_GROUPS[mapping]
        return False

    def flat(data, nent):
        return np.asarray(data).reshape(nbatch, nent)

    for event, arg in _schedule(items, npred, succ, ready, runs_here):
        if event == "map":
            emit(OP_MAP, arg, 0, 0, (), None)
            cur = arg
            continue
        op = items[arg].op
        src = arg
        rb = reg_bits(cur)
        if op.kind in ("diag", "innerd"):
            tb = [nqubits - 1 - q for q in op.targets]
            cb = [nqubits - 1 - q for q in op.controls]
            nt = len(tb)
            d = flat(op.data, 2**nt)
            if op.kind == "diag" and all(b not in rb for b in tb + cb):
                emit(OP_LANE_PHASE, 0, sum(1 << b for b in cb), nt, tb, d)
                continue
            cm, rreq = split_cond(op, cur)
            enc = [(REG_FLAG | rb.index(b)) if b in rb else b for b in tb]
            if op.kind == "diag" and nt == 0 and np.all(d == -1):
                emit(OP_NEG, 0, cm, 0, (), None, rreq)  # CZ-type: a sign flip, no multiply
            elif op.kind == "diag":
                emit(OP_DIAG, 0, cm, nt, enc, d, rreq)
            else:
                enc = enc + [0] * (6 - nt) + [op.slot & 255, op.slot >> 8]
                emit(OP_INNER_DIAG, 0, cm, nt, enc, d, rreq)
                slots.append(op.slot)
            continue
        if op.kind not in ("dense", "inner1") or len(op.targets) != 1:
            raise ValueError(f"batched passes take one-qubit dense and diagonal operations, not {op.kind}")
        lb = local[nqubits - 1 - op.targets[0]]
        j = _REG_GROUPS[cur].index(lb)
        cm, rreq = split_cond(op, cur)
        m = flat(op.data, 4)
        if op.kind == "inner1":
            emit(OP_INNER1, j, cm, 0, [op.slot & 255, op.slot >> 8], m, rreq)
            slots.append(op.slot)
        else:
            real = not np.any(np.imag(m))
            emit(OP_DENSE1_REAL if real else OP_DENSE1, j, cm, 0, (), m, rreq)
    # pack: the same headers for every set, entries per set, records padded to 16
    cols = []
    off = 0
    for hdr, ent, i in zip(headers, entries, sources):
        cols.append(np.broadcast_to(np.frombuffer(hdr, dtype=np.uint8), (nbatch, HEADER_BYTES)))
        off += HEADER_BYTES
        if ent is not None:
            raw = np.ascontiguousarray(ent).view(np.uint8).reshape(nbatch, -1)
            pad = (-raw.shape[1]) % 16
            cols.append(raw)
            if entry_map is not None:
                entry_map.append((off, i, ent.shape[-1]))
            off += raw.shape[1] + pad
            if pad:
                cols.append(np.zeros((nbatch, pad), dtype=np.uint8))
    prog = np.concatenate(cols, axis=1) if cols else np.zeros((nbatch, 0), dtype=np.uint8)
    return np.ascontiguousarray(prog), len(headers), slots


# FP64/FP32 floating-point operations per amplitude of the tile that each
# record form performs (multiply = 1, fused multiply-add = 2), as the kernel
# executes them: real 2x2: 2 mul + 2 fma per component; rotation: 3 fma per
# real pair; complex 2x2: a complex
# multiply (6) and a complex multiply-add (8); diagonal / table: a complex
# multiply; a per-lane factor: one complex multiply per 16 amplitudes;
# dense 2^k block: 6 + 8 (2^k - 1).
_FLOPS = {OP_DENSE1_REAL: 6.0, OP_ROT: 6.0, OP_LANE_REGPHASE: 6.0 / 16, OP_QFLUSH: 3.0, OP_DENSE1: 14.0, OP_DIAG: 6.0, OP_REG_PHASE: 6.0, OP_LANE_PHASE: 6.0 / 16,
          OP_NEG: 0.0, OP_MAP: 0.0}


def program_flops(program: bytes, count: int, itemsize: int) -> float:
    """Floating-point operations per amplitude of one blocked pass's records."""
    off, total = 0, 0.0
    for _ in range(count):
        op, arg = program[off], program[off + 1]
        nent = int.from_bytes(program[off + 4: off + 8], "little")
        if op == OP_DENSE_LDS:
            total += 6.0 + 8.0 * (2**arg - 1)
        elif op == OP_REG_PHASE and arg & 15:
            total += 6.0 * int(_REG_DEP[arg & 15].sum()) / 16
        else:
            total += _FLOPS.get(op, 0.0)
        off += record_bytes(nent, itemsize)
    return total


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
