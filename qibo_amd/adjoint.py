"""Adjoint-method gradients of expectation values, and their autograd wrapper.

For E(theta) = <psi|H|psi> with |psi> = U_G ... U_1 |0>, the adjoint sweep
(no stored intermediate states, gates un-applied in reverse):

    psi <- U_G ... U_1 |0>;   lambda <- H psi;   E = Re <psi|lambda>
    for j = G .. 1:
        psi <- U_j^dagger psi                     (psi is now U_{j-1}...U_1|0>)
        for each parameter theta of gate j:
            dE/dtheta += 2 Re <lambda| dU_j/dtheta |psi>
        lambda <- U_j^dagger lambda

Every step is a kernel of libqsim_mi355x.so: gate passes (the forward state
uses blocked passes), ``qsim_apply_pauli_sum`` for H psi, and
``qsim_inner_matrix`` for the derivative terms, all in FP64 accumulation.

The reference differentiates by re-execution: ``parameter_shift`` and
``finite_differences`` (src/qibo/derivative.py:7-226) run two full
simulations per parameter; this sweep costs about three state passes per
gate for the whole gradient.

Derivative matrices: for exp(-i theta G / 2) rotations (RX, RY, RZ, CRX, CRY,
CRZ, RXX, RYY, RZZ, RZX) dU/dtheta = -i/2 G U; for U1 / CU1 it is
diag(0, i e^{i theta}) on the controlled block. Every other parametrized gate
of the package has matrix entries that depend on each parameter x through
one frequency k only (a constant plus a cos(k x) and sin(k x) or e^{+-i k x}
term: k = 1/2 for half angles, 1 for whole angles; table _FREQUENCIES), so
the shift rule on the matrix is its exact derivative:
dU/dx = k (U(x + pi/(2k)) - U(x - pi/(2k))) / 2. A gate outside the table
falls back to a central difference of the host matrix (step 1e-7).
"""

from typing import List, Sequence, Tuple

import numpy as np
import torch

from . import _lib, blocking, gates, kinds
from .backend import plan_matrix
from .blocking import BatchOp
from .config import raise_error

_X = np.array([[0, 1], [1, 0]], dtype=complex)
_Y = np.array([[0, -1j], [1j, 0]], dtype=complex)
_Z = np.array([[1, 0], [0, -1]], dtype=complex)
_GENERATORS = {
    "RX": _X, "RY": _Y, "RZ": _Z,
    "RXX": np.kron(_X, _X), "RYY": np.kron(_Y, _Y), "RZZ": np.kron(_Z, _Z), "RZX": np.kron(_Z, _X),
}
_CONTROLLED_GEN = {"CRX": _X, "CRY": _Y, "CRZ": _Z}
# the one frequency with which each parameter enters the gate's matrix
# entries (qibo_amd.matrices, after npmatrices.py): 1/2 for half angles
# (cos(theta / 2), e^{i (phi + lam) / 2}), 1 for whole angles
_FREQUENCIES = {
    "PRX": (0.5, 1.0), "GPI": (1.0,), "GPI2": (1.0,), "U2": (0.5, 0.5), "U3": (0.5, 0.5, 0.5),
    "U1q": (0.5, 1.0), "CU2": (0.5, 0.5), "CU3": (0.5, 0.5, 0.5), "fSim": (1.0, 1.0),
    "GeneralizedfSim": (None, 1.0), "RXXYY": (0.5,), "MS": (1.0, 1.0, 0.5), "GIVENS": (1.0,),
    "RBS": (1.0,), "DEUTSCH": (1.0,), "Align": (1.0,),
}


def _gate_qubits(gate) -> Tuple[Tuple[int, ...], Tuple[int, ...]]:
    """(targets, controls) that the gate's matrix acts on / is controlled by."""
    if gate.is_controlled_by:
        return tuple(gate.target_qubits), tuple(gate.control_qubits)
    return tuple(gate.qubits), ()


def derivative_matrices(backend, gate) -> List[np.ndarray]:
    """dU/dtheta_k for each parameter of ``gate`` (matrix over _gate_qubits)."""
    name = gate.__class__.__name__
    u = np.asarray(gate.matrix(backend), dtype=np.complex128)
    if name in _GENERATORS:
        return [-0.5j * _GENERATORS[name] @ u]
    if name in _CONTROLLED_GEN:
        d = np.zeros_like(u)
        d[2:, 2:] = -0.5j * _CONTROLLED_GEN[name] @ u[2:, 2:]
        return [d]
    if name == "U1":
        return [np.diag([0, 1j * u[1, 1]])]
    if name == "CU1":
        return [np.diag([0, 0, 0, 1j * u[3, 3]])]
    params = list(gate.parameters)
    freqs = _FREQUENCIES.get(name)
    if freqs is not None and len(freqs) == len(params):
        out = []
        for k, f in enumerate(freqs):
            if f is None:  # a matrix-valued parameter (GeneralizedfSim's unitary): no scalar derivative
                out.append(np.zeros_like(u))
                continue
            shift = np.pi / (2 * f)
            p_plus = list(params)
            p_minus = list(params)
            p_plus[k] = params[k] + shift
            p_minus[k] = params[k] - shift
            gate.parameters = p_plus
            up = np.asarray(gate.matrix(backend), dtype=np.complex128)
            gate.parameters = p_minus
            um = np.asarray(gate.matrix(backend), dtype=np.complex128)
            out.append(0.5 * f * (up - um))
        gate.parameters = params
        return out
    # a gate outside the table: central difference of the host matrix
    out = []
    eps = 1e-7
    for k in range(len(params)):
        p_plus = list(params)
        p_minus = list(params)
        p_plus[k] = params[k] + eps
        p_minus[k] = params[k] - eps
        gate.parameters = p_plus
        up = np.asarray(gate.matrix(backend), dtype=np.complex128)
        gate.parameters = p_minus
        um = np.asarray(gate.matrix(backend), dtype=np.complex128)
        out.append((up - um) / (2 * eps))
    gate.parameters = params
    return out


def _dagger_op(backend, gate):
    targets, controls = _gate_qubits(gate)
    m = np.asarray(gate.matrix(backend))
    return plan_matrix(np.conj(m.T), targets, controls)


def expectation_and_gradient(backend, circuit, hamiltonian, initial_state=None):
    """(E, dE/dtheta) for the circuit's trainable parameters (flat order of
    ``circuit.get_parameters()``), by one forward pass and one adjoint sweep."""
    if circuit.density_matrix:
        raise_error(NotImplementedError, "adjoint gradients are for state vectors")
    n = circuit.nqubits
    xm, zm, ny, coeffs = hamiltonian.pauli_tables()
    queue = [g for g in circuit.queue if not kinds.is_measurement(g)]
    if initial_state is None:
        psi = backend.zero_state(n)
    else:
        psi = backend.cast(initial_state, copy=True)
    psi = backend.run_queue(queue, psi, n)
    lam = backend.apply_pauli_sum(psi, n, xm, zm, ny, coeffs)
    energy = backend.expectation_pauli(psi, n, xm, zm, ny, coeffs) + hamiltonian.constant
    slots = {}
    k = 0
    for g in circuit.trainable_gates:
        slots[id(g)] = (k, g.nparams)
        k += g.nparams
    grads = [None] * k
    for gate in reversed(queue):
        dop = _dagger_op(backend, gate)
        backend.apply_op(dop, psi, n)
        if id(gate) in slots and gate.trainable:
            first, count = slots[id(gate)]
            targets, controls = _gate_qubits(gate)
            for j, dm in enumerate(derivative_matrices(backend, gate)[:count]):
                grads[first + j] = 2.0 * backend.inner_matrix(lam, psi, n, targets, controls, dm)
        backend.apply_op(dop, lam, n)
    zero = torch.zeros((), dtype=torch.float64, device=psi.device)
    grad = torch.stack([g if g is not None else zero for g in grads]) if grads else zero.reshape(0)
    return energy, grad


def batched_expectation_and_gradient(backend, circuit, hamiltonian, params) -> Tuple[torch.Tensor, torch.Tensor]:
    """Energies [B] and gradients [B, P] for a batch of parameter sets [B, P].

    Circuits of at least 12 qubits whose gates the batched blocked sweep
    covers (one-qubit gates, controlled one-qubit gates, diagonal gates;
    trainable gates RX/RY/RZ/CRX/CRY/CRZ/U1/CU1/RZZ) run all sets at once:
    forward and reverse sweeps as batched blocked passes, one launch per pass
    for the whole batch. Other circuits run set by set."""
    params = np.asarray(params.detach().cpu().numpy() if isinstance(params, torch.Tensor) else params,
                        dtype=np.float64)
    if params.ndim == 1:
        params = params[None, :]
    try:
        return blocked_expectation_and_gradient(backend, circuit, hamiltonian, params)
    except _Unsupported:
        pass
    energies, grads = [], []
    for row in params:
        circuit.set_parameters(row)
        e, g = expectation_and_gradient(backend, circuit, hamiltonian)
        energies.append(e)
        grads.append(g)
    return torch.stack(energies), torch.stack(grads)


# ---------------------------------------------------------------------------
# batched blocked sweep
# ---------------------------------------------------------------------------
class _Unsupported(Exception):
    """The batched blocked sweep does not cover this circuit."""


def _controlled(block: np.ndarray) -> np.ndarray:
    """[[I, 0], [0, block]] for a batch of 2x2 blocks (one control, first)."""
    out = np.zeros(block.shape[:-2] + (4, 4), dtype=block.dtype)
    out[..., 0, 0] = out[..., 1, 1] = 1
    out[..., 2:, 2:] = block
    return out


def _batched_rotation(name: str, th: np.ndarray) -> np.ndarray:
    """Matrices of a parametrized gate for angles ``th`` [B] -> [B, d, d], by the
    formulas of qibo_amd.matrices (npmatrices.py:80-135, 160-205, 269-272)."""
    th = np.asarray(th, dtype=np.float64)
    z = np.zeros_like(th) + 0j
    if name in ("RX", "CRX"):
        c, s = np.cos(th / 2.0) + 0j, -1j * np.sin(th / 2.0)
        m = np.array([[c, s], [s, c]])
    elif name in ("RY", "CRY"):
        c, s = np.cos(th / 2.0) + 0j, np.sin(th / 2.0) + 0j
        m = np.array([[c, -s], [s, c]])
    elif name in ("RZ", "CRZ"):
        ph = np.exp(0.5j * th)
        m = np.array([[np.conj(ph), z], [z, ph]])
    elif name in ("U1", "CU1"):
        m = np.array([[z + 1, z], [z, np.exp(1j * th)]])
    elif name == "RZZ":
        ph = np.exp(0.5j * th)
        cph = np.conj(ph)
        m = np.array([[cph, z, z, z], [z, ph, z, z], [z, z, ph, z], [z, z, z, cph]])
    else:
        raise _Unsupported(name)
    m = np.moveaxis(m, -1, 0)
    return _controlled(m) if name.startswith("C") else m


# M = dU/dtheta U^dagger of the trainable gates the sweep covers, on the
# gate's target qubits (controls of the gate stay conditions)
_HALF_I = -0.5j
_INNER = {
    "RX": ("inner1", _HALF_I * _X), "RY": ("inner1", _HALF_I * _Y), "RZ": ("innerd", _HALF_I * np.array([1, -1])),
    "CRX": ("inner1", _HALF_I * _X), "CRY": ("inner1", _HALF_I * _Y),
    "CRZ": ("innerd", _HALF_I * np.array([1, -1])),
    "U1": ("innerd", np.array([0, 1j])), "CU1": ("innerd", np.array([0, 1j])),
    "RZZ": ("innerd", _HALF_I * np.array([1, -1, -1, 1])),
}


def _batched_ops(backend, circuit, params: np.ndarray, sources=None, groups=None):
    """Forward and reverse BatchOp lists and the trainable-parameter count.

    With ``sources`` and ``groups`` (dicts) it also records, for the ops of
    trainable gates, id(op) -> (gate name, dagger, index among that name's
    gates) and, per gate name, the parameter slots in circuit order."""
    queue = [g for g in circuit.queue if not kinds.is_measurement(g)]
    nb = params.shape[0]
    slot_of, k = {}, 0
    for g in circuit.trainable_gates:
        if g.nparams != 1:
            raise _Unsupported("multi-parameter trainable gate")
        slot_of[id(g)] = k
        k += 1
    if k != params.shape[1]:
        raise _Unsupported("parameter count")
    cdt = np.complex128 if backend.dtype == "complex128" else np.complex64
    fwd, rev = [], []
    fixed = {}  # plans of parameter-free gates, by matrix and qubits
    for g in queue:
        if kinds.runs_outside_passes(g) or kinds.is_fused(g):
            raise _Unsupported("special gate")
        targets, controls = _gate_qubits(g)
        qubits = tuple(g.qubits) if not g.is_controlled_by else tuple(g.target_qubits)
        name = g.__class__.__name__
        if id(g) in slot_of and g.trainable:
            op, dag = _rotation_ops(g, name, params[:, slot_of[id(g)]], cdt)
            if sources is not None:
                members = groups.setdefault(name, [])
                sources[id(op)] = (name, False, len(members))
                sources[id(dag)] = (name, True, len(members))
                members.append(slot_of[id(g)])
        else:
            # the same matrix for every set: plan it once, share its data
            m = np.asarray(g.matrix(backend), dtype=cdt)
            key = (m.tobytes(), qubits, tuple(controls))
            if key not in fixed:
                o1 = plan_matrix(m, qubits, controls)
                o2 = plan_matrix(np.conj(m.T), qubits, controls)
                fixed[key] = tuple(None if o.kind == "noop" else _shared(o, nb) for o in (o1, o2))
            op, dag = fixed[key]
        if op is not None:
            fwd.append(op)
        step = []
        if id(g) in slot_of and g.trainable:
            kind, mm = _INNER[name]
            mm = np.asarray(mm, dtype=cdt).reshape(1, -1)
            step.append(BatchOp(kind, g.target_qubits, g.control_qubits, np.broadcast_to(mm, (nb, mm.shape[1])),
                                slot_of[id(g)]))
        if dag is not None:
            step.append(dag)
        rev.append(step)
    reverse = [op for step in reversed(rev) for op in step]
    return fwd, reverse, k


def _shared(op, nb):
    if op.kind not in ("dense", "diag") or (op.kind == "dense" and len(op.targets) != 1):
        raise _Unsupported(op.kind)
    d = np.asarray(op.data)
    return BatchOp(op.kind, op.targets, op.controls, np.broadcast_to(d, (nb,) + d.shape))


def _rotation_data(name, th, cdt):
    """Data rows [N, nent] of a trainable rotation's forward operation and of
    its dagger for angles th [N]: the 2x2 row-major for X/Y rotations, the
    diagonal for Z rotations, e^{i theta} alone for U1-type phases."""
    mats = _batched_rotation(name, th).astype(cdt)
    if name.startswith("C"):
        mats = mats[:, 2:, 2:]  # the controlled block
    if name in ("RX", "RY", "CRX", "CRY"):
        return mats.reshape(-1, 4), np.conj(np.swapaxes(mats, 1, 2)).reshape(-1, 4)
    d = np.diagonal(mats, axis1=1, axis2=2)
    if name in ("U1", "CU1"):
        d = d[:, 1:]
    return d, np.conj(d)


def _rotation_ops(g, name, th, cdt):
    """Forward and dagger BatchOps of a trainable rotation gate for angles th [B],
    in the form plan_matrix gives each set's matrix (dense 2x2 for X/Y
    rotations, the diagonal for Z rotations, U1-type phases as a phase on the
    subspace where target and controls are 1)."""
    t = g.target_qubits
    cq = tuple(g.control_qubits)
    d, dd = _rotation_data(name, th, cdt)
    if name in ("RX", "RY", "CRX", "CRY"):
        return BatchOp("dense", t, cq, d.reshape(-1, 2, 2)), BatchOp("dense", t, cq, dd.reshape(-1, 2, 2))
    if name in ("U1", "CU1"):  # diag(1, e^{i theta}): the target joins the controls
        return BatchOp("diag", (), cq + t, d), BatchOp("diag", (), cq + t, dd)
    return BatchOp("diag", t, cq, d), BatchOp("diag", t, cq, dd)


# A generic angle: the record form chosen for a trainable gate at it (complex
# or real 2x2) is valid at every angle (RY-type matrices are real for all
# angles, RX-type ones are complex at this one).
_PROBE_ANGLE = 1.0


class _SweepPass:
    """One batched pass of a sweep: tile bits, record count, gradient slots,
    the program row with every parameter-free entry in place, and the
    positions (in entries of the complex dtype) that take columns ``cols`` of
    the per-set data table."""

    __slots__ = ("bits", "count", "slots", "row", "pos", "cols", "dev_slots")

    def __init__(self, bits, count, slots, row, pos, cols):
        self.bits, self.count, self.slots = list(bits), count, list(slots)
        self.row, self.pos, self.cols = row, pos, cols
        self.dev_slots = {}  # device -> int32 tensor of ``slots``


class SweepTemplate:
    """The forward and reverse passes of the batched adjoint sweep of one
    circuit structure, compiled once.

    The records, their order and every parameter-free entry do not depend on
    the trainable angles, so a new batch of parameter sets only computes the
    per-set data table (``data``: one column block per trainable gate and
    direction, all gates of a kind at once) and scatters it into copies of
    the program rows (``programs``). Raises _Unsupported for circuits the
    batched sweep does not cover."""

    def __init__(self, backend, circuit, cdt):
        self.cdt = np.dtype(cdt)
        n = circuit.nqubits
        slot_of = {}
        for k, g in enumerate(circuit.trainable_gates):
            if g.nparams != 1:
                raise _Unsupported("multi-parameter trainable gate")
            slot_of[id(g)] = k
        self.nparams = len(slot_of)
        probe = np.full((1, self.nparams), _PROBE_ANGLE)
        groups = {}  # gate name -> parameter slots, in circuit order
        src = {}  # id(op) -> (gate name, dagger, index in the group)
        fwd, rev, _ = _batched_ops(backend, circuit, probe, src, groups)
        # columns of the data table: per gate name, the forward block then the dagger block
        self.groups, col = [], 0
        for name, slots in groups.items():
            nent = _rotation_data(name, probe[:, 0], self.cdt)[0].shape[1]
            g = len(slots)
            self.groups.append((name, np.array(slots, dtype=np.int64), nent, col, col + g * nent))
            col += 2 * g * nent
        self.ncols = col
        base = {name: (nent, c0, c1) for name, _, nent, c0, c1 in self.groups}

        def first_col(op):
            name, dagger, i = src[id(op)]
            nent, c0, c1 = base[name]
            return (c1 if dagger else c0) + i * nent

        low = 3 if self.cdt == np.complex128 else 4
        isz = self.cdt.itemsize
        self.fwd, self.rev = [], []
        for ops, out in ((fwd, self.fwd), (rev, self.rev)):
            for kind, payload in blocking.partition(ops, n, low, singles=False):
                if kind != "block":
                    raise _Unsupported("single")
                bits, pops = payload
                emap = []
                progs, count, slots = blocking.compile_pass_batched(bits, pops, n, self.cdt, 1, emap)
                pos, cols = [], []
                for off, i, nent in emap:
                    if id(pops[i]) in src:
                        pos.append(off // isz + np.arange(nent))
                        cols.append(first_col(pops[i]) + np.arange(nent))
                cat = (lambda a: np.concatenate(a) if a else np.zeros(0, dtype=np.int64))
                out.append(_SweepPass(bits, count, slots, progs[0].copy(), cat(pos), cat(cols)))

    def data(self, params: np.ndarray) -> np.ndarray:
        """Per-set data table [B, ncols] for parameter sets [B, P]."""
        nb = params.shape[0]
        table = np.empty((nb, self.ncols), dtype=self.cdt)
        for name, slots, nent, c0, c1 in self.groups:
            g = slots.size
            d, dd = _rotation_data(name, params[:, slots].reshape(-1), self.cdt)
            table[:, c0:c1] = d.reshape(nb, g * nent)
            table[:, c1:c1 + g * nent] = dd.reshape(nb, g * nent)
        return table

    def programs(self, p: _SweepPass, table: np.ndarray) -> np.ndarray:
        """Programs [B, nbytes] uint8 of pass ``p`` for the sets of ``table``."""
        out = np.empty((table.shape[0], p.row.size), dtype=np.uint8)
        out[:] = p.row
        if p.pos.size:
            out.view(self.cdt)[:, p.pos] = table[:, p.cols]
        return out


def _upload_programs(passes, template: SweepTemplate, table: np.ndarray, device):
    """One host-to-device copy of the programs of ``passes`` for the sets of
    ``table``. Returns the device buffer and, per pass, (byte offset of its
    first set's program, stride between sets, program bytes); every program
    starts on 256 bytes and the buffer ends with the padding the kernels may
    read past a record stream."""
    nb = table.shape[0]
    layout, total = [], 0
    for p in passes:
        nbytes = p.row.size
        stride = (nbytes + 255) // 256 * 256
        layout.append((total, stride, nbytes))
        total += nb * stride
    host = torch.empty(total + blocking.PAD_BYTES, dtype=torch.uint8).pin_memory()
    buf = host.numpy()
    buf[total:] = 0
    for p, (off, stride, nbytes) in zip(passes, layout):
        buf[off: off + nb * stride].reshape(nb, stride)[:, :nbytes] = template.programs(p, table)
    return host.to(device, non_blocking=True), layout


def _structure_key(backend, circuit):
    """Everything the sweep's records depend on: per gate its class and qubits,
    and the parameter slot of a trainable gate or the parameter values of a
    fixed one. None when some gate cannot be described."""
    slot_of = {id(g): k for k, g in enumerate(circuit.trainable_gates)}
    parts = []
    for g in circuit.queue:
        if kinds.is_measurement(g):
            continue
        if id(g) in slot_of and g.trainable:
            parts.append((type(g).__name__, g.target_qubits, g.control_qubits, g.is_controlled_by, slot_of[id(g)]))
            continue
        k = backend._gate_key(g)
        if k is None:
            return None
        parts.append(k)
    return circuit.nqubits, backend.dtype, tuple(parts)


_TEMPLATE_CACHE_SIZE = 8


def sweep_template(backend, circuit) -> SweepTemplate:
    """The circuit's SweepTemplate, from the backend's cache when a circuit of
    the same structure was seen (circuits that are not described exactly are
    compiled every time). Raises _Unsupported."""
    cdt = np.complex128 if backend.dtype == "complex128" else np.complex64
    key = _structure_key(backend, circuit)
    cache = backend.__dict__.setdefault("_sweep_templates", {})
    if key is not None and key in cache:
        tmpl = cache.pop(key)
        cache[key] = tmpl  # most recent last
    else:
        try:
            tmpl = SweepTemplate(backend, circuit, cdt)
        except _Unsupported as e:
            tmpl = e
        if key is not None:
            cache[key] = tmpl
            while len(cache) > _TEMPLATE_CACHE_SIZE:
                cache.pop(next(iter(cache)))
    if isinstance(tmpl, _Unsupported):
        raise tmpl
    return tmpl


def blocked_expectation_and_gradient(backend, circuit, hamiltonian, params: np.ndarray):
    """Energies [B] and gradients [B, P] of all parameter sets at once (see
    batched_expectation_and_gradient); raises _Unsupported for other circuits."""
    n = circuit.nqubits
    if circuit.density_matrix or n < blocking.TILE_BITS:
        raise _Unsupported("size")
    tmpl = sweep_template(backend, circuit)
    if params.shape[1] != tmpl.nparams:
        raise _Unsupported("parameter count")
    nb_all = params.shape[0]
    nparams = tmpl.nparams
    xm, zm, ny, coeffs = hamiltonian.pauli_tables()
    dev = backend.torch_device
    tdt = torch.complex128 if backend.dtype == "complex128" else torch.complex64
    code = _lib.C128 if backend.dtype == "complex128" else _lib.C64
    lib = backend.lib
    state_bytes = (2**n) * (16 if code == _lib.C128 else 8)
    free = torch.cuda.mem_get_info(dev)[0]
    chunk = max(1, min(nb_all, int(0.4 * free) // (2 * state_bytes)))
    energies = torch.empty(nb_all, dtype=torch.float64, device=dev)
    grads = torch.zeros((nb_all, nparams), dtype=torch.float64, device=dev)
    stream = backend._stream()
    c = np.ascontiguousarray(coeffs, dtype=np.float64)
    xm = np.ascontiguousarray(xm, dtype=np.uint64)
    zm = np.ascontiguousarray(zm, dtype=np.uint64)
    ny = np.ascontiguousarray(ny, dtype=np.int32)
    for b0 in range(0, nb_all, chunk):
        b1 = min(nb_all, b0 + chunk)
        nb = b1 - b0
        table = tmpl.data(params[b0:b1])
        keep = []
        psi = torch.zeros((nb, 2**n), dtype=tdt, device=dev)
        psi[:, 0] = 1
        pdev, layout = _upload_programs(tmpl.fwd, tmpl, table, dev)
        keep.append(pdev)
        for p, (off, stride, nbytes) in zip(tmpl.fwd, layout):
            # the kernel specialised to the pass's records once compiled
            # (every set's program has the template row's records)
            spec = backend._specialized(p.row, p.count, code, _lib.SPEC_PASS, dev) if backend.specialize_sweeps else 0
            if spec:
                rc = lib.qsim_apply_blocked_specialized(spec, psi.data_ptr(), nb, n, code, _lib.int_array(p.bits),
                                                        (1 << 64) - 1, pdev.data_ptr() + off, stride, nbytes,
                                                        p.count, stream)
                backend.specialized_launches += 1
            else:
                rc = lib.qsim_apply_blocked_batched(psi.data_ptr(), nb, n, code, _lib.int_array(p.bits),
                                                    pdev.data_ptr() + off, stride, nbytes, p.count, stream)
            _lib.check(rc, lib)
        lam = torch.empty_like(psi)
        for b in range(nb):
            _lib.check(lib.qsim_apply_pauli_sum(psi[b].data_ptr(), n, code, len(c), xm.ctypes.data, zm.ctypes.data,
                                                ny.ctypes.data, c.ctypes.data, lam[b].data_ptr(), stream), lib)
            energies[b0 + b] = backend.expectation_pauli(psi[b], n, xm, zm, ny, c) + hamiltonian.constant
        gsub = grads[b0:b1]
        pdev, layout = _upload_programs(tmpl.rev, tmpl, table, dev)
        keep.append(pdev)
        for p, (off, stride, nbytes) in zip(tmpl.rev, layout):
            npart = max(1, lib.qsim_adjoint_partials_count(n, nb, len(p.slots)))
            part = torch.empty(npart, dtype=torch.float64, device=dev)
            spec = backend._specialized(p.row, p.count, code, _lib.SPEC_ADJOINT, dev) if backend.specialize_sweeps else 0
            launch = lib.qsim_adjoint_blocked
            args = (psi.data_ptr(), lam.data_ptr(), nb, n, code, _lib.int_array(p.bits), pdev.data_ptr() + off,
                    stride, nbytes, p.count, len(p.slots), part.data_ptr(), npart, stream)
            if spec:
                launch, args = lib.qsim_adjoint_blocked_specialized, (spec,) + args
                backend.specialized_launches += 1
            _lib.check(launch(*args), lib)
            keep.append(part)
            if p.slots:
                pof = p.dev_slots.get(dev)
                if pof is None:
                    pof = p.dev_slots[dev] = torch.tensor(p.slots, dtype=torch.int32).to(dev)
                _lib.check(lib.qsim_adjoint_reduce(part.data_ptr(), npart, nb, n, len(p.slots), pof.data_ptr(), nparams,
                                                   gsub.data_ptr(), stream), lib)
        for t in keep:
            t.record_stream(torch.cuda.current_stream(dev))
    return energies, grads


class AdjointExpectation(torch.autograd.Function):
    """E(params) = <psi(params)|H|psi(params)> for a batch of parameter sets.

    ``forward`` runs the circuit and the adjoint sweep for each row of
    ``params`` ([B, P] float64) and keeps the gradient; ``backward`` scales
    it by the incoming gradient. Usage::

        energies = AdjointExpectation.apply(params, backend, circuit, hamiltonian)
        energies.sum().backward()
    """

    @staticmethod
    def forward(ctx, params, backend, circuit, hamiltonian):
        energies, grads = batched_expectation_and_gradient(backend, circuit, hamiltonian, params)
        ctx.save_for_backward(grads.to(params.device))
        return energies.to(params.device, dtype=params.dtype)

    @staticmethod
    def backward(ctx, grad_out):
        (grads,) = ctx.saved_tensors
        return grad_out[:, None].to(grads.dtype) * grads, None, None, None
