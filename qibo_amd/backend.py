"""``MI355XBackend``: qibo's ``Backend`` contract on top of the HIP library.

The state vector is a torch tensor in HBM (complex128 or complex64) that the
backend owns and updates in place; every gate pass, probability reduction,
cumulative sum and sampling search is one or more calls into
``libqsim_mi355x.so`` on the tensor's current HIP stream. torch is used for
allocation, streams and a few O(nshots) host-side conversions only.

Method-by-method counterpart of qibo 0.3.5 (``src/qibo/backends/abstract.py``):

* ``execute_circuit`` / ``_execute_circuit`` -- :2442-2512, :3306-3344
* ``zero_state`` -- :2243-2275
* ``apply_gate`` (dense, controlled, controlled_by) -- :2322-2361, :3176-3197
* ``matrix`` / ``matrix_parametrized`` / ``matrix_fused`` -- :2649-2717
* ``calculate_probabilities`` -- :2734-2758, :3371-3381
* ``sample_shots`` / ``sample_frequencies`` / ``update_frequencies`` --
  :2760-2801 with ``random_choice`` :1265-1309
* ``samples_to_binary`` / ``samples_to_decimal`` / ``calculate_frequencies``
  -- :2723-2793
* ``exp_value_observable_symbolic`` -- :2946-3054

A gate's matrix is inspected once per call to pick the cheapest exact kernel
form: rows that act as the identity on a qubit being 0 make that qubit a
control (CNOT, CZ, CU1, TOFFOLI, ...), diagonal matrices only touch the
amplitudes they change, and the SWAP permutation only moves two amplitudes
per group. Each form computes the same products as the dense matrix, minus
the terms multiplied by an exact zero.
"""

import collections
import ctypes
import math
import os
import threading
from typing import List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _lib
from . import blocking
from . import config
from . import gates as gates_mod
from . import kinds
from .config import SHOT_BATCH_SIZE, log, raise_error
from .matrices import GateMatrices

_DTYPES = {
    "complex128": (torch.complex128, np.complex128, _lib.C128, torch.float64),
    "complex64": (torch.complex64, np.complex64, _lib.C64, torch.float32),
}

_SWAP4 = np.array([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]])


class _Op:
    """A planned kernel call for one gate."""

    __slots__ = ("kind", "targets", "controls", "data", "native", "masks", "dev")

    def __init__(self, kind, targets, controls, data):
        self.native = None  # flat description for the native record compiler (blocking._desc)
        self.masks = None  # bit masks for the pass selection (blocking._masks)
        self.dev = None  # (device, dtype, tensor): a dense block's matrix uploaded once
        self.kind = kind          # "noop" | "diag" | "swap" | "dense" | "fused"
        self.targets = tuple(targets)
        self.controls = tuple(controls)
        self.data = data          # numpy matrix / diagonal, or None


def _embed(u: np.ndarray, qubits: Tuple[int, ...], targets: Tuple[int, ...]) -> np.ndarray:
    """``u`` acting on ``qubits`` (first = most significant) as a matrix on
    ``targets`` (a superset, same convention), identity on the others."""
    if qubits == targets:
        return u
    rest = tuple(q for q in targets if q not in qubits)
    full = np.kron(u, np.eye(2 ** len(rest), dtype=u.dtype))  # on qubits + rest
    order = qubits + rest
    k = len(targets)
    perm = [order.index(q) for q in targets]
    full = full.reshape((2,) * (2 * k)).transpose(perm + [p + k for p in perm])
    return full.reshape(2**k, 2**k)


def _peel_controls(m: np.ndarray, qubits: Sequence[int]) -> Tuple[np.ndarray, list, list]:
    """Move qubits on which ``m`` is ``[[I, 0], [0, U]]`` into the control list.

    ``m`` acts on ``qubits`` with qubits[0] as the most significant bit.
    Returns (reduced matrix on the remaining targets, targets, controls).
    """
    targets = list(qubits)
    controls = []
    changed = True
    while changed and targets:
        changed = False
        k = len(targets)
        dim = 2 ** (k - 1)
        tensor = m.reshape((2,) * (2 * k))
        for t in range(k):
            def block(rb, cb, t=t):
                sel = [slice(None)] * (2 * k)
                sel[t], sel[k + t] = rb, cb
                return tensor[tuple(sel)].reshape(dim, dim)

            if np.any(block(0, 1) != 0) or np.any(block(1, 0) != 0):
                continue
            if not np.array_equal(block(0, 0), np.eye(dim)):
                continue
            m = block(1, 1)
            controls.append(targets.pop(t))
            changed = True
            break
    return m, targets, controls


def plan_matrix(matrix: np.ndarray, targets: Sequence[int], controls: Sequence[int]) -> _Op:
    """Cheapest exact kernel form of ``matrix`` on ``targets`` under ``controls``."""
    matrix = np.asarray(matrix)
    if matrix.shape == (2, 2):  # one target: the same result as the general path, without it
        a, b, c, d = matrix.ravel().tolist()
        t = targets[0]
        if b != 0 or c != 0:
            return _Op("dense", (t,), controls, matrix)
        if a == 1:  # [[1, 0], [0, d]]: the target becomes a control
            if d == 1:
                return _Op("noop", (), tuple(controls) + (t,), None)
            return _Op("diag", (), tuple(controls) + (t,), np.array([d], dtype=matrix.dtype))
        return _Op("diag", (t,), controls, np.array([a, d], dtype=matrix.dtype))
    if matrix.ndim == 2 and not np.any(matrix - np.diag(np.diagonal(matrix))):
        # diagonal: the same peeling on the diagonal vector (qubit t is a
        # control when every entry with bit t = 0 is 1), without the blocks
        d = np.diagonal(matrix).copy()
        tg, extra = list(targets), []
        changed = True
        while changed and tg:
            changed = False
            k = len(tg)
            cube = d.reshape((2,) * k)
            for t in range(k):
                if np.all(np.take(cube, 0, axis=t) == 1):
                    d = np.take(cube, 1, axis=t).ravel().copy()
                    extra.append(tg.pop(t))
                    changed = True
                    break
        ctrl = list(controls) + extra
        if np.all(d == 1):
            return _Op("noop", tg, ctrl, None)
        if len(tg) > 5:
            raise_error(NotImplementedError, "diagonal gates on more than 5 targets")
        return _Op("diag", tg, ctrl, d)
    m, tg, extra = _peel_controls(matrix, targets)
    ctrl = list(controls) + extra
    dim = m.shape[0]
    if dim == 1 or np.count_nonzero(m - np.diag(np.diagonal(m))) == 0:
        d = np.diagonal(m).copy()
        if np.all(d == 1):
            return _Op("noop", tg, ctrl, None)
        if len(tg) > 5:
            raise_error(NotImplementedError, "diagonal gates on more than 5 targets")
        return _Op("diag", tg, ctrl, d)
    if len(tg) == 2 and np.array_equal(m, _SWAP4):
        return _Op("swap", tg, ctrl, None)
    if len(tg) <= 3:
        return _Op("dense", tg, ctrl, m)
    # dense block on 4-5 targets: the block kernel takes no controls, so keep
    # any controls inside the matrix ([[I, 0], [0, U]] on controls + targets)
    if controls:
        full = np.eye(2 ** (len(controls) + len(targets)), dtype=matrix.dtype)
        full[-len(matrix):, -len(matrix):] = matrix
        matrix, targets = full, tuple(controls) + tuple(targets)
    if len(targets) > 5:
        raise_error(NotImplementedError, "dense blocks on more than 5 qubits are not supported")
    return _Op("fused", targets, (), matrix)


# a fused group of at most 3 qubits with at most this many gates runs as its
# gates inside the blocked passes (GatePlanner.plan_ops)
UNFUSE_MAX_GATES = 12

# plans of parameterless gates, keyed by (class, targets, controls,
# controlled_by, dtype); an op is never modified after planning, only
# annotated with caches that depend on nothing else (its record descriptor,
# bit masks per register size)
_SHARED_PLANS = {}


class GatePlanner:
    """Host half of gate application: matrices and the kernel plan of a gate.

    Needs no device; ``MI355XBackend`` inherits it. ``matrices`` holds the
    numpy gate formulas of the backend dtype.
    """

    def __init__(self, dtype: str = "complex128"):
        self.dtype = dtype
        self.matrices = GateMatrices(dtype)

    # -- matrices (host) -----------------------------------------------------
    def matrix(self, gate):
        name = gate.__class__.__name__
        _matrix = getattr(self.matrices, name)
        if name == "I":
            _matrix = _matrix(2 ** len(gate.target_qubits))
        elif name == "Align":
            _matrix = _matrix(0, 2)
        elif callable(_matrix):
            return self.matrix_parametrized(gate)
        return np.asarray(_matrix)

    def matrix_parametrized(self, gate):
        name = gate.__class__.__name__
        _matrix = getattr(self.matrices, name)
        if name == "GeneralizedRBS":
            _matrix = _matrix(qubits_in=gate.init_args[0], qubits_out=gate.init_args[1],
                              theta=gate.init_kwargs["theta"], phi=gate.init_kwargs["phi"])
        elif name == "FanOut":
            _matrix = _matrix(*gate.init_args)
        else:
            _matrix = _matrix(*gate.parameters)
        return np.asarray(_matrix)

    def matrix_fused(self, fgate):
        """Product of the group's matrices on its sorted qubits
        (abstract.py:2680-2717), formed in the library (qsim_matrix_fused, host
        code): each gate matrix, with its ``controlled_by`` controls as a block
        diagonal, is applied to every column of the running product. The gate
        matrices are this planner's formulas; the product is FP64 and cast to
        the backend dtype."""
        from . import _lib

        lib = _lib.load()
        rank = len(fgate.target_qubits)
        cdt = _DTYPES[self.dtype][1]
        ng = len(fgate.gates)
        desc = np.zeros((max(ng, 1), 2 + rank), dtype=np.int32)
        mats, offs, off = [], np.zeros(max(ng, 1), dtype=np.int64), 0
        for k, gate in enumerate(fgate.gates):
            gm = np.ascontiguousarray(gate.matrix(self), dtype=np.complex128)
            qubits = list(gate.qubits)
            if len(gate.control_qubits) > 0 and len(gm) < 2 ** len(qubits):
                ctrl, targ = list(gate.control_qubits), list(gate.target_qubits)
            else:
                ctrl, targ = [], qubits
            if len(ctrl) + len(targ) > rank or len(gm) != 2 ** len(targ):
                raise_error(ValueError, f"gate {gate} does not fit the fused group on {fgate.target_qubits}")
            desc[k, 0], desc[k, 1] = len(ctrl), len(targ)
            desc[k, 2: 2 + len(ctrl) + len(targ)] = ctrl + targ
            mats.append(gm.ravel())
            offs[k] = off
            off += gm.size
        data = np.concatenate(mats) if mats else np.zeros(1, dtype=np.complex128)
        out = np.empty((2**rank, 2**rank), dtype=np.complex128)
        _lib.check(lib.qsim_matrix_fused(rank, _lib.int_array(fgate.target_qubits), ng, desc.ctypes.data,
                                         data.ctypes.data, offs.ctypes.data, out.ctypes.data), lib)
        return out.astype(cdt)

    def _matrix_fused_python(self, fgate):
        """The reference's construction of matrix_fused in numpy: each gate
        matrix (with its ``controlled_by`` controls as a block diagonal) is
        extended by a Kronecker identity, its axes are permuted onto the
        group's qubits, and the result left-multiplies the running product."""
        rank = len(fgate.target_qubits)
        cdt = _DTYPES[self.dtype][1]
        matrix = np.eye(2**rank, dtype=cdt)
        for gate in fgate.gates:
            gmatrix = np.asarray(gate.matrix(self))
            if len(gate.control_qubits) > 0 and len(gmatrix) < 2 ** len(gate.qubits):
                big = np.eye(2 ** len(gate.qubits), dtype=gmatrix.dtype)
                big[-len(gmatrix):, -len(gmatrix):] = gmatrix
                gmatrix = big
            eye = np.eye(2 ** (rank - len(gate.qubits)), dtype=gmatrix.dtype)
            gmatrix = np.kron(gmatrix, eye)
            shape = gmatrix.shape
            gmatrix = gmatrix.reshape(2 * rank * (2,))
            qubits = list(gate.qubits)
            order = qubits + [q for q in fgate.target_qubits if q not in qubits]
            perm = [int(i) for i in np.argsort(np.asarray(order, dtype=np.int64))]
            gmatrix = np.transpose(gmatrix, perm + [p + rank for p in perm]).reshape(shape)
            matrix = gmatrix @ matrix
        return matrix.astype(cdt)

    # -- gate application ----------------------------------------------------
    def plan_gate(self, gate) -> _Op:
        if kinds.is_measurement(gate):
            raise_error(ValueError, "measurement gates have no matrix and are not planned as gates")
        if kinds.is_fused(gate):
            return plan_matrix(self.matrix_fused(gate), gate.target_qubits, ())
        # gates that keep the base matrix() go straight to the backend formulas;
        # their plan is kept on the gate while its parameter and qubit tuples are
        # the same objects (every setter assigns new ones) and no parameter is an
        # array (which could change in place)
        base = type(gate).matrix in _BASE_MATRIX
        fp = None
        if base:
            fp = (gate._parameters, gate._target_qubits, gate._control_qubits, gate.is_controlled_by, self.dtype)
            cached = gate.__dict__.get("_qplan")
            if (cached is not None and cached[0] is fp[0] and cached[1] is fp[1] and cached[2] is fp[2]
                    and cached[3] == fp[3] and cached[4] == fp[4]):
                return cached[5]
        # a gate without parameters has one plan per class, qubits and dtype,
        # shared by every circuit (the analysis of a 4x4 matrix costs ~30 us)
        skey = None
        if base and not gate._parameters:
            skey = (type(gate).__name__, gate._target_qubits, gate._control_qubits, gate.is_controlled_by,
                    self.dtype)
            op = _SHARED_PLANS.get(skey)
            if op is not None:
                gate.__dict__["_qplan"] = fp + (op,)
                return op
        m = self.matrix(gate) if base else gate.matrix(self)
        if gate.is_controlled_by:
            op = plan_matrix(m, gate.target_qubits, gate.control_qubits)
        else:
            op = plan_matrix(m, gate.qubits, ())
        if base and not any(isinstance(x, (np.ndarray, torch.Tensor)) for x in gate._parameters):
            gate.__dict__["_qplan"] = fp + (op,)
        if skey is not None:
            if len(_SHARED_PLANS) >= 65536:
                _SHARED_PLANS.clear()
            _SHARED_PLANS[skey] = op
        return op

    def plan_ops(self, gate) -> List[_Op]:
        """The ops that apply ``gate`` inside blocked passes. A fused group of
        at most 3 qubits and at most ``UNFUSE_MAX_GATES`` gates whose gates are
        all blockable runs as its gates (rotations, phase tables, sign flips
        in registers) instead of its dense 4x4 / 8x8 product, which a pass
        can only apply through an LDS gather (about 14 rotations' worth): the
        same unitary, rounded differently. Larger groups keep their dense
        matrix (4-5 qubits: one MFMA pass)."""
        if kinds.is_fused(gate) and len(gate.target_qubits) <= 3 and len(gate.gates) <= UNFUSE_MAX_GATES:
            parts = [self.plan_gate(g) for g in gate.gates]
            if all(blocking.blockable(op) or op.kind == "noop" for op in parts):
                return parts
        return [self.plan_gate(gate)]

    def channel_transfer_matrix(self, channel) -> Optional[np.ndarray]:
        """The map of a channel on one or two qubits T as a 4^k x 4^k matrix
        on (row qubits T, column qubits T) of the vectorized density matrix,
        or None when the channel is not of that form.

        A term p U rho U^dagger maps rho[r', c'] to rho[r, c] with weight
        p U[r, r'] conj(U[c, c']), i.e. kron(U, conj(U)); the channel is
        (1 - sum p) I + sum_k p_k kron(U_k, conj(U_k)) (apply_channel,
        abstract.py:2297-2311), each U_k embedded on T. That covers Kraus,
        unitary, Pauli-noise, depolarizing (its Pauli decomposition), reset
        (its Kraus operators), damping and readout channels.
        ThermalRelaxationChannel uses the map its apply writes
        (channels.py:678-699), or the Kraus form on qubit 0 where the
        reference's reset branch puts its Z."""
        targets = tuple(channel.target_qubits)
        k = len(targets)
        if k not in (1, 2) or len(set(targets)) != k:
            return None
        if type(channel).__name__ == "ThermalRelaxationChannel":
            if channel.t_1 < channel.t_2:
                p0, p1, e = (channel.init_kwargs[key] for key in ("p_0", "p_1", "e_t2"))
                return np.array([[1 - p1, 0, 0, p0], [0, e, 0, 0], [0, 0, e, 0], [p1, 0, 0, 1 - p0]],
                                dtype=np.complex128)
            if targets[0] != 0:
                return None
        dim = 2**k
        sop = (1 - channel.coefficient_sum) * np.eye(dim * dim, dtype=np.complex128)
        for coeff, gate in zip(channel.coefficients, channel.gates):
            if gate.is_controlled_by or gate.control_qubits or not set(gate.target_qubits) <= set(targets):
                return None
            u = self.matrix_fused(gate) if kinds.is_fused(gate) else gate.matrix(self)
            u = _embed(np.asarray(u, dtype=np.complex128), tuple(gate.target_qubits), targets)
            sop = sop + coeff * np.kron(u, np.conj(u))
        return sop

    # -- Pauli strings (host) ---------------------------------------------
    def pauli_tables(self, terms: Sequence[str], term_qubits, nqubits: int):
        """X/Z bit masks and Y counts of Pauli strings over amplitude bits."""
        xm, zm, ny = [], [], []
        for term, qubits in zip(terms, term_qubits):
            x = z = y = 0
            for factor, q in zip(term, qubits):
                bit = 1 << (nqubits - 1 - int(q))
                if factor == "X":
                    x ^= bit
                elif factor == "Z":
                    z ^= bit
                elif factor == "Y":
                    x ^= bit
                    z ^= bit
                    y += 1
                elif factor != "I":
                    raise_error(ValueError, f"Unknown Pauli factor {factor}.")
            xm.append(x)
            zm.append(z)
            ny.append(y)
        return (np.asarray(xm, dtype=np.uint64), np.asarray(zm, dtype=np.uint64),
                np.asarray(ny, dtype=np.int32))


class MI355XBackend(GatePlanner):
    """State-vector simulation backend running on one MI355X (gfx950)."""

    def __init__(self, dtype: str = "complex128", device: int = None):
        self.lib = _lib.load()
        if not torch.cuda.is_available():
            raise_error(RuntimeError, "qibo_amd: no HIP device is visible to torch.")
        if device is None:
            device = int(os.environ.get("QIBO_AMD_DEVICE", "0"))
        self.torch_device = torch.device("cuda", device)
        self.device = f"/GPU:{device}"
        self.name = "qibo_amd"
        self.platform = "mi355x"
        self.engine = np
        self.nthreads = 1
        self.numeric_types = (int, float, complex, np.int32, np.int64, np.float32, np.float64,
                              np.complex64, np.complex128)
        self.tensor_types = (torch.Tensor, np.ndarray)
        self.oom_error = MemoryError
        self.supports_multigpu = False
        self.versions = {"qibo_amd": "0.1.0", "torch": torch.__version__}
        self.dtype = None
        self.matrices = None
        self.set_dtype(dtype)
        self._ws = {}  # workspace cache: (device, purpose, stream) -> tensor
        self._timer = None  # (selector, [(start, end) events]) while kernel timing is on
        self.blocking = True  # group gates into blocked passes in run_queue
        # keep circuit order inside blocked passes (bitwise equal to the per-gate
        # path); the default lets commuting gates move ahead, giving fewer passes
        self.blocking_ordered = False
        # index bits every pass keeps in its tile (consecutive amplitudes per
        # memory segment); None: 3 for complex128, 4 for complex64 (128 bytes)
        self.blocking_low_bits = None
        self.block_stats = collections.deque(maxlen=1 << 16)  # gates per blocked pass (diagnostics)
        # (algorithmic HBM bytes, over every tile) per blocked launch (diagnostics)
        self.block_bytes = collections.deque(maxlen=1 << 16)
        # compiled blocked passes of recently executed gate queues, keyed by the
        # gates' kinds, qubits and exact parameter values (see run_queue)
        # one-qubit channels inside density-matrix circuits run as their 4x4
        # transfer matrix on (row qubit, column qubit), inside the blocked
        # passes; False applies them term by term in the reference's order
        self.dm_superoperators = True
        self.program_cache = True
        self.program_cache_size = 16
        # blocked passes specialised to their record sequence (jit.hip): a
        # sequence launched `specialize_after` times is compiled in the
        # background (hiprtc, 1-3 s per pass on the host) and later launches
        # of any program with that sequence (other angles included) run the
        # straight-line kernel, bitwise equal to the interpreted pass
        self.specialize = os.environ.get("QIBO_AMD_SPECIALIZE", "1") != "0"
        self.specialize_after = 2
        # the batched sweeps of adjoint gradients too: off by default, their
        # long runs of 2x2 and diagonal records measured no faster specialised
        # (HEA-20, 64 sets: 852 against 937 gradients/s in complex128)
        self.specialize_sweeps = False
        self._spec_seen = {}
        self._spec_jobs = {}
        self._spec_pool = None
        self._spec_lock = threading.Lock()
        self.specialized_launches = 0
        self.sparse_launches = 0  # blocked passes run on part of the tiles (support below every bit)
        # passes after a start from |0...0> run only the tiles whose amplitudes
        # can be nonzero (the union of the earlier passes' tile bits)
        self.skip_zero_tiles = True
        self._programs = collections.OrderedDict()
        self._programs_lock = threading.Lock()
        self.program_cache_hits = 0
        # attach an accuracy report (infidelity and max-abs amplitude error
        # against the complex128 execution of the same circuit) to every
        # state-vector result of a complex64 backend; off by default (the
        # report runs the circuit a second time)
        self.report_accuracy = False
        rec = self.lib.qsim_blocked_header_bytes()
        if rec != blocking.HEADER_BYTES:
            raise ImportError(f"qibo_amd: blocked-pass record header of {rec} bytes does not match the host layout")

    # -- qibo Backend attributes -----------------------------------------
    @property
    def qubits(self):
        return None

    @property
    def connectivity(self):
        return None

    @property
    def natives(self):
        return None

    def __repr__(self):
        return f"{self.name} ({self.platform})"

    # -- configuration -----------------------------------------------------
    def set_dtype(self, dtype: str) -> None:
        if dtype not in _DTYPES:
            raise_error(
                ValueError,
                f"Unknown ``dtype`` ``{dtype}``. For this backend ({self}), ``dtype`` must be "
                f"either one of the following string: {tuple(_DTYPES)}",
            )
        if dtype != self.dtype:
            self.dtype = dtype
            self.matrices = GateMatrices(dtype)

    def set_device(self, device: str) -> None:
        if not device.startswith("/GPU:"):
            raise_error(ValueError, f"Device {device} is not available for {self} backend.")
        idx = int(device.split(":")[1])
        self.torch_device = torch.device("cuda", idx)
        self.device = device

    def set_threads(self, nthreads: int) -> None:
        if nthreads > 1:
            log.warning("qibo_amd runs on the GPU; set_threads is ignored.")

    def set_seed(self, seed) -> None:
        # Sampling draws its uniforms from numpy's global RandomState, exactly
        # as the reference does through np.random.choice (abstract.py:181-187).
        np.random.seed(seed)

    # -- helpers -----------------------------------------------------------
    @property
    def _tdtype(self):
        return _DTYPES[self.dtype][0]

    def _stream(self, t: torch.Tensor = None):
        dev = t.device if t is not None else self.torch_device
        return ctypes_stream(torch.cuda.current_stream(dev))

    def _workspace(self, nbytes: int, purpose: str, device=None) -> torch.Tensor:
        device = device or self.torch_device
        key = (device, purpose, torch.cuda.current_stream(device).cuda_stream)
        ws = self._ws.get(key)
        if ws is None or ws.numel() < nbytes:
            ws = torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device=device)
            self._ws[key] = ws
        return ws

    def _check_budget(self, nbytes: int):
        budget = config.HBM_BUDGET_BYTES
        if budget is not None and nbytes > budget:
            raise MemoryError(f"state of {nbytes} bytes exceeds the HBM budget of {budget} bytes")

    def _alloc_state(self, nqubits: int, dtype=None) -> torch.Tensor:
        tdt = _DTYPES[dtype or self.dtype][0]
        nbytes = (2**nqubits) * (16 if tdt == torch.complex128 else 8)
        self._check_budget(nbytes)
        try:
            return torch.empty(2**nqubits, dtype=tdt, device=self.torch_device)
        except torch.cuda.OutOfMemoryError as exc:
            raise MemoryError(str(exc)) from exc

    @staticmethod
    def _dtype_code(state: torch.Tensor) -> int:
        if state.dtype == torch.complex128:
            return _lib.C128
        if state.dtype == torch.complex64:
            return _lib.C64
        raise_error(TypeError, f"State dtype {state.dtype} is not complex64/complex128.")

    def cast(self, array, dtype=None, copy: bool = False):
        if dtype is None:
            dtype = self.dtype
        tdt = _torch_dtype(dtype)
        if isinstance(array, torch.Tensor):
            out = array.to(device=self.torch_device, dtype=tdt)
            if copy and out.data_ptr() == array.data_ptr():
                out = out.clone()
            return out
        arr = np.asarray(array)
        return torch.as_tensor(np.ascontiguousarray(arr), device=self.torch_device).to(tdt)

    def to_numpy(self, array):
        if isinstance(array, torch.Tensor):
            return array.detach().cpu().numpy()
        return np.asarray(array)

    def zeros(self, shape, dtype=None):
        return torch.zeros(shape, dtype=_torch_dtype(dtype or self.dtype), device=self.torch_device)

    def concatenate(self, arrays, axis=0):
        return torch.cat([self.cast(a, dtype=a.dtype) if isinstance(a, torch.Tensor) else
                          torch.as_tensor(a, device=self.torch_device) for a in arrays], dim=axis)

    def assert_allclose(self, value, target, rtol=1e-7, atol=0.0):
        np.testing.assert_allclose(self.to_numpy(value), self.to_numpy(target), rtol=rtol, atol=atol)

    # -- state -------------------------------------------------------------
    def _validate_nqubits(self, nqubits, density_matrix=False):
        if not isinstance(nqubits, int) or nqubits < 1:
            raise_error(ValueError, f"nqubits must be a positive integer, but got {nqubits}.")
        if density_matrix:
            if nqubits > _MAX_DM_QUBITS:
                raise_error(ValueError, f"qibo_amd supports density matrices of at most {_MAX_DM_QUBITS} qubits.")
            if config.MAX_QUBITS_DM != -1 and nqubits > config.MAX_QUBITS_DM:
                mem_gib = 4**nqubits * 16 / 1024**3
                raise_error(
                    ValueError,
                    f"Requested {nqubits} qubits for density matrix allocation, which would require "
                    f"~{mem_gib:.1f} GiB of memory. The maximum allowed is {config.MAX_QUBITS_DM} qubits.",
                )
            return
        if nqubits > _MAX_QUBITS:
            raise_error(ValueError, f"qibo_amd supports at most {_MAX_QUBITS} qubits.")
        if config.MAX_QUBITS != -1 and nqubits > config.MAX_QUBITS:
            mem_gib = 2**nqubits * 16 / 1024**3
            raise_error(
                ValueError,
                f"Requested {nqubits} qubits for state vector allocation, which would require "
                f"~{mem_gib:.1f} GiB of memory. The maximum allowed is {config.MAX_QUBITS} qubits.",
            )

    def zero_state(self, nqubits: int, density_matrix: bool = False, dtype=None):
        if density_matrix:
            # |0><0|: the zero state of the 2n-qubit vectorized matrix
            self._validate_nqubits(nqubits, density_matrix=True)
            state = self._alloc_state(2 * nqubits, dtype)
            _lib.check(self.lib.qsim_zero_state(state.data_ptr(), 2 * nqubits, self._dtype_code(state),
                                                self._stream(state)), self.lib)
            return state.view(2**nqubits, 2**nqubits)
        self._validate_nqubits(nqubits)
        state = self._alloc_state(nqubits, dtype)
        _lib.check(self.lib.qsim_zero_state(state.data_ptr(), nqubits, self._dtype_code(state),
                                            self._stream(state)), self.lib)
        return state

    # -- per-launch timing (HIP events on the launch stream) -------------------
    _SELECTORS = {
        "dense1": lambda op: op.kind == "dense" and len(op.targets) == 1 and not op.controls,
        "diag1": lambda op: op.kind == "diag" and len(op.targets) == 1 and not op.controls,
        "fused": lambda op: op.kind == "fused",
        "block": lambda op: op.kind == "block",
        "any": lambda op: op.kind != "noop",
    }

    def enable_kernel_timer(self, selector: str = "dense1"):
        """Record HIP events around every launch whose plan matches ``selector``."""
        self._timer = (self._SELECTORS[selector], [])
        return self._timer

    def disable_kernel_timer(self) -> List[float]:
        """Stop timing; returns the duration (ms) of every timed launch."""
        if self._timer is None:
            return []
        _, pairs = self._timer
        self._timer = None
        for _, end in pairs:
            end.synchronize()
        return [a.elapsed_time(b) for a, b in pairs]

    def apply_op(self, op: _Op, state: torch.Tensor, nqubits: int) -> torch.Tensor:
        if op.kind == "noop":
            return state
        timer = self._timer
        if timer is not None and timer[0](op):
            stream = torch.cuda.current_stream(state.device)
            start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            start.record(stream)
            self._launch(op, state, nqubits)
            end.record(stream)
            timer[1].append((start, end))
            return state
        return self._launch(op, state, nqubits)

    def _launch(self, op: _Op, state: torch.Tensor, nqubits: int) -> torch.Tensor:
        code = self._dtype_code(state)
        npdt = np.complex128 if code == _lib.C128 else np.complex64
        stream = self._stream(state)
        tg = _lib.int_array(op.targets)
        ct = _lib.int_array(op.controls)
        if op.kind == "diag":
            d = np.ascontiguousarray(op.data, dtype=npdt)
            rc = self.lib.qsim_apply_diag(state.data_ptr(), nqubits, code, tg, len(op.targets), ct,
                                          len(op.controls), d.ctypes.data, stream)
        elif op.kind == "swap":
            rc = self.lib.qsim_apply_swap(state.data_ptr(), nqubits, code, op.targets[0], op.targets[1],
                                          ct, len(op.controls), stream)
        elif op.kind == "dense":
            m = np.ascontiguousarray(op.data, dtype=npdt)
            rc = self.lib.qsim_apply_dense(state.data_ptr(), nqubits, code, tg, len(op.targets), ct,
                                           len(op.controls), m.ctypes.data, stream)
        else:  # fused: matrix in device memory
            if op.controls:
                raise_error(NotImplementedError, "controlled blocks on more than 3 targets")
            if len(op.targets) > 5:
                raise_error(NotImplementedError, "dense blocks on more than 5 qubits are not supported")
            # the block's matrix goes to the device once per op (a pinned,
            # stream-ordered copy) and is reused when the op runs again
            cached = op.dev
            cur = torch.cuda.current_stream(state.device)
            if cached is not None and cached[0] == state.device and cached[1] == npdt:
                m = cached[2]
                if cached[4] != cur.cuda_stream:
                    cur.wait_event(cached[3])  # the upload ran on another stream
            else:
                host = torch.from_numpy(np.ascontiguousarray(op.data, dtype=npdt)).pin_memory()
                m = host.to(state.device, non_blocking=True)
                ev = torch.cuda.Event()
                ev.record(cur)
                op.dev = (state.device, npdt, m, ev, cur.cuda_stream)
            rc = self.lib.qsim_apply_fused(state.data_ptr(), nqubits, code, tg, len(op.targets),
                                           m.data_ptr(), stream)
            _keep_alive(m, state)
        _lib.check(rc, self.lib)
        return state

    def apply_gate(self, gate, state, nqubits):
        if state.dim() == 2:
            self._dm_check(state, nqubits, "apply_gate")
            flat = state.view(-1)
            for op in _dm_pair(self.plan_gate(gate), nqubits):
                self.apply_op(op, flat, 2 * nqubits)
            return state
        if state.dim() != 1:
            raise_error(ValueError, f"State of shape {tuple(state.shape)} is neither a vector nor a matrix.")
        return self.apply_op(self.plan_gate(gate), state, nqubits)

    # -- density matrices --------------------------------------------------------
    # A density matrix is a (2^n, 2^n) tensor whose row-major storage is the
    # state vector of 2n qubits (row qubit q = qubit q, column qubit q = qubit
    # n + q): a gate is U on the row qubits and conj(U) on the column qubits,
    # as the reference's pair of einsums (abstract.py:2341-2348, 3199-3235).
    @staticmethod
    def _dm_check(state, nqubits: int, what: str):
        """The kernels read and write 4^n elements: refuse anything that is not a
        contiguous (2^n, 2^n) device matrix before a launch."""
        dim = 2**nqubits
        if not isinstance(state, torch.Tensor) or state.dim() != 2 or tuple(state.shape) != (dim, dim):
            raise_error(ValueError, f"{what} needs a ({dim}, {dim}) density matrix, got shape "
                                    f"{tuple(getattr(state, 'shape', ()))}.")
        if not state.is_contiguous():
            raise_error(ValueError, f"{what} needs a contiguous density matrix.")

    def apply_gate_half_density_matrix(self, gate, state, nqubits):
        """U rho: the gate on the row qubits only (abstract.py:2363-2383)."""
        if gate.is_controlled_by:
            raise_error(NotImplementedError, "Gate density matrix half call is not implemented for "
                                             "``controlled_by``gates.")
        state = self.cast(state, dtype=state.dtype, copy=True).contiguous()
        self._dm_check(state, nqubits, "apply_gate_half_density_matrix")
        self.apply_op(self.plan_gate(gate), state.view(-1), 2 * nqubits)
        return state

    def dm_axpy(self, y: torch.Tensor, x: torch.Tensor, a: float, accumulate: bool = True):
        """y = y + a x (or y = a x), each product rounded before the sum."""
        if y.shape != x.shape or y.dtype != x.dtype:
            raise_error(ValueError, "dm_axpy: operands differ in shape or dtype.")
        if y.data_ptr() == x.data_ptr():
            raise_error(ValueError, "dm_axpy: y and x must be different tensors.")
        _lib.check(self.lib.qsim_dm_axpy(y.data_ptr(), x.data_ptr(), x.numel(), self._dtype_code(x), float(a),
                                         1 if accumulate else 0, self._stream(x)), self.lib)
        return y

    def apply_channel(self, channel, state, nqubits):
        """Density matrices: (1 - sum p) rho + sum_k p_k G_k rho G_k^dagger, in
        the reference's order of terms. State vectors: one of the channel's
        gates (or none) drawn with its probability from numpy's global
        generator (abstract.py:2297-2320)."""
        if state.dim() == 2:
            self._dm_check(state, nqubits, "apply_channel")
            new = torch.empty_like(state)
            self.dm_axpy(new, state, 1 - channel.coefficient_sum, accumulate=False)
            tmp = torch.empty_like(state)
            for coeff, gate in zip(channel.coefficients, channel.gates):
                tmp.copy_(state)
                self.apply_gate(gate, tmp, nqubits)
                self.dm_axpy(new, tmp, coeff)
            return new
        probabilities = tuple(channel.coefficients) + (1 - sum(channel.coefficients),)
        index = int(self.sample_shots(np.asarray(probabilities, dtype=np.float64), 1)[0])
        if index != len(channel.gates):
            state = self.apply_gate(channel.gates[index], state, nqubits)
        return state

    def depolarizing_error_density_matrix(self, gate, state, nqubits):
        """(1 - lam) rho + lam Tr_T(rho) (x) I / 2^k, in place (abstract.py:2095-2131)."""
        self._dm_check(state, nqubits, "depolarizing_error_density_matrix")
        targets = [int(q) for q in gate.target_qubits]
        _lib.check(self.lib.qsim_dm_depolarize(state.data_ptr(), nqubits, self._dtype_code(state),
                                               _lib.int_array(targets), len(targets),
                                               float(gate.init_kwargs["lam"]), self._stream(state)), self.lib)
        return state

    def reset_error_density_matrix(self, gate, state, nqubits):
        """(1 - p0 - p1) rho + Tr_q(rho) (x) (p0 |0><0| + p1 |1><1|), in place
        (abstract.py:2212-2232)."""
        self._dm_check(state, nqubits, "reset_error_density_matrix")
        qubit = int(gate.target_qubits[0])
        _lib.check(self.lib.qsim_dm_reset(state.data_ptr(), nqubits, self._dtype_code(state), qubit,
                                          float(gate.init_kwargs["p_0"]), float(gate.init_kwargs["p_1"]),
                                          self._stream(state)), self.lib)
        return state

    def thermal_error_density_matrix(self, gate, state, nqubits):
        """``gate`` (a 4x4 map on row qubit q and column qubit q) applied to the
        vectorized matrix as a 2n-qubit state (abstract.py:2234-2241)."""
        self._dm_check(state, nqubits, "thermal_error_density_matrix")
        self.apply_op(self.plan_gate(gate), state.view(-1), 2 * nqubits)
        return state

    # -- circuit execution -----------------------------------------------------
    def execute_circuit(self, circuit, initial_state=None, nshots: int = 1000):
        from .circuit import Circuit

        nqubits = circuit.nqubits
        dm = bool(circuit.density_matrix)
        self._validate_nqubits(nqubits, density_matrix=dm)
        if isinstance(initial_state, Circuit):
            return self.execute_circuit(initial_state + circuit, None, nshots)
        source = (circuit, initial_state)  # for accuracy reports (re-execution at complex128)
        if initial_state is not None:
            initial_state = self.cast(initial_state, dtype=self.dtype, copy=True)
            valid = (2**nqubits, 2**nqubits) if dm else (2**nqubits,)
            if tuple(initial_state.shape) != valid:
                raise_error(
                    ValueError,
                    f"Given initial state has shape {tuple(initial_state.shape)}instead of the "
                    f"expected {valid}.",
                )
        if circuit.repeated_execution:
            if not circuit.measurements and not circuit.has_collapse:
                raise_error(
                    RuntimeError,
                    "Attempting to perform noisy simulation with `density_matrix=False` "
                    "and no Measurement gate in the Circuit. If you wish to retrieve the "
                    "statistics of the outcomes please include measurements in the circuit, "
                    "otherwise set `density_matrix=True` to recover the final state.",
                )
            return self.execute_circuit_repeated(circuit, nshots, initial_state)
        try:
            result = self._execute_circuit(circuit, initial_state=initial_state, nshots=nshots)
        except self.oom_error:
            raise_error(
                RuntimeError,
                f"State does not fit in {self.device} memory.Please switch the execution device "
                "to a different one using ``qibo.set_device``.",
            )
        result._source = source
        if self.report_accuracy and not circuit.density_matrix:
            result.accuracy_report = result.accuracy()
        return result

    def reference_backend(self):
        """The complex128 backend on the same device that accuracy reports of
        this backend's results compare with (``self`` when it is complex128)."""
        if self.dtype == "complex128":
            return self
        ref = self.__dict__.get("_reference")
        if ref is None or ref.torch_device != self.torch_device:
            ref = MI355XBackend(dtype="complex128", device=self.torch_device.index)
            ref.blocking = self.blocking
            ref.blocking_ordered = self.blocking_ordered
            self._reference = ref
        return ref

    def _execute_circuit(self, circuit, initial_state=None, nshots: int = 1000):
        from .result import CircuitResult, QuantumState

        nqubits = circuit.nqubits
        if circuit.density_matrix:
            state = self.zero_state(nqubits, density_matrix=True) if initial_state is None else initial_state
            state = self.run_queue_density_matrix(circuit.queue, state, nqubits)
        elif initial_state is None and self.blocking and nqubits >= blocking.TILE_BITS:
            # |0...0> is produced by the first blocked pass itself (no zero fill)
            self._validate_nqubits(nqubits)
            state = self._alloc_state(nqubits)  # budget check and MemoryError as zero_state
            state = self.run_queue(circuit.queue, state, nqubits, zero_pending=True)
        else:
            state = self.zero_state(nqubits) if initial_state is None else initial_state
            state = self.run_queue(circuit.queue, state, nqubits)
        if circuit.measurements:
            circuit._final_state = CircuitResult(state, circuit.measurements, backend=self, nshots=nshots)
        else:
            circuit._final_state = QuantumState(state, backend=self)
        return circuit._final_state

    def execute_circuit_repeated(self, circuit, nshots: int, initial_state=None):
        """Run the circuit once per shot (abstract.py:2532-2636, state vectors).

        Each run starts from a copy of the initial state; a collapsing ``M``
        draws its outcome from the current marginal probabilities (one uniform
        from numpy's global generator, as ``MeasurementResult.add_shot``) and
        collapses the state; the final measurements are sampled once per run.
        The uniforms are drawn in the reference's order, so for the same seed
        the samples are the reference's."""
        from .result import MeasurementOutcomes, QuantumState

        dm = bool(circuit.density_matrix)
        if circuit.has_collapse and not circuit.measurements and not dm:
            raise_error(
                RuntimeError,
                "The circuit contains only collapsing measurements (`collapse=True`) but "
                "`density_matrix=False`. Please set `density_matrix=True` to retrieve "
                "the final state after execution.",
            )
        from .result import CircuitResult

        nqubits = circuit.nqubits
        state_copy = (self.zero_state(nqubits, density_matrix=dm) if initial_state is None
                      else self.cast(initial_state, copy=True))
        results, samples = [], []
        total = None  # density matrices: running sum of the final states, in shot order
        for _ in range(nshots):
            state = state_copy.clone()
            if dm:
                state = self.run_queue_density_matrix(circuit.queue, state, nqubits)
                if total is None:
                    total = self.dm_axpy(torch.empty_like(state), state, 1.0, accumulate=False)
                else:
                    self.dm_axpy(total, state, 1.0)
            else:
                state = self.run_queue(circuit.queue, state, nqubits)
            if circuit.measurements:
                result = CircuitResult(state, circuit.measurements, backend=self, nshots=1)
                sample = result.samples()[0]
                results.append(sample)
                if not dm:
                    samples.append("".join(str(int(x)) for x in sample.tolist()))
                for gate in circuit.measurements:
                    gate.result.reset()
        if dm:
            # the mean of the final states (abstract.py:2603-2623): sum / nshots
            final_state = torch.empty_like(total)
            _lib.check(self.lib.qsim_dm_axpy(final_state.data_ptr(), total.data_ptr(), total.numel(),
                                             self._dtype_code(total), float(nshots), 2, self._stream(total)),
                       self.lib)
            if circuit.measurements:
                final = CircuitResult(final_state, circuit.measurements, backend=self,
                                      samples=self.aggregate_shots(results), nshots=nshots)
            else:
                final = QuantumState(final_state, backend=self)
            circuit._final_state = final
            return final
        final_result = MeasurementOutcomes(
            circuit.measurements, backend=self, samples=self.aggregate_shots(results), nshots=nshots
        )
        final_result._repeated_execution_frequencies = self.calculate_frequencies(samples)
        circuit._final_state = final_result
        return final_result

    def collapse_state(self, state, qubits, shot, nqubits, normalize: bool = True, density_matrix=False):
        """Project ``state`` onto the outcome ``shot`` of measuring ``qubits``
        and renormalise it (abstract.py:2424-2440, 3279-3304): the other
        amplitudes become zero, the kept ones are divided by the square root of
        the outcome's probability."""
        qubits = [int(q) for q in qubits]
        outcome = int(shot[0]) if hasattr(shot, "__len__") else int(shot)
        if density_matrix:
            # keep the (outcome, outcome) block of rows and columns and divide it by
            # its trace (_collapse_density_matrix, abstract.py:3249-3277)
            self._dm_check(state, nqubits, "collapse_state")
            m = len(qubits)
            norm = float(self.calculate_probabilities(state, qubits, nqubits, True)[outcome]) if normalize else 1.0
            both = qubits + [q + nqubits for q in qubits]
            _lib.check(self.lib.qsim_collapse(state.data_ptr(), 2 * nqubits, self._dtype_code(state),
                                              _lib.int_array(both), 2 * m, (outcome << m) | outcome, norm,
                                              self._stream(state)), self.lib)
            return state
        if normalize:
            p = self.calculate_probabilities(state, qubits, nqubits)
            norm = math.sqrt(float(p[outcome]))
        else:
            norm = 1.0
        _lib.check(self.lib.qsim_collapse(state.data_ptr(), nqubits, self._dtype_code(state), _lib.int_array(qubits),
                                          len(qubits), outcome, norm, self._stream(state)), self.lib)
        return state

    def run_queue(self, queue, state: torch.Tensor, nqubits: int, zero_pending: bool = False) -> torch.Tensor:
        """Apply the gates of ``queue`` in order (the loop of abstract.py:3321-3322).

        With ``self.blocking`` (the default) consecutive gates are grouped into
        blocked passes (qibo_amd.blocking) that read and write the state once
        each. Commuting gates may move ahead of each other, which changes the
        rounding of the result by a few units in the last place; with
        ``self.blocking_ordered`` circuit order is kept and the result is bitwise
        that of applying the gates one by one.
        Measurement gates only register the backend (non-collapsing ``M`` is a
        no-op on the state); other special gates run in order between passes.
        With ``zero_pending`` the contents of ``state`` are not yet written and
        stand for |0...0>: a first blocked pass produces it itself, anything
        else fills it first.
        """
        # [|0...0> still to be materialised, support, unwritten]: the support
        # is the mask of index bits that may be 1 in a nonzero amplitude (None:
        # any); the passes from |0...0> grow it by their tile bits, skip the
        # tiles outside it (all zero) and read only the amplitudes inside it.
        # unwritten: the amplitudes outside the support are not in memory yet
        # (written as zeros before anything else reads the state)
        zero = [zero_pending, 0 if zero_pending else None, False]
        if zero_pending and (not self.blocking or nqubits < blocking.TILE_BITS):
            self._fill_zero(state, nqubits)
            zero[0] = False
        if not self.blocking:
            for gate in queue:
                state = gate.apply(self, state, nqubits)
            return state
        if nqubits < blocking.TILE_BITS:
            return self._run_small(queue, state, nqubits)
        key = self._queue_key(queue, nqubits, state, zero_pending) if self.program_cache else None
        if key is not None:
            actions = self._cache_get(key)
            if actions is not None:
                return self._replay(actions, queue, state, nqubits, zero)
            record = []
        else:
            record = None
        gates_left = iter(queue)
        stop = []  # the special gate that ended the current segment

        def segment():
            # planned gates up to the next special gate or collapsing
            # measurement, produced lazily so planning overlaps the passes
            # already launched
            for gate in gates_left:
                if kinds.is_measurement(gate) and not gate.collapse:
                    gate.result.backend = self
                    continue
                if kinds.runs_outside_passes(gate):
                    stop.append(gate)
                    return
                if gate.symbolic_parameters:  # conditioned on an earlier collapse
                    gate.substitute_symbols()
                yield from self.plan_ops(gate)

        while True:
            state = self._run_blocked(segment(), state, nqubits, zero, record)
            self._complete(state, nqubits, zero)  # (no blocked pass produced |0...0>, or left it unwritten)
            if not stop:
                if record is not None:
                    self._cache_put(key, record)
                return state
            state = stop.pop().apply(self, state, nqubits)
            zero[1] = None  # a special gate may touch any amplitude

    # -- compiled-program cache ---------------------------------------------
    @staticmethod
    def _param_key(p):
        if isinstance(p, (float, np.floating)):
            return float(p).hex()  # exact value, sign of zero included
        if isinstance(p, (int, np.integer, bool)):
            return int(p)
        if isinstance(p, (complex, np.complexfloating)):
            return (float(p.real).hex(), float(p.imag).hex())
        if isinstance(p, np.ndarray):
            return (p.dtype.str, p.shape, p.tobytes())
        if isinstance(p, torch.Tensor):
            a = p.detach().cpu().numpy()
            return (a.dtype.str, a.shape, a.tobytes())
        return None  # unknown parameter type: not cacheable

    def _gate_key(self, gate):
        # fast path: the key kept on an ordinary gate (see below) while its
        # parameter and qubit tuples are the same objects
        cached = getattr(gate, "__dict__", {}).get("_qkey")
        if cached is not None and (cached[0] is gate._parameters and cached[1] is gate._target_qubits
                                   and cached[2] is gate._control_qubits
                                   and cached[3] == gate.is_controlled_by):
            return cached[4]
        if kinds.is_measurement(gate):
            return None if gate.collapse else ("M",)
        if getattr(gate, "symbolic_parameters", None):
            # parameters that follow a measurement outcome are substituted at run
            # time (abstract.py:2584-2587): the values seen here may be stale
            return None
        if kinds.is_fused(gate):
            parts = tuple(self._gate_key(g) for g in gate.gates)
            return None if any(k is None for k in parts) else ("Fused", gate.target_qubits, parts)
        if kinds.runs_outside_passes(gate):
            return None  # callbacks and channels run outside the compiled passes
        # the key is kept on the gate while its parameter and qubit tuples are the
        # same objects (every setter assigns new ones); array parameters, which
        # can change in place, are described anew each time
        fp = (gate._parameters, gate._target_qubits, gate._control_qubits, gate.is_controlled_by)
        cached = gate.__dict__.get("_qkey")
        if (cached is not None and cached[0] is fp[0] and cached[1] is fp[1] and cached[2] is fp[2]
                and cached[3] == fp[3]):
            return cached[4]
        params = tuple(getattr(gate, "parameters", ()) or ())
        pk = tuple(self._param_key(x) for x in params)
        if any(k is None for k in pk):
            return None
        key = (type(gate).__name__, gate.target_qubits, gate.control_qubits, gate.is_controlled_by, pk)
        if not any(isinstance(x, (np.ndarray, torch.Tensor)) for x in params):
            gate.__dict__["_qkey"] = fp + (key,)
        return key

    def _queue_key(self, queue, nqubits, state, zero_pending):
        """Exact description of what run_queue would compile for ``queue``:
        every gate's class, qubits and parameter values, the register size,
        dtype, device, the pass options and whether the first pass starts from
        |0...0>. None when some gate (a callback, a collapsing measurement, an
        unknown parameter type) cannot be described."""
        parts = []
        for gate in queue:
            k = self._gate_key(gate)
            if k is None:
                return None
            parts.append(k)
        return (tuple(parts), nqubits, str(state.dtype), str(state.device), self.dtype,
                bool(zero_pending), bool(self.blocking_ordered), self._low_bits(self._dtype_code(state)))

    def _low_bits(self, code: int) -> int:
        """Index bits every blocked pass keeps in its tile: ``blocking_low_bits``,
        or 3 for complex128 and 4 for complex64 (128-byte segments) when None."""
        low = self.blocking_low_bits
        if low is None:
            return 3 if code == _lib.C128 else 4
        if isinstance(low, bool) or not isinstance(low, (int, np.integer)) or not 0 <= int(low) < blocking.TILE_BITS:
            raise_error(ValueError, f"blocking_low_bits must be an integer in [0, {blocking.TILE_BITS}), got {low!r}")
        return int(low)

    def _cache_get(self, key):
        """Compiled launches of ``key`` (marked most recently used), or None.
        The cache is shared by every thread that executes on this backend."""
        with self._programs_lock:
            actions = self._programs.get(key)
            if actions is not None:
                self._programs.move_to_end(key)
                self.program_cache_hits += 1
            return actions

    def _cache_put(self, key, record):
        with self._programs_lock:
            self._programs[key] = record
            self._programs.move_to_end(key)
            while len(self._programs) > self.program_cache_size:
                self._programs.popitem(last=False)

    def _replay(self, actions, queue, state, nqubits, zero):
        """Run the recorded launches of an identical queue: the same blocked
        programs (still on the device) and one-gate kernels, in the same order."""
        for gate in queue:
            if kinds.is_measurement(gate):
                gate.result.backend = self
        code = self._dtype_code(state)
        stream = self._stream(state)
        cur = torch.cuda.current_stream(state.device)
        for act in actions:
            if act[0] == "single":
                self._complete(state, nqubits, zero)
                self.apply_op(act[1], state, nqubits)
                if len(zero) > 1:
                    zero[1] = None
                continue
            _, tile_bits, prog, nbytes, count, ngates, host = act
            self._launch_blocked(state, nqubits, code, tile_bits, prog, nbytes, count, ngates, stream, cur, zero, host)
        self._complete(state, nqubits, zero)
        return state

    def _run_small(self, queue, state: torch.Tensor, nqubits: int) -> torch.Tensor:
        """Circuits on fewer than 12 qubits: run as one blocked pass on a
        12-qubit buffer whose extra (most significant) qubits stay |0>.

        The first 2^n amplitudes of the padded state are the n-qubit state, and
        qubit q becomes qubit q + p of the padded register, so the whole circuit
        is a single kernel launch. Same arithmetic as the per-gate kernels."""
        ops = []
        for gate in queue:
            if kinds.is_measurement(gate) and not gate.collapse:
                gate.result.backend = self
                continue
            if kinds.runs_outside_passes(gate):
                for g in queue:  # callbacks need the real state between gates
                    if g.symbolic_parameters:
                        g.substitute_symbols()
                    state = g.apply(self, state, nqubits)
                return state
            if gate.symbolic_parameters:
                gate.substitute_symbols()
            ops.extend(self.plan_ops(gate))
        return self._run_ops(ops, state, nqubits)

    def _run_ops(self, ops, state: torch.Tensor, nqubits: int) -> torch.Tensor:
        """Planned ops on an nqubits register, in blocked passes; below 12
        qubits the register is the low end of a 12-qubit buffer whose extra
        (most significant) qubits stay |0> (qubit q becomes qubit q + p)."""
        if not ops:
            return state
        if not self.blocking:
            for op in ops:
                self.apply_op(op, state, nqubits)
            return state
        if nqubits >= blocking.TILE_BITS:
            return self._run_blocked(ops, state, nqubits)
        pad = blocking.TILE_BITS - nqubits
        shifted = [_Op(op.kind, tuple(q + pad for q in op.targets), tuple(q + pad for q in op.controls), op.data)
                   for op in ops]
        big = torch.zeros(2**blocking.TILE_BITS, dtype=state.dtype, device=state.device)
        big[: state.numel()].copy_(state)
        big = self._run_blocked(shifted, big, blocking.TILE_BITS)
        state.copy_(big[: state.numel()])
        return state

    def _run_dm_cached(self, queue, state: torch.Tensor, nqubits: int) -> bool:
        """Run a density-matrix queue whose every element is a unitary gate, a
        non-collapsing measurement or a channel with a transfer matrix as one
        op list, reusing the compiled passes of an identical op list (the
        program cache of run_queue, keyed by every op's kind, qubits and exact
        data). Returns False, having done nothing, for other queues."""
        if not (self.program_cache and self.blocking and self.dm_superoperators) or 2 * nqubits < blocking.TILE_BITS:
            return False
        ops, mgates = [], []
        for gate in queue:
            if kinds.is_measurement(gate) and not gate.collapse:
                mgates.append(gate)
                continue
            if kinds.is_channel(gate):
                sop = self.channel_transfer_matrix(gate)
                if sop is None:
                    return False
                t = tuple(gate.target_qubits)
                ops.append(plan_matrix(sop, t + tuple(q + nqubits for q in t), ()))
                continue
            if kinds.runs_outside_passes(gate) or gate.symbolic_parameters:
                return False
            ops.extend(_dm_pair(self.plan_gate(gate), nqubits))
        for gate in mgates:
            gate.result.backend = self
        key = ("density", nqubits, str(state.dtype), str(state.device), bool(self.blocking_ordered),
               self._low_bits(self._dtype_code(state)),
               tuple((op.kind, op.targets, op.controls,
                      None if op.data is None else (op.data.dtype.str, op.data.shape, op.data.tobytes()))
                     for op in ops if op.kind != "noop"))
        flat = state.view(-1)
        actions = self._cache_get(key)
        if actions is not None:
            self._replay(actions, [], flat, 2 * nqubits, [False])
            return True
        record = []
        self._run_blocked(ops, flat, 2 * nqubits, None, record)
        self._cache_put(key, record)
        return True

    def run_queue_density_matrix(self, queue, state: torch.Tensor, nqubits: int) -> torch.Tensor:
        """Apply ``queue`` to a density matrix (the loop of abstract.py:3321-3322).

        Unitary gates become pairs of planned ops on the 2n-qubit vectorized
        matrix (U on the row qubits, conj(U) on the column qubits), grouped into
        blocked passes; channels, collapsing measurements and other special
        gates run in order between the passes through ``gate.apply``."""
        state = state.contiguous()
        self._dm_check(state, nqubits, "run_queue_density_matrix")
        if self._run_dm_cached(queue, state, nqubits):
            return state
        ops = []
        for gate in queue:
            if kinds.is_measurement(gate) and not gate.collapse:
                gate.result.backend = self
                continue
            if kinds.is_channel(gate) and self.dm_superoperators:
                sop = self.channel_transfer_matrix(gate)
                if sop is not None:
                    t = tuple(gate.target_qubits)
                    ops.append(plan_matrix(sop, t + tuple(q + nqubits for q in t), ()))
                    continue
            if kinds.runs_outside_passes(gate):
                self._run_ops(ops, state.view(-1), 2 * nqubits)
                ops = []
                state = gate.apply(self, state, nqubits).contiguous()
                continue
            if gate.symbolic_parameters:  # conditioned on an earlier collapse
                gate.substitute_symbols()
            ops.extend(_dm_pair(self.plan_gate(gate), nqubits))
        self._run_ops(ops, state.view(-1), 2 * nqubits)
        return state

    def _fill_zero(self, state: torch.Tensor, nqubits: int):
        _lib.check(self.lib.qsim_zero_state(state.data_ptr(), nqubits, self._dtype_code(state),
                                            self._stream(state)), self.lib)

    def _complete(self, state: torch.Tensor, nqubits: int, zero):
        """Write every amplitude of a state that passes from |0...0> have left
        partly unwritten (``zero`` as in run_queue), before an operation that
        may read any of them."""
        if zero[0]:
            self._fill_zero(state, nqubits)
            zero[0] = False
        elif len(zero) > 2 and zero[2]:
            _lib.check(self.lib.qsim_zero_outside_support(state.data_ptr(), nqubits, self._dtype_code(state),
                                                          zero[1], self._stream(state)), self.lib)
            zero[2] = False

    def _run_blocked(self, ops, state: torch.Tensor, nqubits: int, zero=None, record=None) -> torch.Tensor:
        """Partition ``ops`` (an iterable of planned gates) into passes and
        launch each pass as soon as it is compiled, so the host plans and
        compiles pass k+1 while the GPU runs pass k. ``zero`` = [True] means the
        state still stands for |0...0>: the first pass starts from it (or the
        state is filled before a one-gate launch); it is cleared once done.
        ``record`` (a list) receives the launches for the program cache."""
        code = self._dtype_code(state)
        cdtype = np.complex128 if code == _lib.C128 else np.complex64
        low = self._low_bits(code)
        ordered = self.blocking_ordered
        stream = self._stream(state)
        cur = torch.cuda.current_stream(state.device)
        for kind, payload in blocking.partition(ops, nqubits, low, ordered):
            if kind == "single":
                if zero is not None:
                    self._complete(state, nqubits, zero)
                self.apply_op(payload, state, nqubits)
                if zero is not None and len(zero) > 1:
                    zero[1] = None
                if record is not None:
                    record.append(("single", payload))
                continue
            tile_bits, pops = payload
            program, count, _ = blocking.compile_pass(tile_bits, pops, nqubits, cdtype, ordered)
            # zero padding: the kernel may read QSIM_BLOCKED_PAD bytes past the stream
            host = torch.frombuffer(bytearray(program) + bytes(blocking.PAD_BYTES), dtype=torch.uint8).pin_memory()
            prog = host.to(state.device, non_blocking=True)
            self._launch_blocked(state, nqubits, code, tuple(tile_bits), prog, len(program), count, len(pops),
                                 stream, cur, zero, program)
            if record is not None:
                record.append(("block", tuple(tile_bits), prog, len(program), count, len(pops), bytes(program)))
        return state

    # -- specialised blocked passes --------------------------------------------
    def _specialized(self, host, count, code, variant, device):
        """Handle of the specialised kernel of this program's record sequence
        (``variant``: _lib.SPEC_PASS, SPEC_PASS_FROM_ZERO or SPEC_ADJOINT), or
        0; counts the sighting and starts the compile in the background once
        the sequence has been launched ``specialize_after`` times."""
        if not self.specialize or host is None:
            return 0
        if not isinstance(host, (bytes, bytearray)):
            host = host.tobytes()
        handle, key_hash = ctypes.c_uint64(0), ctypes.c_uint64(0)
        rc = self.lib.qsim_blocked_specialized_lookup(host, len(host), count, code, int(variant),
                                                      ctypes.byref(handle), ctypes.byref(key_hash))
        if rc != 0:
            return 0  # not specialisable (a conditional mapping change)
        if handle.value:
            return handle.value
        key = key_hash.value
        with self._spec_lock:
            seen = self._spec_seen.get(key, 0) + 1
            self._spec_seen[key] = seen
            if seen < self.specialize_after or key in self._spec_jobs:
                return 0
            if self._spec_pool is None:
                from concurrent.futures import ThreadPoolExecutor

                self._spec_pool = ThreadPoolExecutor(max_workers=4, thread_name_prefix="qibo_amd_spec")
            self._spec_jobs[key] = self._spec_pool.submit(self._specialize_job, bytes(host), count, code,
                                                          int(variant), device)
        return 0

    def _specialize_job(self, host, count, code, variant, device):
        with torch.cuda.device(device):
            handle = ctypes.c_uint64(0)
            rc = self.lib.qsim_blocked_specialize(host, len(host), count, code, variant, ctypes.byref(handle))
            if rc != 0:
                log.warning("specialised blocked pass not built: %s", self.lib.qsim_last_error().decode())
            return rc

    def wait_specializations(self, timeout=None):
        """Wait for the specialised kernels being compiled; returns how many
        compiled successfully so far (all sequences, this backend)."""
        with self._spec_lock:
            jobs = list(self._spec_jobs.values())
        for job in jobs:
            job.result(timeout=timeout)
        return sum(1 for job in jobs if job.done() and job.result() == 0)

    def _launch_blocked(self, state, nqubits, code, tile_bits, prog, nbytes, count, ngates, stream, cur, zero,
                        host=None):
        timer = self._timer
        if timer is not None and timer[0](_BLOCK_OP):
            ev = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
            ev[0].record(cur)
        from_zero = zero is not None and zero[0]
        if from_zero and len(zero) > 2 and nqubits >= _FILL_FROM_ZERO_QUBITS and self.skip_zero_tiles:
            # |0...0> of a large state: write its one nonzero amplitude and run
            # the pass on tile 0 reading it alone; the other amplitudes stay
            # unwritten while the passes that follow keep to the support
            state[:1].fill_(1)
            zero[0] = from_zero = False
            zero[1] = 0
            zero[2] = True
        full = (1 << nqubits) - 1
        tmask = _TILE_MASKS.get(tile_bits)
        if tmask is None:
            tmask = _TILE_MASKS.setdefault(tuple(tile_bits), sum(1 << b for b in tile_bits))
        # amplitudes with a 1 outside the support are zero before the pass and
        # after it: only the tiles inside it run (from |0...0>: tile 0 alone)
        support = zero[1] if zero is not None and len(zero) > 1 and zero[1] is not None else full
        if not self.skip_zero_tiles:
            support = full
        if from_zero:
            support = 0  # (the from-zero entry points fill the state and run tile 0 alone)
        # sparse: tiles outside the support are skipped; partial: the input
        # is read only inside the support (the rest may be unwritten)
        partial = not from_zero and support != full
        sparse = partial and (support | tmask) != full
        variant = _lib.SPEC_PASS_FROM_ZERO if from_zero else _lib.SPEC_PASS
        handle = self._specialized(host, count, code, variant, state.device)
        tba = _lib.int_array(list(tile_bits))
        if handle:
            rc = self.lib.qsim_apply_blocked_specialized(handle, state.data_ptr(), 1, nqubits, code, tba,
                                                         support if partial else _ALL_BITS, prog.data_ptr(), 0,
                                                         nbytes, count, stream)
            self.specialized_launches += 1
        elif from_zero:
            rc = self.lib.qsim_apply_blocked_from_zero(state.data_ptr(), nqubits, code, tba, prog.data_ptr(), nbytes,
                                                       count, stream)
        elif partial:
            rc = self.lib.qsim_apply_blocked_support(state.data_ptr(), nqubits, code, tba, support, prog.data_ptr(),
                                                     nbytes, count, stream)
        else:
            rc = self.lib.qsim_apply_blocked(state.data_ptr(), nqubits, code, tba, prog.data_ptr(), nbytes, count,
                                             stream)
        if from_zero:
            zero[0] = False
        if zero is not None and len(zero) > 1:
            zero[1] = (support | tmask) if (support | tmask) != full else None
            if zero[1] is None and len(zero) > 2:
                zero[2] = False  # every tile ran: the whole state is written
        if sparse:
            self.sparse_launches += 1
        # algorithmic HBM bytes of the launch: the state read and written once
        # (from |0...0>: written once, by the fill), or the support read and
        # the tiles that contain it written
        isz = 16 if code == _lib.C128 else 8
        if from_zero:
            moved = (1 << nqubits) * isz
        elif partial:
            moved = ((1 << bin(support & full).count("1"))
                     + (1 << (bin(support & ~tmask & full).count("1") + len(tile_bits)))) * isz
        else:
            moved = 2 * (1 << nqubits) * isz
        self.block_bytes.append((moved, not (partial or from_zero)))
        _lib.check(rc, self.lib)
        if timer is not None and timer[0](_BLOCK_OP):
            ev[1].record(cur)
            timer[1].append(ev)
        _keep_alive(prog, state)
        self.block_stats.append(ngates)

    def execute_circuits(self, circuits, initial_states=None, nshots=1000, processes=None):
        if initial_states is None:
            initial_states = [None] * len(circuits)
        return [self.execute_circuit(c, s, nshots) for c, s in zip(circuits, initial_states)]

    # -- probabilities and sampling -------------------------------------------
    def calculate_probabilities(self, state, qubits, nqubits, density_matrix=False):
        if not isinstance(state, torch.Tensor) or state.device.type != "cuda":
            state = self.cast(state, dtype=_np_cdtype(state))
        qubits = [int(q) for q in qubits]
        code = self._dtype_code(state)
        rdt = torch.float64 if code == _lib.C128 else torch.float32
        m = len(qubits)
        out = torch.empty(2**m, dtype=rdt, device=state.device)
        if density_matrix:
            # |sum of the diagonal over the unmeasured qubits| (abstract.py:2742-2750)
            state = state.contiguous()
            self._dm_check(state, nqubits, "calculate_probabilities")
            rc = self.lib.qsim_dm_probabilities(state.data_ptr(), nqubits, code, _lib.int_array(qubits), m,
                                                out.data_ptr(), self._stream(state))
            _lib.check(rc, self.lib)
            return out
        if qubits == list(range(nqubits)) and nqubits <= 35:
            # every qubit in order: the pass also writes the tile statistics
            # that sample_shots would otherwise compute from p first
            tiles = torch.empty(self.lib.qsim_tile_stats_bytes(out.numel()), dtype=torch.uint8, device=state.device)
            rc = self.lib.qsim_probabilities_tiles(state.data_ptr(), nqubits, code, out.data_ptr(), tiles.data_ptr(),
                                                   tiles.numel(), self._stream(state))
            _lib.check(rc, self.lib)
            out._qsim_tiles = (tiles, out._version)
            return out
        wsb = self.lib.qsim_probabilities_workspace(nqubits, m)
        ws = self._workspace(wsb, "probs", state.device) if wsb else None
        rc = self.lib.qsim_probabilities(state.data_ptr(), nqubits, code, _lib.int_array(qubits), m,
                                         out.data_ptr(), ws.data_ptr() if ws is not None else None,
                                         wsb, self._stream(state))
        _lib.check(rc, self.lib)
        return out

    def _as_prob_tensor(self, probabilities) -> torch.Tensor:
        if isinstance(probabilities, torch.Tensor):
            p = probabilities.to(self.torch_device)
            if p.dtype not in (torch.float64, torch.float32):
                p = p.to(torch.float64)
            return p.contiguous()
        arr = np.asarray(probabilities)
        if arr.dtype not in (np.float64, np.float32):
            arr = arr.astype(np.float64)
        return torch.as_tensor(np.ascontiguousarray(arr), device=self.torch_device)

    def _validate_probabilities(self, p: torch.Tensor):
        """The checks np.random.choice applies to ``p`` before drawing (one
        pass over ``p``: sum, negative entries, NaNs)."""
        p = p.contiguous()
        n = p.numel()
        out = torch.empty(3, dtype=torch.float64, device=p.device)
        wsb = self.lib.qsim_prob_stats_workspace(n)
        ws = self._workspace(wsb, "pstats", p.device)
        pcode = _lib.F64 if p.dtype == torch.float64 else _lib.F32
        _lib.check(self.lib.qsim_prob_stats(p.data_ptr(), pcode, n, out.data_ptr(), ws.data_ptr(), wsb,
                                            self._stream(p)), self.lib)
        self._check_prob_stats(*out.tolist(), p.dtype)

    def exact_cdf(self, p: torch.Tensor) -> torch.Tensor:
        """Sequential-order cumulative sum of ``p`` as float64 (numpy's cumsum)."""
        n = p.numel()
        cdf = torch.empty(n, dtype=torch.float64, device=p.device)
        wsb = self.lib.qsim_cumsum_workspace(n)
        ws = self._workspace(wsb, "cumsum", p.device)
        pcode = _lib.F64 if p.dtype == torch.float64 else _lib.F32
        rc = self.lib.qsim_cumsum_exact(p.data_ptr(), pcode, n, cdf.data_ptr(), ws.data_ptr(), wsb,
                                        self._stream(p))
        _lib.check(rc, self.lib)
        return cdf

    def search_cdf(self, cdf: torch.Tensor, uniforms: np.ndarray) -> torch.Tensor:
        k = len(uniforms)
        out = torch.empty(k, dtype=torch.int64, device=cdf.device)
        if k == 0:
            return out
        u = torch.from_numpy(np.ascontiguousarray(uniforms, dtype=np.float64)).pin_memory()
        u = u.to(cdf.device, non_blocking=True)
        rc = self.lib.qsim_search_sorted(cdf.data_ptr(), cdf.numel(), u.data_ptr(), k, out.data_ptr(),
                                         self._stream(cdf))
        _lib.check(rc, self.lib)
        _keep_alive(u, cdf)
        return out

    @staticmethod
    def _check_prob_stats(total, nneg, nnan, dtype):
        """np.random.choice's checks on p (sum within sqrt(eps) of 1, no
        negative entry, no NaN), from the device's statistics of p."""
        atol = math.sqrt(np.finfo(np.float64).eps)
        if dtype == torch.float32:
            atol = max(atol, math.sqrt(np.finfo(np.float32).eps))
        if nnan or math.isnan(total):
            raise ValueError("probabilities contain NaN")
        if nneg:
            raise ValueError("probabilities are not non-negative")
        if abs(total - 1.0) > atol:
            raise ValueError("probabilities do not sum to 1")

    def sample_shots(self, probabilities, nshots: int):
        """``nshots`` indices drawn as np.random.choice(n, nshots, p=p) draws
        them from numpy's global generator (abstract.py:2774-2781), in one
        library call (``qsim_sample_exact``: the exact sequential cumsum read
        only where the bisections compare it). If p fails np.random.choice's
        checks the generator is put back, so a failed call consumes nothing,
        as the checks run before numpy's draws."""
        p = self._as_prob_tensor(probabilities)
        n, k = p.numel(), int(nshots)
        rng_state = np.random.get_state()
        u = np.random.random_sample(k)
        out = torch.empty(k, dtype=torch.int64, device=p.device)
        stats = torch.empty(3, dtype=torch.float64, device=p.device)
        ud = torch.from_numpy(np.ascontiguousarray(u, dtype=np.float64)).pin_memory().to(p.device, non_blocking=True)
        wsb = self.lib.qsim_sample_workspace(n)
        ws = self._workspace(wsb, "sample", p.device)
        pcode = _lib.F64 if p.dtype == torch.float64 else _lib.F32
        tiles = getattr(p, "_qsim_tiles", None)
        tiles = tiles[0].data_ptr() if tiles is not None and tiles[1] == p._version else None
        rc = self.lib.qsim_sample_exact(p.data_ptr(), pcode, n, tiles, ud.data_ptr(), k, out.data_ptr(),
                                        stats.data_ptr(), ws.data_ptr(), wsb, self._stream(p))
        _lib.check(rc, self.lib)
        _keep_alive(ud, p)
        try:
            self._check_prob_stats(*stats.tolist(), p.dtype)
        except ValueError:
            np.random.set_state(rng_state)
            raise
        return out

    def np_sum(self, p: torch.Tensor) -> torch.Tensor:
        """numpy's ``np.sum`` of a power-of-two-long probability vector, bit for bit
        (one device value of p's dtype)."""
        n = p.numel()
        out = torch.empty(1, dtype=p.dtype, device=p.device)
        wsb = self.lib.qsim_np_sum_workspace(n)
        ws = self._workspace(wsb, "npsum", p.device)
        pcode = _lib.F64 if p.dtype == torch.float64 else _lib.F32
        _lib.check(self.lib.qsim_np_sum(p.data_ptr(), pcode, n, out.data_ptr(), ws.data_ptr(), wsb,
                                        self._stream(p)), self.lib)
        return out

    def sample_frequencies(self, probabilities, nshots: int):
        """Frequencies of ``nshots`` draws (abstract.py:2760-2772).

        The normalisation ``p / sum(p)`` uses numpy's own summation order, and
        the draws the same batches of SHOT_BATCH_SIZE uniforms from numpy's
        global generator, so the counts equal the reference's for the same
        seed and probabilities."""
        p = self._as_prob_tensor(probabilities)
        n = p.numel()
        if n & (n - 1) == 0:
            total = self.np_sum(p)
            nprobs = torch.empty_like(p)
            pcode = _lib.F64 if p.dtype == torch.float64 else _lib.F32
            _lib.check(self.lib.qsim_divide_by(p.data_ptr(), pcode, n, total.data_ptr(), nprobs.data_ptr(),
                                               self._stream(p)), self.lib)
        else:
            nprobs = p / p.sum()
        self._validate_probabilities(nprobs)
        cdf = self.exact_cdf(nprobs)
        freqs = torch.zeros(nprobs.numel(), dtype=torch.int64, device=p.device)
        batches = [SHOT_BATCH_SIZE] * (nshots // SHOT_BATCH_SIZE) + [nshots % SHOT_BATCH_SIZE]
        for k in batches:
            samples = self.search_cdf(cdf, np.random.random_sample(k))
            if k:
                _lib.check(self.lib.qsim_histogram(samples.data_ptr(), k, freqs.data_ptr(),
                                                   freqs.numel(), self._stream(freqs)), self.lib)
        nz = torch.nonzero(freqs).flatten()
        vals = freqs[nz]
        return collections.Counter(dict(zip(nz.cpu().tolist(), vals.cpu().tolist())))

    def update_frequencies(self, frequencies, probabilities, nsamples):
        samples = self.sample_shots(probabilities, nsamples)
        res, counts = torch.unique(samples, return_counts=True)
        frequencies[res] += counts
        return frequencies

    def samples_to_binary(self, samples, nqubits):
        s = samples if isinstance(samples, torch.Tensor) else torch.as_tensor(np.asarray(samples))
        s = s.to(device=self.torch_device, dtype=torch.int64).contiguous()
        out = torch.empty((s.numel(), nqubits), dtype=torch.int64, device=s.device)
        _lib.check(self.lib.qsim_samples_to_binary(s.data_ptr(), s.numel(), nqubits, out.data_ptr(),
                                                   self._stream(s)), self.lib)
        return out

    def apply_bitflips(self, noiseless_samples, bitflip_probabilities):
        """Readout bitflips (abstract.py:2279-2295): one uniform per sample bit
        from numpy's global generator, in the samples' row-major order, then
        s + (1 - s) [u < p0] - s [u < p1] on the device."""
        s = noiseless_samples if isinstance(noiseless_samples, torch.Tensor) else \
            torch.as_tensor(np.asarray(noiseless_samples))
        s = s.to(device=self.torch_device, dtype=torch.int64).contiguous()
        if s.dim() != 2:
            raise_error(ValueError, "apply_bitflips expects samples of shape (nshots, nbits).")
        nshots, nbits = s.shape
        probs = np.asarray(self.to_numpy(bitflip_probabilities), dtype=np.float64).reshape(2, -1)
        if probs.shape[1] != nbits:
            raise_error(ValueError, f"{nbits} measured bits but {probs.shape[1]} bitflip probabilities.")
        u = np.random.random_sample((nshots, nbits))
        ud = torch.from_numpy(np.ascontiguousarray(u)).pin_memory().to(s.device, non_blocking=True)
        p0 = np.ascontiguousarray(probs[0])
        p1 = np.ascontiguousarray(probs[1])
        out = torch.empty_like(s)
        _lib.check(self.lib.qsim_apply_bitflips(s.data_ptr(), nshots, nbits, ud.data_ptr(), p0.ctypes.data,
                                                p1.ctypes.data, out.data_ptr(), self._stream(s)), self.lib)
        _keep_alive(ud, out)
        return out

    def samples_to_decimal(self, samples, nqubits):
        s = samples if isinstance(samples, torch.Tensor) else torch.as_tensor(np.asarray(samples))
        s = s.to(device=self.torch_device, dtype=torch.int64)
        w = (2 ** torch.arange(nqubits - 1, -1, -1, device=s.device, dtype=torch.int64))
        # int32 arithmetic of the reference (wraps for >= 32 measured qubits)
        return (s * w).sum(dim=1).to(torch.int32)

    def calculate_frequencies(self, samples):
        if isinstance(samples, (list, tuple)) and samples and isinstance(samples[0], str):
            # bitstrings of repeated execution (abstract.py:2727-2732 on a list of str)
            res, counts = np.unique(samples, return_counts=True)
            return collections.Counter(dict(zip(res, counts)))
        s = samples if isinstance(samples, torch.Tensor) else torch.as_tensor(np.asarray(samples))
        res, counts = torch.unique(s, return_counts=True)
        return collections.Counter(dict(zip(res.cpu().tolist(), counts.cpu().tolist())))

    def aggregate_shots(self, shots):
        return torch.stack([self.cast(s, dtype=s.dtype) for s in shots])

    # -- expectation values ------------------------------------------------------
    def expectation_pauli(self, state: torch.Tensor, nqubits: int, xm, zm, ny, coeffs) -> torch.Tensor:
        """Device float64 scalar ``sum_t c_t <psi|P_t|psi>`` (no host sync); for a
        density matrix ``sum_t c_t Re Tr(P_t rho)``."""
        nterms = len(coeffs)
        out = torch.empty(1, dtype=torch.float64, device=state.device)
        if state.dim() == 2:
            state = state.contiguous()
            self._dm_check(state, nqubits, "expectation")
            wsb = self.lib.qsim_dm_expval_workspace(nterms)
            ws = self._workspace(wsb, "dm_expval", state.device)
            c = np.ascontiguousarray(coeffs, dtype=np.float64)
            xm = np.ascontiguousarray(xm, dtype=np.uint64)
            zm = np.ascontiguousarray(zm, dtype=np.uint64)
            ny = np.ascontiguousarray(ny, dtype=np.int32)
            rc = self.lib.qsim_dm_expval_pauli(state.data_ptr(), nqubits, self._dtype_code(state), nterms,
                                               xm.ctypes.data, zm.ctypes.data, ny.ctypes.data, c.ctypes.data,
                                               out.data_ptr(), ws.data_ptr(), wsb, self._stream(state))
            _lib.check(rc, self.lib)
            return out[0]
        wsb = self.lib.qsim_expval_workspace(nqubits, nterms)
        ws = self._workspace(wsb, "expval", state.device)
        c = np.ascontiguousarray(coeffs, dtype=np.float64)
        xm = np.ascontiguousarray(xm, dtype=np.uint64)
        zm = np.ascontiguousarray(zm, dtype=np.uint64)
        ny = np.ascontiguousarray(ny, dtype=np.int32)
        rc = self.lib.qsim_expval_pauli(state.data_ptr(), nqubits, self._dtype_code(state), nterms,
                                        xm.ctypes.data, zm.ctypes.data, ny.ctypes.data, c.ctypes.data,
                                        out.data_ptr(), ws.data_ptr(), wsb, self._stream(state))
        _lib.check(rc, self.lib)
        return out[0]

    def apply_pauli_sum(self, state: torch.Tensor, nqubits: int, xm, zm, ny, coeffs) -> torch.Tensor:
        """New state ``sum_t c_t P_t |psi>`` (the Hamiltonian-vector product)."""
        out = torch.empty_like(state)
        c = np.ascontiguousarray(coeffs, dtype=np.float64)
        xm = np.ascontiguousarray(xm, dtype=np.uint64)
        zm = np.ascontiguousarray(zm, dtype=np.uint64)
        ny = np.ascontiguousarray(ny, dtype=np.int32)
        rc = self.lib.qsim_apply_pauli_sum(state.data_ptr(), nqubits, self._dtype_code(state), len(c),
                                           xm.ctypes.data, zm.ctypes.data, ny.ctypes.data, c.ctypes.data,
                                           out.data_ptr(), self._stream(state))
        _lib.check(rc, self.lib)
        return out

    def inner_matrix(self, bra: torch.Tensor, ket: torch.Tensor, nqubits: int, targets, controls,
                     matrix) -> torch.Tensor:
        """Device float64 scalar ``Re <bra| M_targets |ket>`` over the controlled subspace."""
        code = self._dtype_code(ket)
        npdt = np.complex128 if code == _lib.C128 else np.complex64
        m = np.ascontiguousarray(matrix, dtype=npdt)
        out = torch.empty(1, dtype=torch.float64, device=ket.device)
        wsb = self.lib.qsim_expval_workspace(nqubits, 1)
        ws = self._workspace(wsb, "inner", ket.device)
        rc = self.lib.qsim_inner_matrix(bra.data_ptr(), ket.data_ptr(), nqubits, code,
                                        _lib.int_array(targets), len(targets), _lib.int_array(controls),
                                        len(controls), m.ctypes.data, out.data_ptr(), ws.data_ptr(), wsb,
                                        self._stream(ket))
        _lib.check(rc, self.lib)
        return out[0]

    def exp_value_observable_symbolic(self, circuit, terms, term_qubits, term_coefficients, nqubits):
        result = circuit._final_state if circuit._final_state is not None else self.execute_circuit(circuit)
        state = result.state()
        xm, zm, ny = self.pauli_tables(terms, term_qubits, nqubits)
        coeffs = [float(np.real(c)) for c in term_coefficients]
        return float(self.expectation_pauli(state, nqubits, xm, zm, ny, coeffs))


_MAX_QUBITS = 36
_MAX_DM_QUBITS = 17  # 2^34 elements: 256 GiB in complex128


def _dm_pair(op: _Op, nqubits: int) -> List[_Op]:
    """The planned gate on the row qubits and its complex conjugate on the
    column qubits (qubit q + n) of a vectorized density matrix."""
    if op.kind == "noop":
        return []
    data = None if op.data is None else np.conj(op.data)
    col = _Op(op.kind, tuple(q + nqubits for q in op.targets), tuple(q + nqubits for q in op.controls), data)
    return [op, col]
_BLOCK_OP = _Op("block", (), (), None)
_ALL_BITS = (1 << 64) - 1  # support mask of a pass over every tile
_FILL_FROM_ZERO_QUBITS = 22  # from this size on the first pass from |0...0> runs tile 0 alone
_TILE_MASKS = {}  # tile bits -> their index mask


def ctypes_stream(stream) -> int:
    return stream.cuda_stream


def _torch_dtype(dtype):
    if isinstance(dtype, torch.dtype):
        return dtype
    name = getattr(dtype, "name", None) or (dtype if isinstance(dtype, str) else np.dtype(dtype).name)
    table = {
        "complex128": torch.complex128, "complex64": torch.complex64, "float64": torch.float64,
        "float32": torch.float32, "int64": torch.int64, "int32": torch.int32, "int8": torch.int8,
    }
    if name not in table:
        raise_error(ValueError, f"Unsupported dtype {dtype}.")
    return table[name]


def _np_cdtype(array):
    dt = getattr(array, "dtype", None)
    if dt is not None and np.dtype(dt) == np.complex64:
        return "complex64"
    return "complex128"


def _keep_alive(tmp: torch.Tensor, user: torch.Tensor):
    """Keep a temporary device buffer alive until the stream has consumed it."""
    tmp.record_stream(torch.cuda.current_stream(user.device))


_BASE_MATRIX = (gates_mod.Gate.matrix, gates_mod.ParametrizedGate.matrix)
