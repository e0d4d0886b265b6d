"""Continuous-batching circuit server on one MI355X.

Requests (a circuit, a precision, and either an expectation of sum_i Z_i or a
number of shots) arrive over time. The server admits a waiting request as
soon as an execution slot and enough HBM for its state vector are free, runs
it on a HIP stream of its own so that its kernels overlap with those of the
requests already on the GPU, and completes it when the stream's event fires.
States stay resident in HBM from allocation to completion; nothing goes back
to the host but the final expectation value or shot indices.

* Admission by footprint: a request needs 2^n x itemsize bytes for its state
  plus the workspace of its measurement; the server tracks the bytes of all
  admitted requests against ``hbm_budget`` and admits first-fit from the
  queue (a large request at the head does not block small ones behind it for
  more than ``max_bypass`` admissions).
* Determinism: each request carries its own seed; its uniforms come from
  ``np.random.RandomState(seed)`` and all reductions have a fixed order, so a
  request's result is bitwise the same alone and in any mix.
* Latency: arrival -> completion, reported as p50 / p99 with throughput.

The executor interface is small (``start(request, stream) -> handle``,
``done(handle) -> bool``, ``result(handle)``) so that the admission logic can
be exercised on the CPU with a fake executor.
"""

import collections
import math
import time
from dataclasses import dataclass, field
from typing import Callable, List, Optional

import numpy as np


@dataclass
class Request:
    rid: int
    circuit: object
    dtype: str                 # "complex64" | "complex128"
    kind: str                  # "expval" | "shots"
    nshots: int = 0
    seed: int = 0
    arrival: float = 0.0       # seconds on the server clock
    # filled by the server
    admitted: float = math.nan
    completed: float = math.nan
    result: object = None

    @property
    def nqubits(self) -> int:
        return self.circuit.nqubits

    def footprint(self) -> int:
        itemsize = 16 if self.dtype == "complex128" else 8
        n = self.nqubits
        state = (2**n) * itemsize
        if self.kind == "shots":
            probs = (2**n) * (8 if itemsize == 16 else 4)
            cdf = (2**n) * 8
            return state + probs + cdf + 64 * (2 ** max(n - 12, 0)) + self.nshots * 16
        return state + 8 * 1024 * n


@dataclass
class ServeStats:
    latencies: List[float] = field(default_factory=list)
    wall: float = 0.0
    completed: int = 0
    peak_bytes: int = 0
    peak_inflight: int = 0

    def summary(self) -> dict:
        lat = np.asarray(self.latencies) if self.latencies else np.zeros(1)
        return {
            "circuits": self.completed,
            "wall_s": self.wall,
            "circuits_per_s": self.completed / self.wall if self.wall > 0 else 0.0,
            "p50_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "peak_hbm_bytes": self.peak_bytes,
            "peak_inflight": self.peak_inflight,
        }


class CircuitServer:
    def __init__(self, executor, hbm_budget: int, max_inflight: int = 8, max_bypass: int = 4,
                 clock: Callable[[], float] = time.perf_counter):
        self.executor = executor
        self.hbm_budget = int(hbm_budget)
        self.max_inflight = max_inflight
        self.max_bypass = max_bypass
        self.clock = clock

    def run(self, requests: List[Request], realtime: bool = True) -> ServeStats:
        """Serve ``requests`` (sorted by arrival); returns latency statistics.

        With ``realtime`` the arrival times are honoured against the clock
        (Poisson stream); without it every request is available at t = 0.
        """
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        queue: List[Request] = []
        inflight = []   # (request, handle, bytes)
        used = 0
        bypassed = 0
        stats = ServeStats()
        for r in requests:
            if r.footprint() > self.hbm_budget:
                raise MemoryError(f"request {r.rid} needs {r.footprint()} bytes, budget {self.hbm_budget}")
        t0 = self.clock()
        while pending or queue or inflight:
            now = self.clock() - t0
            while pending and (not realtime or pending[0].arrival <= now):
                queue.append(pending.popleft())
            # completions
            still = []
            for req, h, nbytes in inflight:
                if self.executor.done(h):
                    req.result = self.executor.result(h)
                    req.completed = self.clock() - t0
                    stats.latencies.append(req.completed - (req.arrival if realtime else 0.0))
                    stats.completed += 1
                    used -= nbytes
                else:
                    still.append((req, h, nbytes))
            inflight = still
            # admissions: first fit, bounded bypass of the head of the queue
            admitted_any = True
            while admitted_any and queue and len(inflight) < self.max_inflight:
                admitted_any = False
                for i, req in enumerate(queue):
                    if i > 0 and bypassed >= self.max_bypass:
                        break
                    nbytes = req.footprint()
                    if used + nbytes <= self.hbm_budget:
                        queue.pop(i)
                        bypassed = bypassed + 1 if i > 0 else 0
                        req.admitted = self.clock() - t0
                        h = self.executor.start(req)
                        inflight.append((req, h, nbytes))
                        used += nbytes
                        stats.peak_bytes = max(stats.peak_bytes, used)
                        stats.peak_inflight = max(stats.peak_inflight, len(inflight))
                        admitted_any = True
                        break
            if not admitted_any or not queue:
                self.executor.idle()
        stats.wall = self.clock() - t0
        return stats


def _zsum_tables(n):
    """Pauli tables of sum_i Z_i on n qubits (qsim_expval_pauli layout)."""
    xm = np.zeros(n, dtype=np.uint64)
    zm = np.asarray([1 << (n - 1 - q) for q in range(n)], dtype=np.uint64)
    ny = np.zeros(n, dtype=np.int32)
    return xm, zm, ny, np.ones(n, dtype=np.float64)


class _Tail:
    """Device buffers of one request's measurement (``<sum Z_i>`` or shot
    indices) and the launches that fill them: the same library calls as
    ``calculate_probabilities`` -> ``sample_shots`` and ``expectation_pauli``,
    into buffers owned here (so that a captured graph
    replays into them), ending with the copy of the result to pinned host
    memory."""

    def __init__(self, backend, n, dtype_code, kind, nshots):
        import torch

        from . import _lib

        self.lib = backend.lib
        self.n, self.code, self.kind, self.nshots = n, dtype_code, kind, nshots
        dev = backend.torch_device
        if kind == "expval":
            self.tables = _zsum_tables(n)
            self.wsb = self.lib.qsim_expval_workspace(n, n)
            self.ws = torch.empty(max(self.wsb, 256), dtype=torch.uint8, device=dev)
            self.out = torch.empty(1, dtype=torch.float64, device=dev)
            self.host = torch.empty(1, dtype=torch.float64, pin_memory=True)
        else:
            pdt = torch.float64 if dtype_code == _lib.C128 else torch.float32
            self.pcode = _lib.F64 if dtype_code == _lib.C128 else _lib.F32
            self.probs = torch.empty(2**n, dtype=pdt, device=dev)
            self.pwsb = self.lib.qsim_probabilities_workspace(n, n)
            self.pws = torch.empty(max(self.pwsb, 256), dtype=torch.uint8, device=dev)
            self.swsb = self.lib.qsim_sample_workspace(2**n)
            self.sws = torch.empty(max(self.swsb, 256), dtype=torch.uint8, device=dev)
            self.stats = torch.empty(3, dtype=torch.float64, device=dev)
            self.u_host = torch.empty(max(nshots, 1), dtype=torch.float64, pin_memory=True)
            self.u_dev = torch.empty(max(nshots, 1), dtype=torch.float64, device=dev)
            self.out = torch.empty(max(nshots, 1), dtype=torch.int64, device=dev)
            self.host = torch.empty(max(nshots, 1), dtype=torch.int64, pin_memory=True)
        self.qubits = None

    def set_uniforms(self, seed):
        if self.kind == "shots":
            self.u_host.numpy()[: self.nshots] = np.random.RandomState(seed).random_sample(self.nshots)

    def launch(self, state, stream_ptr):
        """Enqueue the measurement of ``state`` (2^n amplitudes) on the current stream."""
        from . import _lib

        lib, n = self.lib, self.n
        if self.kind == "expval":
            xm, zm, ny, c = self.tables
            rc = lib.qsim_expval_pauli(state.data_ptr(), n, self.code, n, xm.ctypes.data, zm.ctypes.data,
                                       ny.ctypes.data, c.ctypes.data, self.out.data_ptr(), self.ws.data_ptr(),
                                       self.wsb, stream_ptr)
            _lib.check(rc, lib)
        else:
            if self.qubits is None:
                self.qubits = _lib.int_array(range(n))
            rc = lib.qsim_probabilities(state.data_ptr(), n, self.code, self.qubits, n, self.probs.data_ptr(),
                                        self.pws.data_ptr(), self.pwsb, stream_ptr)
            _lib.check(rc, lib)
            self.u_dev.copy_(self.u_host, non_blocking=True)
            rc = lib.qsim_sample_exact(self.probs.data_ptr(), self.pcode, 2**n, None, self.u_dev.data_ptr(), self.nshots,
                                       self.out.data_ptr(), self.stats.data_ptr(), self.sws.data_ptr(), self.swsb,
                                       stream_ptr)
            _lib.check(rc, lib)
        self.host.copy_(self.out, non_blocking=True)

    def result(self):
        if self.kind == "expval":
            return float(self.host[0].item())
        return self.host.numpy()[: self.nshots].copy()


class _Graph:
    """A request structure's captured launch sequence: the upload of the
    compiled records from a pinned buffer, every blocked pass (the first from
    |0...0>), and the measurement. A later request of the same structure
    writes its own records and uniforms into the pinned inputs and replays it."""

    def __init__(self, signature, nbytes, state, prog_host, prog_dev, tail, graph):
        self.signature = signature
        self.nbytes = nbytes
        self.state, self.prog_host, self.prog_dev, self.tail, self.graph = state, prog_host, prog_dev, tail, graph
        self.busy = None  # event of the replay in flight


class GPUExecutor:
    """Runs admitted requests on the MI355X backend.

    * Host work -- gate planning and record compilation -- runs on worker
      threads (the native planner and compiler release the GIL), so the
      server's admission and completion loop never waits for it.
    * Every request runs on a HIP stream of its own, so its passes overlap
      with those of the other requests in flight.
    * Requests on fewer than 12 qubits are batched: each is one blocked pass
      on a 12-qubit buffer, and those ready at a flush run in one
      ``qsim_apply_blocked_batched`` launch (grid z = request) over one
      contiguous buffer, their programs padded to a common length with no-op
      records; each then measures its own slice.
    * Graphs: a request on 12..``graph_max_qubits`` qubits whose passes are
      all blocked and whose structure (the gates' classes and qubits, register
      size, dtype and measurement) ran before is captured as a HIP graph. A
      later request of the same structure -- a new parameter set of the same
      ansatz, or the same circuit again -- writes its compiled records and
      uniforms into the graph's pinned inputs and replays it: one launch for
      the program upload, every pass, the measurement and the result copy.
      Graph nodes are kernels and copies only.
    Results are bitwise the same on every path (the same kernels, records and
    reductions)."""

    def __init__(self, backends: dict, nstreams: int = 8, workers: int = 2, batch_small: bool = True,
                 graphs: bool = True, graph_max_qubits: int = 22, graph_cache: int = 64):
        import threading
        from concurrent.futures import ThreadPoolExecutor

        import torch

        self.torch = torch
        self.backends = backends                      # dtype -> MI355XBackend
        dev = next(iter(backends.values())).torch_device
        self.device = dev
        self.streams = [torch.cuda.Stream(device=dev) for _ in range(nstreams)]
        self.free_streams = list(range(nstreams))
        self.pool = ThreadPoolExecutor(max_workers=workers, thread_name_prefix="qsim-serve")
        self.batch_small = batch_small
        self.graphs = graphs
        self.graph_max_qubits = graph_max_qubits
        self.graph_cache = graph_cache
        self._graphs = collections.OrderedDict()      # structure key -> [_Graph]
        self._seen = collections.OrderedDict()        # structures run once (captured when seen again)
        self._lock = threading.Lock()
        self._small = []                              # handles of small requests awaiting a batched launch
        self.batch_stream = torch.cuda.Stream(device=dev)
        self.stats = collections.Counter()            # path counts: graph_capture, graph_replay, direct, batched
        # per completed request: (nqubits, host seconds from start to the last
        # launch, seconds from then to completion seen by done())
        self.timings = collections.deque(maxlen=1 << 16)

    def _count(self, key, k=1):
        with self._lock:  # worker threads count too
            self.stats[key] += k

    # -- the executor interface ------------------------------------------------
    def start(self, req: Request):
        h = {"req": req, "sid": None, "event": None, "tail": None, "keep": None, "future": None,
             "graph": None, "t_start": time.perf_counter(), "t_host": None}
        if self.batch_small and req.nqubits < 12:
            h["future"] = self.pool.submit(self._compile_small, req)
            with self._lock:
                self._small.append(h)
            return h
        h["sid"] = self.free_streams.pop()
        h["future"] = self.pool.submit(self._run, req, h)
        return h

    def done(self, h) -> bool:
        if "value" in h:
            return True
        fut = h["future"]
        if fut is not None and not fut.done():
            return False
        if fut is not None:
            fut.result()  # re-raise a failure of the host work
        if h["event"] is None and h["sid"] is None:
            self.flush_small()  # a compiled small request still waiting for its batch
        ev = h["event"]
        if ev is None or not ev.query():
            return False
        # read the result before the graph (and its pinned output) is released
        h["value"] = h["tail"].result()
        t_done = time.perf_counter()
        t_host = h["t_host"] if h["t_host"] is not None else t_done
        self.timings.append((h["req"].nqubits, t_host - h["t_start"], t_done - t_host))
        h["keep"] = h["tail"] = None
        if h["sid"] is not None:
            self.free_streams.append(h["sid"])
            h["sid"] = None
        if h["graph"] is not None:
            h["graph"].busy = None
            h["graph"] = None
        return True

    def result(self, h):
        return h["value"]

    def idle(self):
        self.flush_small()
        # the loop polls completions; a longer nap leaves the interpreter to
        # the worker threads' planning (the polling thread competes for it)
        time.sleep(500e-6)

    def close(self):
        self.pool.shutdown(wait=True)

    # -- requests on >= 12 qubits ---------------------------------------------
    @staticmethod
    def _structure(req):
        parts = tuple((type(g).__name__, tuple(g.qubits)) for g in req.circuit.queue)
        return (req.nqubits, req.dtype, req.kind, req.nshots, parts)

    def _compile(self, backend, req):
        from . import _lib, blocking

        n = req.nqubits
        ops = [op for g in req.circuit.queue for op in backend.plan_ops(g)]
        code = _lib.C128 if req.dtype == "complex128" else _lib.C64
        cdtype = np.complex128 if code == _lib.C128 else np.complex64
        plan, program = blocking.compile_circuit(ops, n, cdtype, backend._low_bits(code),
                                                 backend.blocking_ordered)
        return code, plan, program

    def _run(self, req, h):
        torch = self.torch
        from . import blocking

        backend = self.backends[req.dtype]
        stream = self.streams[h["sid"]]
        n = req.nqubits
        if n < blocking.TILE_BITS:
            self._run_small_alone(req, h, backend, stream)
            return
        if not self.graphs or n > self.graph_max_qubits:
            # large registers: the backend's own loop, which launches each pass
            # as soon as it is compiled (planning overlaps the GPU) and replays
            # the compiled passes of a repeated circuit
            from . import _lib

            code = _lib.C128 if req.dtype == "complex128" else _lib.C64
            with torch.cuda.stream(stream):
                state = backend._alloc_state(n, req.dtype)
                state = backend.run_queue(list(req.circuit.queue), state, n, zero_pending=True)
                tail = _Tail(backend, n, code, req.kind, req.nshots)
                tail.set_uniforms(req.seed)
                tail.launch(state, stream.cuda_stream)
                ev = torch.cuda.Event()
                ev.record(stream)
            h.update(tail=tail, event=ev, keep=(state,), t_host=time.perf_counter())
            self._count("direct", 1)
            return
        code, plan, program = self._compile(backend, req)
        all_blocked = plan and all(p[0] == "block" for p in plan)
        signature = tuple((p[1], p[2], p[3]) for p in plan if p[0] == "block")
        key = self._structure(req) if (self.graphs and all_blocked and n <= self.graph_max_qubits) else None
        with torch.cuda.stream(stream):
            sptr = stream.cuda_stream
            if key is not None:
                g = self._graph_for(key, signature, len(program))
                if g is not None:  # replay with this request's records and uniforms
                    host = g.prog_host.numpy()
                    host[: len(program)] = np.frombuffer(program, dtype=np.uint8)
                    g.tail.set_uniforms(req.seed)
                    g.graph.replay()
                    ev = torch.cuda.Event()
                    ev.record(stream)
                    g.busy = ev
                    h.update(tail=g.tail, graph=g, event=ev, t_host=time.perf_counter())
                    self._count("graph_replay", 1)
                    return
                g = self._capture(backend, req, key, code, plan, program, signature) if self._repeat(key) else None
                if g is not None:
                    g.tail.set_uniforms(req.seed)
                    g.graph.replay()
                    ev = torch.cuda.Event()
                    ev.record(stream)
                    g.busy = ev
                    h.update(tail=g.tail, graph=g, event=ev, t_host=time.perf_counter())
                    self._count("graph_capture", 1)
                    return
            # direct launches
            state = backend._alloc_state(n, req.dtype)
            zero = [True]
            if program:
                host = torch.frombuffer(bytearray(program) + bytes(blocking.PAD_BYTES), dtype=torch.uint8)
                prog = host.pin_memory().to(backend.torch_device, non_blocking=True)
            else:
                prog = None
            cur = torch.cuda.current_stream(backend.torch_device)
            for p in plan:
                if p[0] == "single":
                    if zero[0]:
                        backend._fill_zero(state, n)
                        zero[0] = False
                    backend.apply_op(p[1], state, n)
                else:
                    _, tile_bits, off, count, ngates = p
                    backend._launch_blocked(state, n, code, tile_bits, prog[off:], self._pass_bytes(plan, p, program),
                                            count, ngates, sptr, cur, zero)
            if zero[0]:
                backend._fill_zero(state, n)
            tail = _Tail(backend, n, code, req.kind, req.nshots)
            tail.set_uniforms(req.seed)
            tail.launch(state, sptr)
            ev = torch.cuda.Event()
            ev.record(stream)
            h.update(tail=tail, event=ev, keep=(state, prog), t_host=time.perf_counter())
            self._count("direct", 1)

    @staticmethod
    def _pass_bytes(plan, p, program):
        """Bytes of pass p's records in the concatenated program."""
        offs = [q[2] for q in plan if q[0] == "block"] + [len(program)]
        i = offs.index(p[2])
        return offs[i + 1] - p[2]

    def _repeat(self, key) -> bool:
        """True when the structure ran before: a graph is captured only for a
        structure seen a second time (a one-off circuit is not worth one)."""
        with self._lock:
            if key in self._seen:
                return True
            self._seen[key] = True
            while len(self._seen) > 4096:
                self._seen.popitem(last=False)
            return False

    def _graph_for(self, key, signature, nbytes):
        with self._lock:
            for g in self._graphs.get(key, ()):
                if g.busy is None and g.signature == signature and g.nbytes == nbytes:
                    g.busy = True  # claimed until the replay's event is recorded
                    self._graphs.move_to_end(key)
                    return g
        return None

    def _capture(self, backend, req, key, code, plan, program, signature):
        torch = self.torch
        from . import _lib, blocking

        n = req.nqubits
        with self._lock:
            total = sum(len(v) for v in self._graphs.values())
            if total >= self.graph_cache:
                return None
        state = backend._alloc_state(n, req.dtype)
        prog_host = torch.zeros(len(program) + blocking.PAD_BYTES, dtype=torch.uint8, pin_memory=True)
        prog_host.numpy()[: len(program)] = np.frombuffer(program, dtype=np.uint8)
        prog_dev = torch.empty_like(prog_host, device=backend.torch_device)
        tail = _Tail(backend, n, code, req.kind, req.nshots)
        graph = torch.cuda.CUDAGraph()
        side = torch.cuda.Stream(device=backend.torch_device)
        side.wait_stream(torch.cuda.current_stream(backend.torch_device))
        # Captured with capture_begin / capture_end on a side stream, not with
        # the torch.cuda.graph context (whose entry synchronizes the device and
        # empties the allocator cache: under load that stalled every request in
        # flight). Thread-local capture: the other workers and the server loop
        # keep launching and polling their own streams meanwhile.
        with torch.cuda.stream(side):
            graph.capture_begin(capture_error_mode="thread_local")
            try:
                self._launch_captured(backend, n, code, plan, program, state, prog_host, prog_dev, tail,
                                      side.cuda_stream)
            except BaseException:
                try:
                    graph.capture_end()  # leave the stream usable; the graph is dropped
                except Exception:  # noqa: BLE001
                    pass
                raise
            graph.capture_end()
        g = _Graph(signature, len(program), state, prog_host, prog_dev, tail, graph)
        g.busy = True
        with self._lock:
            self._graphs.setdefault(key, []).append(g)
            self._graphs.move_to_end(key)
        return g

    def _launch_captured(self, backend, n, code, plan, program, state, prog_host, prog_dev, tail, sptr):
        """The launches a graph records: program upload, passes, measurement."""
        from . import _lib

        prog_dev.copy_(prog_host, non_blocking=True)
        zero = True
        for p in plan:
            _, tile_bits, off, count, ngates = p
            launch = backend.lib.qsim_apply_blocked_from_zero if zero else backend.lib.qsim_apply_blocked
            zero = False
            rc = launch(state.data_ptr(), n, code, _lib.int_array(list(tile_bits)), prog_dev.data_ptr() + off,
                        self._pass_bytes(plan, p, program), count, sptr)
            _lib.check(rc, backend.lib)
        tail.launch(state, sptr)

    # -- requests below 12 qubits: one batched launch ----------------------------
    def _run_small_alone(self, req, h, backend, stream):
        """A request below 12 qubits on its own stream: the same padded
        12-qubit pass as in a batched launch (so the results agree bitwise)."""
        torch = self.torch
        from . import _lib, blocking

        code, compiled = self._compile_small(req)
        n = req.nqubits
        with torch.cuda.stream(stream):
            sptr = stream.cuda_stream
            if compiled is None:
                state = backend.zero_state(n)
                state = backend.run_queue(list(req.circuit.queue), state, n)
                keep = (state,)
            else:
                program, count = compiled
                tdt = torch.complex128 if code == _lib.C128 else torch.complex64
                big = torch.zeros(2**blocking.TILE_BITS, dtype=tdt, device=self.device)
                big[0] = 1
                keep = (big,)
                if count:
                    host = torch.frombuffer(bytearray(program) + bytes(blocking.PAD_BYTES), dtype=torch.uint8)
                    prog = host.pin_memory().to(self.device, non_blocking=True)
                    rc = backend.lib.qsim_apply_blocked(big.data_ptr(), blocking.TILE_BITS, code,
                                                        _lib.int_array(range(blocking.TILE_BITS)), prog.data_ptr(),
                                                        len(program), count, sptr)
                    _lib.check(rc, backend.lib)
                    keep = (big, prog)
                state = big[: 2**n]
            tail = _Tail(backend, n, code, req.kind, req.nshots)
            tail.set_uniforms(req.seed)
            tail.launch(state, sptr)
            ev = torch.cuda.Event()
            ev.record(stream)
        h.update(tail=tail, event=ev, keep=keep, t_host=time.perf_counter())
        self._count("small_alone", 1)

    def _compile_small(self, req):
        from . import _lib, blocking
        from .backend import _Op

        backend = self.backends[req.dtype]
        n = req.nqubits
        pad = blocking.TILE_BITS - n
        ops = []
        for g in req.circuit.queue:
            for op in backend.plan_ops(g):
                ops.append(_Op(op.kind, tuple(q + pad for q in op.targets), tuple(q + pad for q in op.controls),
                               op.data))
        code = _lib.C128 if req.dtype == "complex128" else _lib.C64
        cdtype = np.complex128 if code == _lib.C128 else np.complex64
        plan, program = blocking.compile_circuit(ops, blocking.TILE_BITS, cdtype, backend._low_bits(code),
                                                 backend.blocking_ordered)
        if len(plan) > 1 or (plan and plan[0][0] != "block"):
            return code, None  # not a single pass (e.g. a 4-qubit dense block): run it alone
        count = plan[0][3] if plan else 0
        return code, (program, count)

    def flush_small(self):
        """Launch every compiled small request in one batched pass per dtype."""
        torch = self.torch
        from . import _lib, blocking

        with self._lock:
            ready = [h for h in self._small if h["future"].done()]
            if not ready:
                return
            self._small = [h for h in self._small if not h["future"].done()]
        groups = collections.defaultdict(list)
        for h in ready:
            code, compiled = h["future"].result()
            if compiled is None:  # not one pass: run on its own stream
                req = h["req"]
                backend = self.backends[req.dtype]
                with torch.cuda.stream(self.batch_stream):
                    state = backend.zero_state(req.nqubits)
                    state = backend.run_queue(list(req.circuit.queue), state, req.nqubits)
                    tail = _Tail(backend, req.nqubits, code, req.kind, req.nshots)
                    tail.set_uniforms(req.seed)
                    tail.launch(state, self.batch_stream.cuda_stream)
                    ev = torch.cuda.Event()
                    ev.record(self.batch_stream)
                h.update(tail=tail, keep=(state,), event=ev, future=None, t_host=time.perf_counter())
                self._count("small_alone", 1)
                continue
            groups[code].append((h, compiled))
        for code, members in groups.items():
            backend = self.backends[members[0][0]["req"].dtype]
            nb = len(members)
            maxcount = max(c for _, (_, c) in members)
            maxbytes = max(len(p) + (maxcount - c) * blocking.HEADER_BYTES for _, (p, c) in members)
            stride = (maxbytes + 255) // 256 * 256
            progs = np.zeros(stride * nb + blocking.PAD_BYTES, dtype=np.uint8)
            for i, (_, (p, c)) in enumerate(members):
                progs[i * stride: i * stride + len(p)] = np.frombuffer(p, dtype=np.uint8)
                # no-op records up to the common count (header only, no entries)
                for k in range(maxcount - c):
                    progs[i * stride + len(p) + k * blocking.HEADER_BYTES] = blocking.OP_NOP
            tdt = torch.complex128 if code == _lib.C128 else torch.complex64
            with torch.cuda.stream(self.batch_stream):
                sptr = self.batch_stream.cuda_stream
                states = torch.zeros(nb, 2**blocking.TILE_BITS, dtype=tdt, device=self.device)
                states[:, 0] = 1
                prog = torch.from_numpy(progs).pin_memory().to(self.device, non_blocking=True)
                if maxcount:
                    rc = backend.lib.qsim_apply_blocked_batched(
                        states.data_ptr(), nb, blocking.TILE_BITS, code,
                        _lib.int_array(range(blocking.TILE_BITS)), prog.data_ptr(), stride, maxbytes, maxcount, sptr)
                    _lib.check(rc, backend.lib)
                evs = []
                for i, (h, _) in enumerate(members):
                    req = h["req"]
                    tail = _Tail(backend, req.nqubits, code, req.kind, req.nshots)
                    tail.set_uniforms(req.seed)
                    tail.launch(states[i, : 2**req.nqubits], sptr)
                    h.update(tail=tail, keep=(states, prog), future=None, t_host=time.perf_counter())
                    evs.append(h)
                ev = torch.cuda.Event()
                ev.record(self.batch_stream)
            for h in evs:
                h["event"] = ev
                h["t_host"] = time.perf_counter()
            self._count("batched", nb)
            self._count("batched_launches", 1)


def poisson_stream(nrequests: int, qps: float, seed: int, nmin: int = 10, nmax: int = 30,
                   dmin: int = 10, dmax: int = 50) -> List[Request]:
    """The mixed stream of the serving benchmark: n ~ U[nmin, nmax], depth ~
    U[dmin, dmax] layers of the seeded layered generator, half complex64 and
    half complex128, half <sum Z_i> and half nshots ~ U[1000, 10000]."""
    from .models import random_layered_circuit

    rng = np.random.default_rng(seed)
    t = 0.0
    reqs = []
    for i in range(nrequests):
        t += rng.exponential(1.0 / qps)
        n = int(rng.integers(nmin, nmax + 1))
        depth = int(rng.integers(dmin, dmax + 1))
        dtype = "complex64" if rng.random() < 0.5 else "complex128"
        kind = "expval" if rng.random() < 0.5 else "shots"
        nshots = int(rng.integers(1000, 10001)) if kind == "shots" else 0
        circ = random_layered_circuit(n, depth, seed=int(rng.integers(1 << 30)), measure=False)
        reqs.append(Request(i, circ, dtype, kind, nshots, seed=int(rng.integers(1 << 30)), arrival=t))
    return reqs
