"""Benchmark of the MI355X state-vector backend.

Workload (one "step"): the seeded random layered circuit -- per layer RY on
every qubit, RZ on every qubit, CZ on the brick pairs -- on 30 qubits in
complex128, 20 layers (600 RY + 600 RZ + 290 CZ = 1490 gate applications),
followed by M(all) with 100 000 shots: the final state, its probabilities
and the sampled bitstrings, through the drop-in path ``circuit(nshots=...)``
-> ``CircuitResult.samples()``. Inputs are synthetic (seeded angles); the
state stays resident in HBM for the whole step.

Reported: gate applications per second (``value``); the HBM bytes the
step's kernels must move (one read and one write of the state per blocked
pass, the first pass only writing; the probabilities and the CDF) over the
step time; the roofline of the dominant kernel (``k_blocked``: one blocked
pass of many gates, 2 * 2^n * 16 bytes per launch) timed with HIP events on
its stream inside the timed region; and a CPU baseline: the numpy oracle of
qibo's NumpyBackend timed on a bounded sample of the same circuit's gates on
this host. The timed steps execute the same circuit again, so they replay the
blocked passes compiled at the first execution (the backend's program cache);
the same steps with the cache off -- planning and compiling every step -- are
reported beside them.

Beside the headline line the same JSON object carries the other parts of
the metric, each on its own bounded workload:
* "layered28": the same step at 28 qubits (20 layers, 100 000 shots);
* "qft": qibo's QFT circuit at 30 qubits in complex128 and complex64;
* "qft15": qibo's QFT at 15 qubits, complex128, microseconds per circuit;
* "precision": complex64 and its ``fuse(max_qubits=5)`` form against the
  complex128 path at 32 qubits (layered circuit, QFT, dense 5-qubit blocks on
  MFMA), infidelity and max-abs amplitude error at 30 qubits, and the target
  check (>= 2x at infidelity <= 1e-6) per workload, passed or failed;
* "serving": a Poisson stream of layered circuits (n ~ U[10, 30], depth ~
  U[10, 50], half complex64, half <sum Z> / shots) through the continuous-
  batching server: circuits/s, p50 and p99 latency;
* "adjoint": adjoint gradients of the TFIM energy for the 20-qubit, 10-layer
  RY/RZ + CNOT-ladder ansatz over 64 parameter sets, gradients/s;
* "density": the layered circuit with a depolarizing channel on every qubit
  after every layer on a 14-qubit density matrix (4 GiB, complex128):
  operations/s, the depolarizing pass in GB/s, and the oracle beside the GPU
  on the same circuit at 9 qubits.

    python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--nqubits 30] [--no-cpu] [--no-extras]
                    [--extras NAME,NAME,...]
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published


def cpu_baseline(nqubits, nlayers, seed, budget_s):
    """Oracle (numpy NumpyBackend restatement) on a stratified gate sample, one process."""
    import numpy as np

    from oracle import OracleBackend
    from qibo_amd import models

    blas_name = "unknown"
    try:
        from threadpoolctl import threadpool_info

        info = [i for i in threadpool_info() if i.get("user_api") == "blas"]
        blas_threads = max((i.get("num_threads", 1) for i in info), default=1)
        if info:
            blas_name = f"{info[0].get('internal_api')} {info[0].get('version')}"
    except Exception:  # noqa: BLE001
        blas_threads = 1
    c = models.random_layered_circuit(nqubits, nlayers, seed=seed, measure=False)
    layer0 = list(c.queue)[: 2 * nqubits + (nqubits // 2)]
    ry, rz, cz = layer0[:nqubits], layer0[nqubits: 2 * nqubits], layer0[2 * nqubits:]
    order = []
    for q in (nqubits - 1, 0, nqubits // 2, nqubits - 2, 1, nqubits // 3):
        order += [ry[q], rz[q]]
        if cz:
            order.append(cz[min(q // 2, len(cz) - 1)])
    ob = OracleBackend()
    state = ob.zero_state(nqubits)
    done, t0 = 0, time.perf_counter()
    names = []
    for g in order:
        state = ob.apply_gate(g, state, nqubits)
        done += 1
        names.append(f"{g.__class__.__name__}{g.qubits}")
        if time.perf_counter() - t0 > budget_s:
            break
    elapsed = time.perf_counter() - t0
    return {
        "value": done / elapsed,
        "unit": "gate-apps/s",
        "cores": int(blas_threads),
        "kind": "port",
        "host": {"os_cpu_count": os.cpu_count(), "blas": blas_name,
                 "OPENBLAS_NUM_THREADS": os.environ.get("OPENBLAS_NUM_THREADS")},
        "sample": f"{done} gates of layer 0 at n={nqubits} complex128 ({', '.join(names)}); "
                  f"{elapsed:.1f} s; numpy {np.__version__} matmul with {blas_threads} BLAS threads, "
                  "transposes single-threaded",
    }


def _timed(fn):
    import torch

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = fn()
    torch.cuda.synchronize()
    return res, time.perf_counter() - t0


def _block_circuit(n, nblocks, seed):
    """Brickwork of Haar-random 5-qubit unitaries (dense 32 x 32 blocks, the
    shape Circuit.fuse(max_qubits=5) produces on dense circuits)."""
    import numpy as np

    from qibo_amd import Circuit, gates

    rng = np.random.default_rng(seed)
    c = Circuit(n)
    starts = list(range(0, n - 4, 5))
    k = 0
    while k < nblocks:
        for s0 in starts:
            if k == nblocks:
                break
            z = (rng.normal(size=(32, 32)) + 1j * rng.normal(size=(32, 32))) / np.sqrt(2)
            q, r = np.linalg.qr(z)
            u = q * (np.diagonal(r) / np.abs(np.diagonal(r)))
            off = (k // len(starts)) % 5
            qubits = [(s0 + off + i) % n for i in range(5)]
            c.add(gates.Unitary(u, *qubits, check_unitary=False))
            k += 1
    return c


def _time_forms(forms):
    """Seconds of one execution of each (key, backend, circuit): one warm-up
    run (allocator, program cache), then a timed run that reuses the state
    block the warm-up left in torch's cache."""
    import torch

    out = {}
    for key, backend, circ in forms:
        circ._final_state = None
        backend.execute_circuit(circ)
        circ._final_state = None
        r, dt = _timed(lambda: backend.execute_circuit(circ))
        del r
        circ._final_state = None
        torch.cuda.empty_cache()
        out[key] = dt
    return out


def _accuracy(b64, b128, c64_circ, c128_circ):
    """qibo's infidelity 1 - |<a|b>|^2, the same for the normalised states and
    the max-abs amplitude error of the complex64 execution of ``c64_circ``
    against the complex128 execution of ``c128_circ`` (the same circuit,
    possibly in another fused form), computed on the GPU."""
    import torch

    from qibo_amd.accuracy import compare_states

    a = b64.execute_circuit(c64_circ).state()
    c64_circ._final_state = None
    b = b128.execute_circuit(c128_circ).state()
    c128_circ._final_state = None
    rep = compare_states(b64, a, b, reference="complex128 GPU execution (NumpyBackend to 1e-12)")
    del a, b
    torch.cuda.empty_cache()
    return rep


def _blocks5_timing(b64, b128, seed, n, nblocks):
    """Dense 5-qubit blocks at n qubits: one qsim_apply_fused launch each,
    complex64 on FP32 MFMA and complex128 on FP64 MFMA."""
    blocks = _block_circuit(n, nblocks, seed)
    t = _time_forms([("c128_s", b128, blocks), ("c64_s", b64, blocks)])
    blk = {"nqubits": n, "blocks": nblocks, **t, "speedup_c64_vs_c128": t["c128_s"] / t["c64_s"]}
    flops = 8 * 32 * 2**n  # 32 complex multiply-adds per amplitude
    for key, isz, peak in (("c64", 8, 157.3), ("c128", 16, 78.6)):
        per = t[f"{key}_s"] / nblocks
        blk[key] = {"ms_per_block": 1e3 * per, "hbm_GBs": 2 * 2**n * isz / per / 1e9,
                    "TFLOPs": flops / per / 1e12, "mfma_peak_TFLOPs": peak}
    return blk


def bench_blocks5(seed, n=32, nblocks=24):
    """The dense 5-qubit block workload alone (the MFMA kernels' profile runs)."""
    from qibo_amd.backend import MI355XBackend

    return _blocks5_timing(MI355XBackend(dtype="complex64"), MI355XBackend(dtype="complex128"), seed, n, nblocks)


def bench_precision(seed, n_time=32, n_acc=30, nlayers=20, nblocks=24):
    """complex64 and its fused form against the plain complex128 GPU path.

    Three workloads, each timed at ``n_time`` qubits in every form and checked
    for accuracy at ``n_acc`` qubits (the complex128 state of 2^32 amplitudes
    is 64 GiB, the accuracy pair at 30 qubits 24 GiB):
    * "layered": the layered RY/RZ/CZ generator, 20 layers -- unfused
      complex128, unfused complex64, and ``circuit.fuse(max_qubits=5)`` in
      complex64 (the configuration as written);
    * "qft": qibo's QFT -- unfused complex128 against ``QFT.fuse(max_qubits=5)``
      in complex64 (its seven 5-qubit groups run on MFMA);
    * "blocks5": 24 Haar-random dense 5-qubit blocks -- complex64 on FP32 MFMA
      against complex128 on FP64 MFMA, with ms per block, GB/s and TFLOP/s.
    The check the target names -- the complex64 form at least 2x the complex128
    path at infidelity <= 1e-6 -- is stated per workload, passed or failed."""
    import torch

    from qibo_amd import models
    from qibo_amd.backend import MI355XBackend

    b64 = MI355XBackend(dtype="complex64")
    b128 = MI355XBackend(dtype="complex128")
    b64.specialize = b128.specialize = False  # interpreted passes: the forms compared are the kernels'
    res = {"nqubits_timed": n_time, "nqubits_accuracy": n_acc}

    # layered circuit, as configured
    lay = models.random_layered_circuit(n_time, nlayers, seed=seed, measure=False)
    lay_f = lay.fuse(max_qubits=5)
    sizes = {}
    for g in lay_f.queue:
        sizes[len(g.qubits)] = sizes.get(len(g.qubits), 0) + 1
    t = _time_forms([("c128_s", b128, lay), ("c64_s", b64, lay), ("c64_fused5_s", b64, lay_f)])
    acc_c = models.random_layered_circuit(n_acc, nlayers, seed=seed, measure=False)
    res["layered"] = {
        "layers": nlayers, "gates": lay.ngates, "fused5_group_sizes": {str(k): v for k, v in sorted(sizes.items())},
        **t,
        "speedup_c64_vs_c128": t["c128_s"] / t["c64_s"],
        "speedup_c64_fused5_vs_c128": t["c128_s"] / t["c64_fused5_s"],
        "speedup_c64_fused5_vs_c64": t["c64_s"] / t["c64_fused5_s"],
        "accuracy_c64": _accuracy(b64, b128, acc_c, acc_c),
        "accuracy_c64_fused5": _accuracy(b64, b128, acc_c.fuse(max_qubits=5), acc_c),
    }
    del lay, lay_f

    # the unfused pair again with the passes specialised to their record
    # sequences (the path the headline step runs): compiled during a warm-up
    # run and waited for; bitwise the interpreted results, so the accuracy
    # above holds for them
    s64, s128 = MI355XBackend(dtype="complex64"), MI355XBackend(dtype="complex128")
    s64.specialize_after = s128.specialize_after = 1
    lay = models.random_layered_circuit(n_time, nlayers, seed=seed, measure=False)
    ts = {}
    for key, be in (("c128", s128), ("c64", s64)):
        lay._final_state = None
        be.execute_circuit(lay)
        be.wait_specializations()
        lay._final_state = None
        r, dt = _timed(lambda: be.execute_circuit(lay))
        del r
        lay._final_state = None
        torch.cuda.empty_cache()
        ts[key + "_s"] = dt
        ts[key + "_specialized_launches"] = be.specialized_launches
    res["layered_specialized"] = {**ts, "speedup_c64_vs_c128": ts["c128_s"] / ts["c64_s"]}
    del lay, s64, s128

    # QFT and its fused form
    q = models.QFT(n_time)
    qf = q.fuse(max_qubits=5)
    sizes = {}
    for g in qf.queue:
        sizes[len(g.qubits)] = sizes.get(len(g.qubits), 0) + 1
    t = _time_forms([("c128_s", b128, q), ("c64_s", b64, q), ("c64_fused5_s", b64, qf)])
    qa = models.QFT(n_acc)
    res["qft"] = {
        "gates": q.ngates, "fused5_group_sizes": {str(k): v for k, v in sorted(sizes.items())}, **t,
        "speedup_c64_vs_c128": t["c128_s"] / t["c64_s"],
        "speedup_c64_fused5_vs_c128": t["c128_s"] / t["c64_fused5_s"],
        "accuracy_c64": _accuracy(b64, b128, qa, qa),
        "accuracy_c64_fused5": _accuracy(b64, b128, qa.fuse(max_qubits=5), qa),
    }
    del q, qf

    blk = _blocks5_timing(b64, b128, seed, n_time, nblocks)
    ba = _block_circuit(n_acc, nblocks, seed)
    blk["accuracy_c64"] = _accuracy(b64, b128, ba, ba)
    res["blocks5"] = blk
    del ba

    checks = []
    for name, key, acc_key in (("layered, fuse(max_qubits=5)", "layered", "accuracy_c64_fused5"),
                               ("layered, unfused", "layered", "accuracy_c64"),
                               ("layered, unfused, specialised passes", "layered_specialized", "accuracy_c64"),
                               ("QFT, fuse(max_qubits=5)", "qft", "accuracy_c64_fused5"),
                               ("dense 5-qubit blocks", "blocks5", "accuracy_c64")):
        w = res[key]
        sp = w["speedup_c64_fused5_vs_c128"] if "fused5" in acc_key else w["speedup_c64_vs_c128"]
        inf = (res["layered"] if key == "layered_specialized" else w)[acc_key]["infidelity"]
        checks.append({"workload": name, "speedup_vs_c128": sp, "infidelity": inf,
                       "passed": bool(sp >= 2.0 and abs(inf) <= 1e-6)})
    res["target_c64_2x_at_1e-6"] = checks
    torch.cuda.empty_cache()
    return res


def bench_qft(seed, n=30):
    """qibo's QFT (models.QFT) at 30 qubits: H, controlled phases and the final
    SWAPs, complex128 and complex64 (the circuit of qibo's own benchmarks)."""
    import torch

    from qibo_amd import models
    from qibo_amd.backend import MI355XBackend

    res = {"nqubits": n}
    c = models.QFT(n)
    res["gates"] = c.ngates
    for dtype in ("complex128", "complex64"):
        backend = MI355XBackend(dtype=dtype)
        backend.specialize = False  # one timed run: nothing is repeated
        c._final_state = None
        backend.execute_circuit(c)
        c._final_state = None  # the timed run reuses the cached state block
        r, dt = _timed(lambda: backend.execute_circuit(c))
        del r
        c._final_state = None
        torch.cuda.empty_cache()
        res[f"{dtype}_s"] = dt
        res[f"{dtype}_gate_apps_per_s"] = c.ngates / dt
    return res


def bench_qft15(seed, reps=50):
    """qibo's QFT at 15 qubits, complex128, from |0...0> (the reference's
    CPU-runnable benchmark circuit): launch-bound, reported per circuit."""
    import torch

    from qibo_amd import models
    from qibo_amd.backend import MI355XBackend

    backend = MI355XBackend(dtype="complex128")
    backend.specialize_after = 1  # compiled during the warm-up execution
    c = models.QFT(15)
    c._final_state = None
    backend.execute_circuit(c)
    c._final_state = None
    backend.wait_specializations()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        r = backend.execute_circuit(c)
        c._final_state = None
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    del r
    # the same executions with the program cache off: every circuit is planned
    # and compiled again on the host
    backend.program_cache = False
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        r = backend.execute_circuit(c)
        c._final_state = None
    torch.cuda.synchronize()
    dt_nc = (time.perf_counter() - t0) / reps
    backend.program_cache = True
    del r
    return {"nqubits": 15, "gates": c.ngates, "us_per_circuit": dt * 1e6, "gate_apps_per_s": c.ngates / dt,
            "circuits_timed": reps, "program_cache": "replay of the compiled passes of the identical circuit",
            "us_per_circuit_compiling_every_circuit": dt_nc * 1e6}


def bench_layered28(seed, nshots=100000, steps=3):
    """The layered circuit at 28 qubits, complex128, 20 layers, M(all) with
    100 000 shots: the same step as the headline at the register size of the
    reference's layered-circuit configuration."""
    import torch

    from qibo_amd import models
    from qibo_amd.backend import MI355XBackend

    backend = MI355XBackend(dtype="complex128")
    backend.specialize_after = 1  # compiled during the warm-up step
    c = models.random_layered_circuit(28, 20, seed=seed)

    def step():
        c._final_state = None
        for gate in c.measurements:
            gate.result.reset()
        res = backend.execute_circuit(c, nshots=nshots)
        return res.samples()

    step()
    backend.wait_specializations()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        s = step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    del s
    c._final_state = None
    torch.cuda.empty_cache()
    ngates = sum(1 for g in c.queue if g.__class__.__name__ != "M")
    return {"nqubits": 28, "gates": ngates, "nshots": nshots, "ms_per_step": dt * 1e3,
            "gate_apps_per_s": ngates / dt}


def bench_serving(seed, nrequests=48, qps_list=(8.0, 16.0, 32.0, 64.0), ncheck=8):
    """The mixed stream through the continuous-batching server at increasing
    offered load (Poisson arrivals): circuits/s, p50 and p99 latency per
    offered QPS, up to saturation. A sample of each stream's requests is then
    served again alone and its results compared (bitwise) with the stream's."""
    import numpy as np
    import torch

    from qibo_amd.backend import MI355XBackend
    from qibo_amd.serving import CircuitServer, GPUExecutor, poisson_stream

    backends = {"complex128": MI355XBackend(dtype="complex128"), "complex64": MI355XBackend(dtype="complex64")}
    for b in backends.values():
        b.specialize = False  # requests of many structures: no background compiles while serving
    ex = GPUExecutor(backends, nstreams=8)
    budget = int(0.8 * torch.cuda.mem_get_info()[0])
    sweep = []
    mismatches = checked = 0
    for i, qps in enumerate(qps_list):
        # the same requests at every level (the seed fixes the circuits; the
        # offered rate only scales the arrival times)
        reqs = poisson_stream(nrequests, qps=qps, seed=seed)
        ex.timings.clear()
        stats = CircuitServer(ex, budget, max_inflight=8).run(reqs, realtime=True)
        row = stats.summary()
        row["qps_offered"] = qps
        tm = np.asarray([(h, d) for _, h, d in ex.timings]) if ex.timings else np.zeros((1, 2))
        # per request: host time from start to the last launch, then the wait
        # for the device
        row["host_ms_p50"] = float(np.percentile(tm[:, 0], 50) * 1e3)
        row["host_ms_max"] = float(tm[:, 0].max() * 1e3)
        row["device_wait_ms_p50"] = float(np.percentile(tm[:, 1], 50) * 1e3)
        sweep.append(row)
        rng = np.random.default_rng(seed + i)
        for r in rng.choice(reqs, size=min(ncheck, len(reqs)), replace=False):
            mixed = r.result
            h = ex.start(r)
            while not ex.done(h):
                pass
            alone = ex.result(h)
            checked += 1
            same = (mixed == alone) if r.kind == "expval" else np.array_equal(mixed, alone)
            mismatches += 0 if same else 1
        torch.cuda.empty_cache()
    ex.close()
    best = max(sweep, key=lambda r: r["circuits_per_s"])
    return {"requests_per_level": nrequests, "nqubits_range": [10, 30], "depth_range": [10, 50],
            "hbm_budget_bytes": budget, "sweep": sweep, "circuits_per_s_max": best["circuits_per_s"],
            "p50_ms_at_max": best["p50_ms"], "p99_ms_at_max": best["p99_ms"],
            "executor_paths": dict(ex.stats),
            "alone_vs_stream_checked": checked, "alone_vs_stream_mismatches": mismatches}


def _noisy_layered(n, nlayers, seed, lam=0.01):
    """The layered RY/RZ/CZ circuit on a density matrix with a one-qubit
    DepolarizingChannel on every qubit after every layer."""
    from qibo_amd import Circuit, gates, models

    angles = models.layered_angles(n, nlayers, seed)
    c = Circuit(n, density_matrix=True)
    for layer in range(nlayers):
        for q in range(n):
            c.add(gates.RY(q, float(angles[layer, 0, q])))
        for q in range(n):
            c.add(gates.RZ(q, float(angles[layer, 1, q])))
        for q in range(layer % 2, n - 1, 2):
            c.add(gates.CZ(q, q + 1))
        for q in range(n):
            c.add(gates.DepolarizingChannel((q,), lam))
    return c


def bench_density(seed, n=14, nlayers=6, n_cpu=9, reps=3):
    """Noisy layered circuit on an n-qubit density matrix (a 2n-qubit vector:
    4 GiB at n = 14 in complex128). Unitary gates run as blocked passes of
    conj pairs, each DepolarizingChannel as one qsim_dm_depolarize pass.
    The oracle (numpy restatement of the reference's density path) is timed on
    the same circuit at n_cpu qubits, with the GPU beside it at that size."""
    import numpy as np
    import torch

    from oracle import OracleDensity
    from qibo_amd.backend import MI355XBackend

    backend = MI355XBackend(dtype="complex128")
    backend.specialize = False  # one timed run per size: nothing is repeated
    out = {"nqubits": n, "layers": nlayers}
    for size, key in ((n, "gpu"), (n_cpu, "gpu_small")):
        c = _noisy_layered(size, nlayers, seed)
        ngates = sum(1 for g in c.queue if not hasattr(g, "coefficients"))
        nch = len(c.queue) - ngates
        backend.execute_circuit(c)  # warm-up
        best = None
        for _ in range(reps):
            c._final_state = None
            _, dt = _timed(lambda: backend.execute_circuit(c))
            best = dt if best is None else min(best, dt)
        out[key] = {"nqubits": size, "gates": ngates, "channels": nch, "s": best,
                    "operations_per_s": (ngates + nch) / best}
        del c
        torch.cuda.empty_cache()
    # one depolarizing pass over the 4^n elements: read + write of the matrix
    rho = backend.zero_state(n, density_matrix=True)
    from qibo_amd import gates as g_

    ch = g_.DepolarizingChannel((n // 2,), 0.01)
    stream = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    backend.depolarizing_error_density_matrix(ch, rho, n)
    ev[0].record(stream)
    for _ in range(10):
        backend.depolarizing_error_density_matrix(ch, rho, n)
    ev[1].record(stream)
    torch.cuda.synchronize()
    ms = ev[0].elapsed_time(ev[1]) / 10
    nbytes = 2 * (4**n) * 16
    out["depolarize_pass"] = {"ms": ms, "bytes": nbytes, "GBs": nbytes / (ms * 1e-3) / 1e9}
    del rho
    torch.cuda.empty_cache()
    # the oracle on the small size, same circuit
    c = _noisy_layered(n_cpu, nlayers, seed)
    t0 = time.perf_counter()
    want = OracleDensity().execute_dm(c)
    cpu_s = time.perf_counter() - t0
    got = backend.execute_circuit(c).state(numpy=True)
    out["cpu_oracle_small"] = {"nqubits": n_cpu, "s": cpu_s, "cores": 1,
                               "max_abs_err_gpu_vs_oracle": float(np.max(np.abs(got - want)))}
    out["gpu_small_speedup_vs_cpu"] = cpu_s / out["gpu_small"]["s"]
    return out


def bench_adjoint(seed, n=20, nlayers=10, nsets=64):
    import numpy as np
    import torch

    from qibo_amd import hamiltonians, models
    from qibo_amd.adjoint import batched_expectation_and_gradient
    from qibo_amd.backend import MI355XBackend

    rng = np.random.default_rng(seed)
    params = rng.uniform(0, 2 * np.pi, (nsets, 2 * n * nlayers))
    out = {"nqubits": n, "layers": nlayers, "parameters": 2 * n * nlayers, "parameter_sets": nsets}
    for dtype in ("complex128", "complex64"):
        backend = MI355XBackend(dtype=dtype)
        # the sweeps' passes run the kernels specialised to their records,
        # compiled during the one-set warm-up (the same records as 64 sets),
        # where they fit the compile budget (measured here, off by default)
        backend.specialize_after, backend.specialize_sweeps = 1, True
        c = models.hardware_efficient_ansatz(n, nlayers)
        ham = hamiltonians.TFIM(n, h=1.0, closed_boundary=False, backend=backend)
        batched_expectation_and_gradient(backend, c, ham, params[:1])  # warm-up
        t0 = time.perf_counter()
        backend.wait_specializations()
        out[f"{dtype}_compile_wait_s"] = time.perf_counter() - t0
        launches0 = backend.specialized_launches
        (e2, g2), dts = _timed(lambda: batched_expectation_and_gradient(backend, c, ham, params))
        out[f"{dtype}_specialized_launches"] = backend.specialized_launches - launches0
        out[f"{dtype}_gradients_per_s_specialized"] = nsets / dts
        # the default path: sweeps interpreted (specialize_sweeps off)
        backend.specialize_sweeps = False
        (e, g), dt = _timed(lambda: batched_expectation_and_gradient(backend, c, ham, params))
        out[f"{dtype}_s"] = dt
        out[f"{dtype}_gradients_per_s"] = nsets / dt
        out[f"{dtype}_specialized_equals_interpreted"] = bool(torch.equal(e, e2) and torch.equal(g, g2))
        del e2, g2
        if dtype == "complex128":
            # central differences (eps = 1e-4) on four parameters of set 0
            c.set_parameters(params[0])
            errs = []
            for k in rng.choice(params.shape[1], 4, replace=False):
                p = params[0].copy()
                p[k] += 1e-4
                c.set_parameters(p)
                fp = ham.expectation(c)
                p[k] -= 2e-4
                c.set_parameters(p)
                fm = ham.expectation(c)
                errs.append(abs(float(g[0, k]) - (fp - fm) / 2e-4))
            out["max_abs_err_vs_central_fd"] = max(errs)
        del e, g
        torch.cuda.empty_cache()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nqubits", type=int, default=30)
    ap.add_argument("--nlayers", type=int, default=20)
    ap.add_argument("--nshots", type=int, default=100000)
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--extras", default=None, help="comma-separated subset of the extras to run")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)

    import numpy as np
    import torch

    from qibo_amd import models
    from qibo_amd.backend import MI355XBackend

    backend = MI355XBackend(dtype="complex128", device=0)
    # compile the passes' specialised kernels at their first launch (in the
    # warm-up), not at the second
    backend.specialize_after = 1
    n = args.nqubits
    circuit = models.random_layered_circuit(n, args.nlayers, seed=args.seed)
    ngates = sum(1 for g in circuit.queue if g.__class__.__name__ != "M")
    itemsize = 16


    def step():
        circuit._final_state = None
        # a non-collapsing M keeps the samples of the previous run on the gate
        # (qibo's MeasurementResult); clear them so every step samples afresh
        for gate in circuit.measurements:
            gate.result.reset()
        result = backend.execute_circuit(circuit, nshots=args.nshots)
        samples = result.samples()
        return result, samples

    # copy calibration: a plain device-to-device copy of one state-sized buffer
    # (torch's copy kernel), the reference point for the HBM counter scaling
    src = torch.empty(2**n, dtype=torch.complex128, device=backend.torch_device)
    dst = torch.empty_like(src)
    src.zero_()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    copy_ms = []
    for _ in range(3):
        ev0.record()
        dst.copy_(src)
        ev1.record()
        ev1.synchronize()
        copy_ms.append(ev0.elapsed_time(ev1))
    copy_bytes = 2 * src.numel() * itemsize
    del src, dst
    torch.cuda.empty_cache()

    for _ in range(args.warmup):
        r, s = step()
        del r, s
    torch.cuda.synchronize()
    # the kernels specialised to the step's record sequences were compiled in
    # the background during the warm-up (hiprtc on the host); wait for them
    t_spec = time.perf_counter()
    nspec = backend.wait_specializations() if backend.specialize else 0
    spec_wait_s = time.perf_counter() - t_spec
    spec0 = backend.specialized_launches

    # events around every blocked-pass launch on the stream it runs on
    backend.block_stats.clear()
    backend.block_bytes.clear()
    hits0 = backend.program_cache_hits
    timer = backend.enable_kernel_timer("block")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r, s = step()
        del r, s
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    kernel_ms = backend.disable_kernel_timer()
    passes = list(backend.block_stats)
    launch_bytes = list(backend.block_bytes)  # (algorithmic bytes, over every tile) per launch
    cache_hits = backend.program_cache_hits - hits0
    spec_launches = backend.specialized_launches - spec0

    # the same steps with the compiled-program cache off: every step plans and
    # compiles its passes on the host again (overlapped with the GPU after the
    # first pass)
    backend.program_cache = False
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r, s = step()
        del r, s
    torch.cuda.synchronize()
    elapsed_nocache = time.perf_counter() - t0
    backend.program_cache = True

    # the same steps with every pass interpreted by k_blocked
    spec_on = backend.specialize
    backend.specialize = False
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r, s = step()
        del r, s
    torch.cuda.synchronize()
    elapsed_interp = time.perf_counter() - t0
    backend.specialize = spec_on

    ms_per_step = 1000.0 * elapsed / args.steps
    # HBM bytes the step's kernels move: each blocked pass reads and writes the
    # state once (the first ones, from |0...0>, read only the support built so
    # far and write the tiles that contain it); the probabilities
    # read the state and write 8 B per amplitude; the sampling maps read the
    # probabilities again and write 1 B per amplitude (16 B per 16 entries)
    npass = len(passes) / args.steps if passes else 0.0
    # the passes' bytes as launched: the first (from |0...0>) runs tile 0 and
    # the next ones read only the support built so far and run only the tiles
    # that contain it (the state outside is zero); see backend.block_bytes
    pass_bytes = sum(b for b, _ in launch_bytes) / args.steps if launch_bytes else 0.0
    step_bytes = pass_bytes + (2**n) * (itemsize + 8) + (2**n) * 9
    # HBM bytes per k_blocked launch, from the committed rocprofv3 counter runs of
    # this same workload (FETCH_SIZE x 2 + WRITE_SIZE x 1, the scales read off a
    # plain copy of one state buffer); only quoted for the workload it measured
    # FP64 work of the timed passes (the same compile the backend performs)
    from qibo_amd import blocking

    ops = [backend.plan_gate(g) for g in circuit.queue if g.__class__.__name__ != "M"]
    plan, program = blocking.compile_circuit(ops, n, np.complex128, 3)
    pass_flops = [blocking.program_flops(program[e[2]:], e[3], 16) for e in plan if e[0] == "block"]
    step_dense = [d for _, d in launch_bytes[: len(pass_flops)]]
    if len(step_dense) == len(pass_flops) and any(step_dense):  # the passes over every tile
        pass_flops = [f for f, d in zip(pass_flops, step_dense) if d]
    flops_per_launch = float(np.mean(pass_flops)) * 2**n if pass_flops else 0.0
    traffic, traffic_src = None, None
    # (the specialised passes' counters when the timed steps ran them)
    prof_name = "qsim_blocked_spec_layered30_c128.json" if spec_launches else "k_blocked_layered30_c128.json"
    prof = os.path.join(ROOT, "profiles", prof_name)
    if n == 30 and args.nlayers == 20 and os.path.exists(prof):
        with open(prof) as fh:
            traffic = json.load(fh)["traffic_bytes"]
        traffic_src = "profiles/" + prof_name
    value = ngates * args.steps / elapsed
    # the roofline of the pass over every tile (the state read and written
    # once): the launches that ran on part of the tiles are left out
    dense_ms = [ms for ms, (_, dense) in zip(kernel_ms, launch_bytes) if dense] if len(kernel_ms) == len(
        launch_bytes) else list(kernel_ms)
    avg_dense_ms = float(np.mean(dense_ms)) if dense_ms else float("nan")
    dense_bytes = 2 * (2**n) * itemsize
    achieved = dense_bytes / (avg_dense_ms * 1e-3) / 1e9

    out = {
        "metric": "gate-apps/s + HBM GB/s at 30q c128; fused c64 speedup; served circuits/s, p99",
        "value": value,
        "unit": "gate-apps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "complex128",
        "data": "synthetic (seeded layered circuit angles)",
        "config": {
            "workload": f"random layered circuit RY/RZ/CZ-brick, {args.nlayers} layers, M(all), "
                        f"{args.nshots} shots",
            "nqubits": n,
            "gates_per_step": ngates,
            "nshots": args.nshots,
            "parallelism": "single GPU",
        },
        "hbm_GBs_step": {
            "bytes_per_step": step_bytes,
            "GBs": step_bytes / (elapsed / args.steps) / 1e9,
            "counted": "per blocked pass one read and one write of the tiles it runs (the first "
                       "ones, from |0...0>, read only the support built so far and write the tiles "
                       "that contain it), probabilities (16 B read + 8 B written per amplitude), "
                       "sampling maps (8 B read + 1 B written per amplitude)",
        },
        # the backend keeps the compiled passes of a repeated identical gate
        # queue (same gates, qubits and parameter values); every timed step
        # still runs all passes, probabilities and sampling on the GPU
        "host_program_cache": {
            "timed_steps": "replay of the passes compiled at the first execution of the identical circuit",
            "hits_in_timed_steps": cache_hits,
            "ms_per_step_compiling_every_step": 1000.0 * elapsed_nocache / args.steps,
            "value_compiling_every_step": ngates * args.steps / elapsed_nocache,
        },
        # every pass runs the kernel specialised to its record sequence
        # (qsim_blocked_spec, compiled by hiprtc during the warm-up): the same
        # operations as k_blocked's interpreter, bitwise the same amplitudes
        "specialized_passes": {
            "kernels": nspec,
            "compile_wait_after_warmup_s": spec_wait_s,
            "specialized_launches_in_timed_steps": spec_launches,
            "launches_in_timed_steps": len(passes),
            "ms_per_step_interpreted": 1000.0 * elapsed_interp / args.steps,
            "value_interpreted": ngates * args.steps / elapsed_interp,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": ("qsim_blocked_spec (blocked pass specialised to its records: state read + written once)"
                       if spec_launches else "k_blocked<double> (blocked pass: state read + written once)"),
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_source": traffic_src,
            "launches_timed": len(dense_ms),
            "launches_on_part_of_the_tiles": len(kernel_ms) - len(dense_ms),
            "avg_launch_ms": avg_dense_ms,
            "gates_per_launch": (sum(passes) / len(passes)) if passes else None,
            "passes_per_step": len(passes) / args.steps,
            # the pass is bound by FP64 VALU work, not HBM: its FP64 rate against
            # the MI355X FP64 peak (78.6 TFLOP/s counts every operation as an FMA)
            "compute": {
                "flops_per_launch": flops_per_launch,
                "achieved_TFLOPs": flops_per_launch / (avg_dense_ms * 1e-3) / 1e12,
                "peak_TFLOPs_fp64": 78.6,
                "frac": flops_per_launch / (avg_dense_ms * 1e-3) / 78.6e12,
            },
        },
        "copy_calibration": {
            "bytes": copy_bytes,
            "best_ms": min(copy_ms),
            "GBs": copy_bytes / (min(copy_ms) * 1e-3) / 1e9,
        },
    }
    if not args.no_extras:
        chosen = None if args.extras is None else set(args.extras.split(","))
        for key, fn in (("layered28", bench_layered28), ("qft", bench_qft), ("qft15", bench_qft15),
                        ("precision", bench_precision), ("blocks5", bench_blocks5), ("serving", bench_serving), ("adjoint", bench_adjoint),
                        ("density", bench_density)):
            if (chosen is not None and key not in chosen) or (chosen is None and key == "blocks5"):
                continue  # blocks5 is part of "precision"; alone only on request (profile runs)
            try:
                out[key] = fn(args.seed)
            except Exception as exc:  # noqa: BLE001 - an extra must not hide the headline line
                out[key] = {"error": f"{type(exc).__name__}: {exc}"}
                print(f"bench: {key} failed: {exc}", file=sys.stderr, flush=True)
            torch.cuda.empty_cache()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(n, args.nlayers, args.seed, args.cpu_budget)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
