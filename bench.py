"""Benchmark of the MI355X state-vector backend.

Workload (one "step"): the seeded random layered circuit -- per layer RY on
every qubit, RZ on every qubit, CZ on the brick pairs -- on 30 qubits in
complex128, 20 layers (600 RY + 600 RZ + 290 CZ = 1490 gate applications),
followed by M(all) with 100 000 shots: the final state, its probabilities
and the sampled bitstrings, through the drop-in path ``circuit(nshots=...)``
-> ``CircuitResult.samples()``. Inputs are synthetic (seeded angles); the
state stays resident in HBM for the whole step.

Reported: gate applications per second (``value``), the algorithmic HBM
bandwidth of the whole step, the roofline of the dominant kernel (the dense
one-qubit pass, 2 * 2^n * 16 bytes per launch) timed with HIP events on its
stream inside the timed region, and a CPU baseline: the numpy oracle of
qibo's NumpyBackend timed on a bounded sample of the same circuit's gates on
this host.

    python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--nqubits 30] [--no-cpu]
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published


def gate_bytes(kind, ntargets, ncontrols, nqubits, itemsize):
    """Algorithmic bytes of one planned gate: each changed amplitude read + written once."""
    s = (2**nqubits) * itemsize
    if kind == "noop":
        return 0
    if kind == "swap":
        return 2 * s / 2 / (2**ncontrols)
    return 2 * s / (2**ncontrols)


def cpu_baseline(nqubits, nlayers, seed, budget_s):
    """Oracle (numpy NumpyBackend restatement) on a stratified gate sample, one process."""
    import numpy as np

    from oracle import OracleBackend
    from qibo_amd import models

    try:
        from threadpoolctl import threadpool_info

        blas_threads = max((i.get("num_threads", 1) for i in threadpool_info()), default=1)
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
        "sample": f"{done} gates of layer 0 at n={nqubits} complex128 ({', '.join(names)}); "
                  f"{elapsed:.1f} s; numpy {np.__version__} matmul with {blas_threads} BLAS threads, "
                  "transposes single-threaded",
    }


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
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)

    import numpy as np
    import torch

    from qibo_amd import models
    from qibo_amd.backend import MI355XBackend

    backend = MI355XBackend(dtype="complex128", device=0)
    n = args.nqubits
    circuit = models.random_layered_circuit(n, args.nlayers, seed=args.seed)
    ngates = sum(1 for g in circuit.queue if g.__class__.__name__ != "M")
    itemsize = 16

    # algorithmic bytes of one step
    step_bytes = 0.0
    for g in circuit.queue:
        if g.__class__.__name__ == "M":
            continue
        op = backend.plan_gate(g)
        step_bytes += gate_bytes(op.kind, len(op.targets), len(op.controls), n, itemsize)
    probs_bytes = (2**n) * itemsize + (2**n) * 8
    cumsum_bytes = 2 * (2**n) * 8
    step_bytes += probs_bytes + cumsum_bytes

    def step():
        circuit._final_state = None
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

    # events around every dense one-qubit launch on the stream it runs on
    timer = backend.enable_kernel_timer("dense1")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r, s = step()
        del r, s
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    kernel_ms = backend.disable_kernel_timer()

    ms_per_step = 1000.0 * elapsed / args.steps
    value = ngates * args.steps / elapsed
    avg_dense_ms = float(np.mean(kernel_ms)) if kernel_ms else float("nan")
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
        "hbm_gbs_step": step_bytes / (elapsed / args.steps) / 1e9,
        "roofline": {
            "bound": "hbm",
            "kernel": "k_dense<double,1> (RY pass)",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "launches_timed": len(kernel_ms),
            "avg_launch_ms": avg_dense_ms,
        },
        "copy_calibration": {
            "bytes": copy_bytes,
            "best_ms": min(copy_ms),
            "GBs": copy_bytes / (min(copy_ms) * 1e-3) / 1e9,
        },
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(n, args.nlayers, args.seed, args.cpu_budget)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
