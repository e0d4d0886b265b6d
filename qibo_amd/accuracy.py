"""Accuracy reports: infidelity and max-abs amplitude error of a result.

qibo defines the fidelity of two pure states as |conj(state) @ target|^2 and
the infidelity as 1 - fidelity (src/qibo/quantum_info/metrics.py:200-291).
A complex64 result is reported against the complex128 execution of the same
circuit from the same initial state on the same device; that path matches
qibo's NumpyBackend to 1e-12 in max-abs amplitude. Both states are read once
by one HIP kernel (``qsim_compare_states``) that accumulates the overlap, the
two norms and the largest amplitude difference in FP64, so no widened copy of
a complex64 state is made.

The report holds:

* ``infidelity``: 1 - |<state|reference>|^2, qibo's formula as written (a
  complex64 norm that drifts by d moves it by about -d);
* ``infidelity_normalized``: the same for the normalised states, i.e. the
  error of the direction of the state alone;
* ``max_abs_error``: max_i |state_i - reference_i|;
* ``norm_error``: <state|state> - 1;
* ``reference``: what the state was compared with.
"""

import math

import torch

from . import _lib
from .config import raise_error


def compare_states(backend, state: torch.Tensor, target: torch.Tensor, reference: str = "target") -> dict:
    """Accuracy of the device state ``state`` against ``target`` (same length,
    complex64 or complex128 each) on the GPU."""
    if state.dim() != 1 or target.dim() != 1 or state.numel() != target.numel():
        raise_error(ValueError, "compare_states takes two state vectors of the same length")
    if state.device != target.device:
        raise_error(ValueError, "compare_states: the states are on different devices")
    n = int(state.numel()).bit_length() - 1
    if 1 << n != state.numel():
        raise_error(ValueError, "compare_states: the length is not a power of two")
    lib = backend.lib
    codes = []
    for t in (state, target):
        if t.dtype == torch.complex128:
            codes.append(_lib.C128)
        elif t.dtype == torch.complex64:
            codes.append(_lib.C64)
        else:
            raise_error(TypeError, f"compare_states: unsupported dtype {t.dtype}")
    state = state.contiguous()
    target = target.contiguous()
    out = torch.empty(5, dtype=torch.float64, device=state.device)
    nbytes = lib.qsim_compare_workspace(n)
    ws = backend._workspace(nbytes, "compare", device=state.device)
    rc = lib.qsim_compare_states(state.data_ptr(), codes[0], target.data_ptr(), codes[1], n, out.data_ptr(),
                                 ws.data_ptr(), nbytes, backend._stream(state))
    _lib.check(rc, lib)
    re, im, naa, nbb, mx = out.cpu().tolist()
    ov2 = re * re + im * im
    denom = naa * nbb
    return {
        "infidelity": 1.0 - ov2,
        "infidelity_normalized": (1.0 - ov2 / denom) if denom > 0 else math.nan,
        "max_abs_error": mx,
        "norm_error": naa - 1.0,
        "reference": reference,
    }


def result_accuracy(result) -> dict:
    """Accuracy of a state-vector result against the complex128 execution of
    the circuit that produced it (re-run on the same device). A complex128
    result is the reference precision itself: its report is exact zeros."""
    backend = result.backend
    state = result.state()
    if state.dim() != 1:
        raise_error(NotImplementedError, "accuracy reports cover state vectors, not density matrices")
    if state.dtype == torch.complex128:
        return {"infidelity": 0.0, "infidelity_normalized": 0.0, "max_abs_error": 0.0,
                "norm_error": None, "reference": "self (complex128)"}
    source = getattr(result, "_source", None)
    if source is None:
        raise_error(ValueError, "this result does not know the circuit that produced it")
    circuit, initial_state = source
    ref_backend = backend.reference_backend()
    saved_final = getattr(circuit, "_final_state", None)
    saved_backends = [(g, g.result.backend) for g in circuit.measurements]
    try:
        ref = ref_backend.execute_circuit(circuit, initial_state=initial_state, nshots=1)
        ref_state = ref.state()
        report = compare_states(backend, state, ref_state, reference="complex128 execution of the same circuit")
    finally:
        circuit._final_state = saved_final
        for gate, b in saved_backends:
            gate.result.backend = b
    return report
