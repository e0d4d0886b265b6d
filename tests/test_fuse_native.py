"""The native fusion planner (qsim_fuse_plan, behind Circuit.fuse) and the
native fused-matrix product (qsim_matrix_fused, behind matrix_fused) against
their Python restatements (Circuit._fuse_python, step for step the
reference's planner; GatePlanner._matrix_fused_python, the reference's
kron / transpose / multiply construction): the same groups, gate for gate
and in the same order, on layered, QFT, hardware-efficient and random
circuits with measurements, callbacks, controlled_by gates and three-qubit
gates, for max_qubits 2..5; matrices within 1e-13 (complex128) and 1e-5
(complex64). Host code only: no GPU."""

import numpy as np
import pytest

from qibo_amd import Circuit, gates, models
from qibo_amd.backend import GatePlanner


def _random_circuit(n, ngates, rng, specials=True):
    c = Circuit(n)
    for i in range(ngates):
        k = int(rng.integers(0, 12))
        q = [int(x) for x in rng.choice(n, 4, replace=False)]
        th = float(rng.uniform(-6, 6))
        if k == 0:
            c.add(gates.RY(q[0], th))
        elif k == 1:
            c.add(gates.RZ(q[0], th))
        elif k == 2:
            c.add(gates.CZ(q[0], q[1]))
        elif k == 3:
            c.add(gates.CNOT(q[0], q[1]))
        elif k == 4:
            c.add(gates.TOFFOLI(q[0], q[1], q[2]))
        elif k == 5:
            c.add(gates.H(q[0]))
        elif k == 6:
            c.add(gates.CU1(q[0], q[1], th))
        elif k == 7:
            c.add(gates.SWAP(q[0], q[1]))
        elif k == 8:
            c.add(gates.RX(q[0], th).controlled_by(q[1], q[2]))
        elif k == 9:
            c.add(gates.U3(q[0], th, 0.3, -0.7))
        elif k == 10 and specials and i % 7 == 0:
            c.add(gates.M(q[0]))
        else:
            c.add(gates.fSim(q[0], q[1], th, 0.4))
    return c


def _circuits():
    rng = np.random.default_rng(5)
    yield models.random_layered_circuit(12, 6, seed=3, measure=True)
    yield models.QFT(9)
    yield models.hardware_efficient_ansatz(8, 3, params=rng.uniform(0, 6, 48))
    for n in (4, 6, 9, 13):
        yield _random_circuit(n, 120, rng)


def _signature(queue):
    out = []
    for g in queue:
        if isinstance(g, gates.FusedGate):
            out.append(("F", tuple(g.target_qubits), tuple(id(h) for h in g.gates)))
        else:
            out.append(("G", id(g)))
    return out


@pytest.mark.parametrize("maxq", [2, 3, 4, 5])
def test_native_fuse_equals_python(maxq):
    ngroups = 0
    for c in _circuits():
        got = c.fuse(max_qubits=maxq).queue
        want = c._fuse_python(max_qubits=maxq).queue
        assert _signature(got) == _signature(want)
        ngroups += sum(1 for g in got if isinstance(g, gates.FusedGate))
    assert ngroups > 20


def test_native_fuse_with_callback():
    from qibo_amd.special import CallbackGate

    class Cb:
        def apply(self, backend, state):
            pass

    c = Circuit(4)
    c.add(gates.H(0))
    c.add(gates.CNOT(0, 1))
    c.add(CallbackGate(Cb()))
    c.add(gates.H(1))
    c.add(gates.CZ(1, 2))
    assert _signature(c.fuse(3).queue) == _signature(c._fuse_python(3).queue)


@pytest.mark.parametrize("dtype,tol", [("complex128", 1e-13), ("complex64", 1e-5)])
def test_native_matrix_fused_equals_python(dtype, tol):
    gp = GatePlanner(dtype)
    count = 0
    for c in _circuits():
        for g in c.fuse(max_qubits=5).queue:
            if not isinstance(g, gates.FusedGate):
                continue
            a = gp.matrix_fused(g)
            b = gp._matrix_fused_python(g)
            assert a.dtype == b.dtype and a.shape == b.shape
            assert np.max(np.abs(a - b)) <= tol
            count += 1
    assert count > 20
