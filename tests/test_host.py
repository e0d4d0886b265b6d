"""Host-side logic of the backend, checked on the CPU against the oracle.

* gate matrices of qibo_amd.matrices are bitwise those of the oracle's
  NumpyMatrices restatement (both dtypes);
* the kernel plan chosen for a gate (controls peeled off, diagonal, swap,
  dense) reproduces the dense product when emulated in numpy;
* fusion grouping follows the structural goldens of
  tests/test_models_circuit_fuse.py:24-54 of the reference;
* the fused-matrix construction equals the oracle's sparse one;
* circuit bookkeeping (registers, parameters, inversion) behaves like qibo's.
"""

import numpy as np
import pytest

from oracle import OracleBackend
from qibo_amd import Circuit, gates, hamiltonians, models
from qibo_amd.backend import GatePlanner, plan_matrix

GATE_FACTORIES = [
    lambda: gates.H(0), lambda: gates.X(1), lambda: gates.Y(2), lambda: gates.Z(0),
    lambda: gates.S(1), lambda: gates.SDG(1), lambda: gates.T(2), lambda: gates.TDG(2),
    lambda: gates.SX(0), lambda: gates.SXDG(0), lambda: gates.I(1),
    lambda: gates.RX(0, 0.123), lambda: gates.RY(1, 2.5), lambda: gates.RZ(2, -1.7),
    lambda: gates.U1(0, 0.3), lambda: gates.U2(1, 0.4, -0.9), lambda: gates.U3(2, 0.1, 0.2, 0.3),
    lambda: gates.CNOT(0, 2), lambda: gates.CY(1, 0), lambda: gates.CZ(2, 1), lambda: gates.CH(0, 1),
    lambda: gates.CRX(0, 1, 0.7), lambda: gates.CRY(1, 2, -0.7), lambda: gates.CRZ(2, 0, 1.9),
    lambda: gates.CU1(0, 2, 0.25), lambda: gates.SWAP(0, 2), lambda: gates.iSWAP(1, 2),
    lambda: gates.FSWAP(0, 1), lambda: gates.fSim(0, 2, 0.4, 1.3), lambda: gates.RXX(0, 1, 0.3),
    lambda: gates.RZZ(1, 2, 0.9), lambda: gates.TOFFOLI(0, 1, 2), lambda: gates.CCZ(2, 0, 1),
]


@pytest.mark.parametrize("dtype", ["complex128", "complex64"])
@pytest.mark.parametrize("make", GATE_FACTORIES)
def test_matrices_bitwise_equal_oracle(make, dtype):
    g = make()
    mine = GatePlanner(dtype).matrix(g)
    ref = OracleBackend(dtype).matrix(g)
    assert mine.dtype == ref.dtype
    np.testing.assert_array_equal(mine, ref)


def emulate(op, state, n):
    """numpy emulation of the kernel semantics of a planned op."""
    psi = state.copy()
    if op.kind == "noop":
        return psi
    idx = np.arange(2**n)
    cmask = 0
    for c in op.controls:
        cmask |= 1 << (n - 1 - c)
    active = (idx & cmask) == cmask
    if op.kind == "diag":
        r = np.zeros_like(idx)
        for t in op.targets:
            r = (r << 1) | ((idx >> (n - 1 - t)) & 1)
        out = psi.copy()
        out[active] = psi[active] * op.data[r[active]]
        return out
    if op.kind == "swap":
        a, b = op.targets
        ma, mb = 1 << (n - 1 - a), 1 << (n - 1 - b)
        sel = active & ((idx & ma) != 0) & ((idx & mb) == 0)
        out = psi.copy()
        i1 = idx[sel]
        i2 = (i1 & ~ma) | mb
        out[i1], out[i2] = psi[i2], psi[i1]
        return out
    k = len(op.targets)
    tmask = 0
    for t in op.targets:
        tmask |= 1 << (n - 1 - t)
    out = psi.copy()
    for base in idx[active & ((idx & tmask) == 0)]:
        ids = []
        for r in range(2**k):
            i = base
            for j, t in enumerate(op.targets):
                if (r >> (k - 1 - j)) & 1:
                    i |= 1 << (n - 1 - t)
            ids.append(i)
        out[ids] = op.data @ psi[ids]
    return out


@pytest.mark.parametrize("make", GATE_FACTORIES + [
    lambda: gates.RX(1, 0.8).controlled_by(0, 2),
    lambda: gates.SWAP(1, 2).controlled_by(0),
    lambda: gates.U1(2, 0.4).controlled_by(0, 1),
    lambda: gates.Unitary(np.linalg.qr(np.arange(16).reshape(4, 4) + 1j)[0], 2, 0),
])
def test_plan_reproduces_dense_product(make, rng):
    n = 4
    psi = rng.normal(size=2**n) + 1j * rng.normal(size=2**n)
    g = make()
    op = GatePlanner().plan_gate(g)
    want = OracleBackend().apply_gate(g, psi.copy(), n)
    np.testing.assert_allclose(emulate(op, psi, n), want, atol=1e-14, rtol=0)


def test_plan_kinds():
    p = GatePlanner()
    cz = p.plan_gate(gates.CZ(0, 1))
    assert cz.kind == "diag" and cz.targets == () and set(cz.controls) == {0, 1}
    cnot = p.plan_gate(gates.CNOT(2, 0))
    assert cnot.kind == "dense" and cnot.targets == (0,) and cnot.controls == (2,)
    assert p.plan_gate(gates.SWAP(0, 1)).kind == "swap"
    assert p.plan_gate(gates.RZ(0, 0.3)).kind == "diag"
    u1 = p.plan_gate(gates.U1(1, 0.2))
    assert u1.kind == "diag" and u1.targets == () and u1.controls == (1,)
    assert p.plan_gate(gates.I(0)).kind == "noop"
    tof = p.plan_gate(gates.TOFFOLI(0, 1, 2))
    assert tof.kind == "dense" and tof.targets == (2,) and set(tof.controls) == {0, 1}
    assert plan_matrix(np.eye(16), (0, 1, 2, 3), ()).kind == "noop"


def test_single_fusion_gate():
    queue = [gates.H(0), gates.X(1), gates.CZ(0, 1)]
    c = Circuit(2)
    c.add(queue)
    c = c.fuse()
    assert len(c.queue) == 1
    fg = c.queue[0]
    assert fg.gates[0] == queue[1] and fg.gates[1] == queue[0] and fg.gates[2] == queue[2]


def test_two_fusion_gate():
    queue = [gates.X(0), gates.H(1), gates.RX(2, theta=0.1234).controlled_by(1), gates.H(2),
             gates.Y(1), gates.H(0)]
    c = Circuit(3)
    c.add(queue)
    c = c.fuse()
    assert len(c.queue) == 2
    f1, f2 = c.queue
    assert f2.gates[0] == queue[0] and f2.gates[1] == queue[-1]
    assert f1.gates == [queue[1], queue[2], queue[4], queue[3]]


def test_fusedgate_matrix_calculation():
    queue = [gates.H(0), gates.H(1), gates.CNOT(0, 1), gates.X(0), gates.X(1)]
    c = Circuit(2)
    c.add(queue)
    c = c.fuse()
    assert len(c.queue) == 1
    x = np.array([[0, 1], [1, 0]])
    h = np.array([[1, 1], [1, -1]]) / np.sqrt(2)
    cnot = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]])
    np.testing.assert_allclose(GatePlanner().matrix_fused(c.queue[0]),
                               np.kron(x, x) @ cnot @ np.kron(h, h), atol=1e-15)


@pytest.mark.parametrize("dtype", ["complex128", "complex64"])
def test_matrix_fused_equals_oracle(dtype):
    c = models.random_layered_circuit(6, 4, seed=11, measure=False)
    c.add(gates.RX(3, 0.2).controlled_by(1))
    c.add(gates.TOFFOLI(0, 4, 5))
    tol = 1e-14 if dtype == "complex128" else 2e-6
    for k in (2, 3, 4, 5):
        for g in c.fuse(max_qubits=k).queue:
            if isinstance(g, gates.FusedGate):
                np.testing.assert_allclose(GatePlanner(dtype).matrix_fused(g),
                                           OracleBackend(dtype).matrix_fused(g), atol=tol)


def test_fusion_keeps_measurements():
    c = Circuit(3)
    c.add(gates.X(i) for i in range(3))
    c.add(gates.CNOT(0, 1))
    c.add(gates.H(2))
    c.add(gates.M(0, 1, 2))
    f = c.fuse(max_qubits=2)
    assert f.measurements == c.measurements
    assert isinstance(f.queue[-1], gates.M)


def test_circuit_bookkeeping():
    c = Circuit(3)
    c.add(gates.RX(0, 0.1))
    c.add(gates.fSim(0, 2, 0.3, 0.4))
    c.add(gates.H(1))
    r = c.add(gates.M(0, 2, register_name="a"))
    c.add(gates.M(1))
    assert c.measurements[0].register_name == "a" and c.measurements[1].register_name == "register1"
    assert r is c.measurements[0].result
    c.set_parameters([0.5, (0.6, 0.7)])
    assert c.queue[0].parameters == (0.5,) and c.queue[1].parameters == (0.6, 0.7)
    c.set_parameters([1.0, 2.0, 3.0])
    assert c.get_parameters() == [(1.0,), (2.0, 3.0)]
    with pytest.raises(TypeError):  # a scalar for a two-parameter gate, as in qibo
        c.set_parameters([1.0, 2.0])
    with pytest.raises(ValueError):
        c.set_parameters([1.0, 2.0, 3.0, 4.0])
    with pytest.raises(ValueError):
        Circuit(2).add(gates.H(5))
    with pytest.raises(KeyError):
        c.add(gates.M(1, register_name="a"))
    inv = models.QFT(3).invert()
    assert inv.ngates == models.QFT(3).ngates


def test_invert_is_inverse(rng):
    n = 4
    c = models.random_layered_circuit(n, 2, seed=5, measure=False)
    c.add(gates.CU1(0, 3, 0.3))
    c.add(gates.U3(1, 0.1, 0.2, 0.3))
    psi = rng.normal(size=2**n) + 1j * rng.normal(size=2**n)
    ob = OracleBackend()
    back = ob.execute(c.invert(), ob.execute(c, psi))
    np.testing.assert_allclose(back, psi, atol=1e-12)


def test_layered_generator_counts():
    c = models.random_layered_circuit(28, 20, seed=1)
    names = c.gate_names
    assert names["ry"] == 560 and names["rz"] == 560 and names["cz"] == 270 and names["measure"] == 1
    h = models.hardware_efficient_ansatz(20, 10)
    assert len(h.trainable_gates) == 400 and h.ngates == 590


def test_tfim_terms():
    ham = hamiltonians.TFIM(4, h=1.0, closed_boundary=False)
    coeffs, strings, qubits = ham.simple_terms
    assert sorted(zip(strings, qubits)) == sorted(
        [("ZZ", (0, 1)), ("ZZ", (1, 2)), ("ZZ", (2, 3)), ("X", (0,)), ("X", (1,)), ("X", (2,)), ("X", (3,))])
    assert all(c == -1.0 for c in coeffs)


def test_diagonal_planning_matches_general_peeling():
    """plan_matrix's diagonal path gives the op the general block peeling gives."""
    from qibo_amd.backend import _peel_controls, plan_matrix

    rng = np.random.default_rng(0)
    for k in range(1, 5):
        for _ in range(100):
            cube = np.exp(1j * rng.normal(size=2**k)).reshape((2,) * k)
            for t in range(k):
                if rng.random() < 0.5:
                    idx = [slice(None)] * k
                    idx[t] = 0
                    cube[tuple(idx)] = 1
            d = cube.ravel()
            m = np.diag(d)
            targets = [int(x) for x in rng.choice(20, size=k, replace=False)]
            got = plan_matrix(m, targets, [25])
            red, tg, extra = _peel_controls(m, targets)
            want_d = np.diagonal(red)
            kind = "noop" if np.all(want_d == 1) else "diag"
            assert (got.kind, got.targets, got.controls) == (kind, tuple(tg), tuple([25] + extra))
            if kind == "diag":
                assert np.array_equal(got.data, want_d)


def test_program_cache_keys_follow_gate_changes():
    """The per-gate program-cache key kept on a gate changes whenever its
    parameters or qubits do (setters assign new tuples), and is not kept for
    array parameters that could change in place."""
    import numpy as np

    from qibo_amd import gates, models
    from qibo_amd.backend import MI355XBackend

    be = MI355XBackend.__new__(MI355XBackend)  # host methods only: no device
    c = models.random_layered_circuit(6, 2, seed=1, measure=False)
    before = [be._gate_key(g) for g in c.queue]
    assert [be._gate_key(g) for g in c.queue] == before
    c.set_parameters(np.array(c.get_parameters()).ravel() + 0.1)
    after = [be._gate_key(g) for g in c.queue]
    assert sum(a != b for a, b in zip(before, after)) == sum(1 for g in c.queue if g.__class__.__name__ in ("RY", "RZ"))
    h = gates.H(1)
    k1 = be._gate_key(h)
    h.controlled_by(0, 2)
    assert be._gate_key(h) == ("H", (1,), (0, 2), True, ())
    assert k1 != be._gate_key(h)
    u = gates.Unitary(np.eye(2), 0)
    be._gate_key(u)
    assert "_qkey" not in u.__dict__


def test_program_cache_key_none_for_symbolic_gates():
    """A gate whose angle follows a measurement outcome is never cached, also
    after substitute_symbols turned its parameters into floats: the next run
    substitutes a possibly different outcome (abstract.py:2584-2587)."""
    import numpy as np

    from qibo_amd import gates
    from qibo_amd.backend import MI355XBackend

    be = MI355XBackend.__new__(MI355XBackend)
    m = gates.M(0, collapse=True)
    g = gates.RX(1, theta=np.pi * m.result.symbols[0])
    assert be._gate_key(g) is None
    m.result.register_samples(np.array([[1]]))
    g.substitute_symbols()
    assert isinstance(g.parameters[0], float)
    assert be._gate_key(g) is None


def test_low_bits_resolved_and_checked():
    import pytest

    from qibo_amd import _lib, blocking
    from qibo_amd.backend import MI355XBackend

    be = MI355XBackend.__new__(MI355XBackend)
    be.blocking_low_bits = None
    assert be._low_bits(_lib.C128) == 3 and be._low_bits(_lib.C64) == 4
    for ok in (0, 2, blocking.TILE_BITS - 1):
        be.blocking_low_bits = ok
        assert be._low_bits(_lib.C128) == ok
    for bad in (-1, blocking.TILE_BITS, 2.0, True, "3"):
        be.blocking_low_bits = bad
        with pytest.raises(ValueError):
            be._low_bits(_lib.C128)


def test_standin_gates_planned_by_duck_type():
    """Gates that are not qibo_amd classes but carry qibo's attributes (what
    qibo itself passes after set_backend) are planned like the package's own;
    a measurement is recognised and never asked for its matrix."""
    import numpy as np
    import pytest

    from qibo_amd import gates, kinds
    from qibo_amd.backend import GatePlanner, MI355XBackend
    from tests import standins as si

    gp = GatePlanner("complex128")
    pairs = [(si.RY(2, 0.7), gates.RY(2, 0.7)), (si.RZ(0, 1.3), gates.RZ(0, 1.3)), (si.CZ(1, 3), gates.CZ(1, 3)),
             (si.H(1), gates.H(1)), (si.CNOT(3, 0), gates.CNOT(3, 0)),
             (si.H(2).controlled_by(0, 1), gates.H(2).controlled_by(0, 1))]
    for foreign, own in pairs:
        a, b = gp.plan_gate(foreign), gp.plan_gate(own)
        assert (a.kind, a.targets, a.controls) == (b.kind, b.targets, b.controls)
        assert np.array_equal(a.data, b.data)
    m = si.M(0, 1, 2)
    assert kinds.is_measurement(m) and not kinds.runs_outside_passes(m)
    with pytest.raises(ValueError):
        gp.plan_gate(m)
    fg, own = si.FusedGate(0, 2), gates.FusedGate(0, 2)
    for g, h in ((si.H(0), gates.H(0)), (si.CZ(0, 2), gates.CZ(0, 2)), (si.RY(2, 0.3), gates.RY(2, 0.3))):
        fg.append(g)
        own.append(h)
    assert kinds.is_fused(fg) and not kinds.is_channel(fg)
    assert np.array_equal(gp.matrix_fused(fg), gp.matrix_fused(own))
    a = gp.plan_gate(fg)
    assert a.kind == gp.plan_gate(own).kind and np.array_equal(a.data, gp.plan_gate(own).data)
    be = MI355XBackend.__new__(MI355XBackend)
    assert be._gate_key(m) == ("M",)
    assert be._gate_key(si.M(0, collapse=True)) is None
    assert be._gate_key(si.RY(2, 0.7)) == be._gate_key(gates.RY(2, 0.7))
    assert be._gate_key(fg) is not None


def test_standin_measurement_bookkeeping():
    """Result bookkeeping on stand-in measurement gates: samples are
    registered per gate in its qubit order and read back per register."""
    import numpy as np

    from qibo_amd.result import MeasurementOutcomes
    from tests import standins as si

    m1, m2 = si.M(0, 2, register_name="a"), si.M(1, register_name="b")
    samples = np.array([[0, 1, 1], [1, 0, 0], [1, 1, 0]], dtype=np.int64)
    out = MeasurementOutcomes([m1, m2], backend=None, samples=samples, nshots=3)
    assert out.measurement_gate.qubits == (0, 2, 1)
    assert np.array_equal(m1.result.samples(), samples[:, [0, 1]])
    assert np.array_equal(m2.result.samples(), samples[:, [2]])
    regs = out.samples(registers=True)
    assert np.array_equal(regs["a"], samples[:, [0, 1]]) and np.array_equal(regs["b"], samples[:, [2]])


def test_shared_plans_of_parameterless_gates():
    """Gates without parameters share one plan per class, qubits and dtype;
    parametrized gates, other qubits, controls and dtypes get their own."""
    from qibo_amd.backend import GatePlanner

    g128, g64 = GatePlanner("complex128"), GatePlanner("complex64")
    a, b = g128.plan_gate(gates.CZ(3, 4)), g128.plan_gate(gates.CZ(3, 4))
    assert a is b
    assert g128.plan_gate(gates.CZ(4, 5)) is not a
    assert g64.plan_gate(gates.CZ(3, 4)) is not a
    h = g128.plan_gate(gates.H(4))
    hc = g128.plan_gate(gates.H(4).controlled_by(3))
    assert hc is not h and hc.controls == (3,) and h.controls == ()
    assert g128.plan_gate(gates.RY(3, 0.5)) is not g128.plan_gate(gates.RY(3, 0.5))
    cn = g128.plan_gate(gates.CNOT(1, 2))
    assert cn.kind == "dense" and cn.controls == (1,) and cn.targets == (2,)
