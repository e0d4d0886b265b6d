"""The native pass selection (qsim_plan_pass) against the planner's Python
rule (blocking._scan from every seed): the same passes, tile bits and gate
lists on layered, QFT, random-gate and adjoint-sweep circuits, ordered and
not. Host code only: needs the built library, no GPU."""

import numpy as np
import pytest

from qibo_amd import blocking, gates, models
from qibo_amd.backend import GatePlanner


def _random_circuit(n, ngates, rng):
    c = models.Circuit(n) if hasattr(models, "Circuit") else None
    from qibo_amd import Circuit

    c = Circuit(n)
    for _ in range(ngates):
        k = rng.integers(0, 6)
        q = [int(x) for x in rng.choice(n, 3, replace=False)]
        if k == 0:
            c.add(gates.RY(q[0], float(rng.uniform(0, 6))))
        elif k == 1:
            c.add(gates.RZ(q[0], float(rng.uniform(0, 6))))
        elif k == 2:
            c.add(gates.CZ(q[0], q[1]))
        elif k == 3:
            c.add(gates.CNOT(q[0], q[1]))
        elif k == 4:
            c.add(gates.TOFFOLI(q[0], q[1], q[2]))
        else:
            c.add(gates.H(q[0]))
    return c


def _circuits():
    rng = np.random.default_rng(9)
    yield models.random_layered_circuit(20, 8, seed=3, measure=False)
    yield models.random_layered_circuit(30, 4, seed=4, measure=False)
    yield models.QFT(16) if hasattr(models, "QFT") else models.random_layered_circuit(16, 3, seed=1, measure=False)
    yield _random_circuit(14, 400, rng)
    yield _random_circuit(25, 600, rng)


@pytest.mark.parametrize("ordered", [False, True])
@pytest.mark.parametrize("k", range(5))
def test_native_selection_equals_python(k, ordered):
    c = list(_circuits())[k]
    n = c.nqubits
    gp = GatePlanner()
    ops = [gp.plan_gate(g) for g in c.queue if g.__class__.__name__ != "M"]
    for low in (3, 4):
        py = list(blocking.partition(iter(ops), n, low, ordered, native=False))
        nat = list(blocking.partition(iter(ops), n, low, ordered, native=True))
        assert len(py) == len(nat)
        for (k1, p1), (k2, p2) in zip(py, nat):
            assert k1 == k2
            if k1 == "single":
                assert p1 is p2
            else:
                assert p1[0] == p2[0] and len(p1[1]) == len(p2[1])
                assert all(a is b for a, b in zip(p1[1], p2[1]))
