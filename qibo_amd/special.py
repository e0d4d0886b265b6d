"""Fusion groups and callback gates.

``FusedGate`` is the node of qibo's greedy neighbour-fusion planner
(``src/qibo/gates/special.py:28-153``). The planner's result (which gates end
up in which group, and in which order) decides the fused matrices, so the
decision rules are the same as there:

* a group can absorb a neighbour if neither is marked and the union of their
  qubits stays within ``max_qubits`` (special.py:81-91);
* the two groups fuse only if no third group blocks them on qubits outside the
  shared ones; the one with more outside neighbours becomes the parent and the
  other is appended (parent earlier in time) or prepended (parent later), and
  the neighbour links of the absorbed group are re-pointed to the parent
  (special.py:106-153).

The fused matrix itself is built by the backend (``matrix_fused``).
"""

from . import gates as _g  # gates imports this module last, after SpecialGate exists


class CallbackGate(_g.SpecialGate):
    """Run ``callback(backend, state)`` at this point of the circuit."""

    def __init__(self, callback):
        super().__init__()
        self.name = callback.__class__.__name__
        self.callback = callback
        self.init_args = [callback]

    def apply(self, backend, state, nqubits):
        self.callback.nqubits = nqubits
        self.callback.apply(backend, state)
        return state


class FusedGate(_g.SpecialGate):
    """Ordered list of gates applied as one dense block on ``target_qubits``."""

    def __init__(self, *q):
        super().__init__()
        self.name = "Fused Gate"
        self.draw_label = "[]"
        self.target_qubits = tuple(sorted(q))
        self.init_args = list(q)
        self.qubit_set = set(q)
        self.gates = []
        self.marked = False
        self.fused = False
        self.left_neighbors = {}
        self.right_neighbors = {}

    @classmethod
    def from_gate(cls, gate):
        fgate = cls(*gate.qubits)
        fgate.append(gate)
        if isinstance(gate, (_g.M, _g.SpecialGate)):
            fgate.marked = True  # special gates never take part in fusion
        return fgate

    def _grow(self, gate):
        self.qubit_set = self.qubit_set | set(gate.qubits)
        self.init_args = sorted(self.qubit_set)
        self.target_qubits = tuple(self.init_args)

    def prepend(self, gate):
        self._grow(gate)
        inner = gate.gates if isinstance(gate, self.__class__) else [gate]
        self.gates = list(inner) + self.gates

    def append(self, gate):
        self._grow(gate)
        if isinstance(gate, self.__class__):
            self.gates.extend(gate.gates)
        else:
            self.gates.append(gate)

    def _dagger(self):
        dagger = self.__class__(*self.init_args)
        for gate in reversed(self.gates):
            dagger.append(gate.dagger())
        return dagger

    def dagger(self):
        return self._dagger()

    def can_fuse(self, gate, max_qubits):
        if gate is None or self.marked or gate.marked:
            return False
        return len(self.qubit_set | gate.qubit_set) <= max_qubits

    def matrix(self, backend=None):
        from .backends import _check_backend

        return _check_backend(backend).matrix_fused(self)

    def fuse(self, gate):
        """Fuse ``gate`` (the later group in time) with ``self``."""
        blocking_left = set(self.right_neighbors.values()) - {gate}
        blocking_right = set(gate.left_neighbors.values()) - {self}
        if blocking_left and blocking_right:
            return
        shared = self.qubit_set & gate.qubit_set
        if len(blocking_left) > len(blocking_right):
            parent, child = self, gate
            if {parent.right_neighbors.get(q) for q in shared} == {child}:
                child.marked = True
                parent.append(child)
                for q in shared:
                    nxt = child.right_neighbors.get(q)
                    if nxt is not None:
                        parent.right_neighbors[q] = nxt
                        nxt.left_neighbors[q] = parent
                    else:
                        parent.right_neighbors.pop(q)
        else:
            parent, child = gate, self
            if {parent.left_neighbors.get(q) for q in shared} == {child}:
                child.marked = True
                parent.prepend(child)
                for q in shared:
                    prv = child.left_neighbors.get(q)
                    if prv is not None:
                        parent.left_neighbors[q] = prv
                        prv.right_neighbors[q] = parent
        if child.marked:
            for q in child.qubit_set - shared:
                nxt = child.right_neighbors.get(q)
                if nxt is not None:
                    parent.right_neighbors[q] = nxt
                    nxt.left_neighbors[q] = parent
                prv = child.left_neighbors.get(q)
                if prv is not None:
                    parent.left_neighbors[q] = prv
                    prv.right_neighbors[q] = parent
