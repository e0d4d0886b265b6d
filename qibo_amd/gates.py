"""Gate objects of the caller surface.

These mirror qibo's gate classes (``src/qibo/gates/abstract.py`` and
``src/qibo/gates/gates.py``) in everything the state-vector path reads:
class names (the backend looks matrices up by class name), ``qubits`` =
``control_qubits + target_qubits`` with controls sorted (abstract.py:440-442),
``is_controlled_by`` set only by ``controlled_by`` (abstract.py:154-180),
``parameters``/``nparams``/``trainable`` (abstract.py:701-744) and ``dagger``.
A gate never touches the state itself: ``apply`` hands it to the backend
(abstract.py:95-96), which owns the device buffer.
"""

import math
from collections.abc import Iterable
from typing import Sequence, Tuple

import numpy as np

from .config import PRECISION_TOL, raise_error


def _check_backend(backend):
    from .backends import _check_backend as _cb  # local import: avoids a cycle

    return _cb(backend)


class Gate:
    """Base class: qubits, controls, matrix lookup and dispatch to the backend."""

    def __init__(self):
        self.name = None
        self.draw_label = None
        self.is_controlled_by = False
        self.init_args = []
        self.init_kwargs = {}
        self.unitary = False
        self._target_qubits = ()
        self._control_qubits = ()
        self._parameters = ()
        self.symbolic_parameters = {}
        self.device_gates = set()
        self.original_gate = None

    # -- qubits ----------------------------------------------------------
    @property
    def target_qubits(self) -> Tuple[int, ...]:
        return self._target_qubits

    @target_qubits.setter
    def target_qubits(self, qubits: Sequence[int]) -> None:
        qubits = tuple(qubits)
        if len(set(qubits)) != len(qubits):
            raise_error(
                ValueError,
                f"Target qubit {_first_repeated(qubits)} was given twice for gate "
                f"{self.__class__.__name__}.",
            )
        self._target_qubits = qubits
        self._check_overlap()

    @property
    def control_qubits(self) -> Tuple[int, ...]:
        return tuple(sorted(self._control_qubits))

    @control_qubits.setter
    def control_qubits(self, qubits: Sequence[int]) -> None:
        qubits = tuple(qubits)
        if len(set(qubits)) != len(qubits):
            raise_error(
                ValueError,
                f"Control qubit {_first_repeated(qubits)} was given twice for gate "
                f"{self.__class__.__name__}.",
            )
        self._control_qubits = qubits
        self._check_overlap()

    @property
    def qubits(self) -> Tuple[int, ...]:
        return self.control_qubits + self.target_qubits

    def _check_overlap(self):
        both = tuple(self._control_qubits) + tuple(self._target_qubits)
        if len(set(both)) != len(both):
            raise_error(
                ValueError,
                f"{set(self._target_qubits) & set(self._control_qubits)}"
                f"qubits are both targets and controls for gate {self.__class__.__name__}.",
            )

    # -- parameters ------------------------------------------------------
    @property
    def parameters(self):
        return self._parameters

    @property
    def nparams(self) -> int:
        return getattr(self, "_nparams", 0)

    @nparams.setter
    def nparams(self, value):
        self._nparams = value

    # -- behaviour -------------------------------------------------------
    def apply(self, backend, state, nqubits):
        return backend.apply_gate(self, state, nqubits)

    def matrix(self, backend=None):
        backend = _check_backend(backend)
        return backend.matrix(self)

    def _controls_guard(self):
        if self.control_qubits:
            raise_error(
                RuntimeError,
                f"Cannot use `controlled_by` method on gate {self} because it is "
                f"already controlled by {self.control_qubits}.",
            )

    def controlled_by(self, *qubits: int):
        """Control the gate on ``qubits`` (abstract.py:154-180)."""
        self._controls_guard()
        if qubits:
            self.is_controlled_by = True
            self.control_qubits = qubits
        return self

    def _dagger(self):
        # Hermitian gates (H, X, Y, Z, CNOT, CZ, SWAP, ...) are their own dagger.
        return self.__class__(*self.init_args, **self.init_kwargs)

    def dagger(self):
        new_gate = self._dagger()
        new_gate.is_controlled_by = self.is_controlled_by
        new_gate.control_qubits = self.control_qubits
        return new_gate

    def generator_eigenvalue(self):
        raise_error(
            NotImplementedError,
            f"Generator eigenvalue is not implemented for {self.__class__.__name__}",
        )

    def __repr__(self):
        return f"{self.__class__.__name__}({', '.join(map(str, self.qubits))})"


def _first_repeated(qubits):
    seen = set()
    for q in qubits:
        if q in seen:
            return q
        seen.add(q)
    return None


class SpecialGate(Gate):
    """Gates without a matrix (callbacks, fused groups)."""

    def matrix(self, backend=None):  # pragma: no cover
        raise_error(NotImplementedError, "Special gates do not have matrix representation.")


class ParametrizedGate(Gate):
    """Gate whose matrix depends on ``parameters`` (abstract.py:676-752)."""

    def __init__(self, trainable: bool = True):
        super().__init__()
        self.parameter_names = "theta"
        self.nparams = 1
        self.trainable = trainable

    @Gate.parameters.setter
    def parameters(self, x):
        if isinstance(self.parameter_names, str):
            nparams = 1
            names = [self.parameter_names]
            if not isinstance(x, Iterable):
                x = [x]
            else:
                try:
                    if len(x) != 1:
                        x = [x]
                except TypeError:
                    x = [x]
        else:
            nparams = len(self.parameter_names)
            names = self.parameter_names
        x = list(x)
        if len(x) != nparams:
            raise_error(
                ValueError,
                f"Parametrized gate has {nparams} parameters but {len(x)} update values "
                "were given.",
            )
        for i, v in enumerate(x):
            if type(v).__module__.startswith("sympy"):  # measurement-conditioned (abstract.py:733-735)
                self.symbolic_parameters[i] = v
        self._parameters = tuple(x)
        self.init_kwargs.update(
            {n: v for n, v in zip(names, self._parameters) if n in self.init_kwargs}
        )

    def substitute_symbols(self):
        """Replace measurement symbols by the latest outcomes (abstract.py:746-752)."""
        params = list(self._parameters)
        for i, param in self.symbolic_parameters.items():
            for symbol in param.free_symbols:
                param = symbol.evaluate(param)
            params[i] = float(param)
        self.parameters = tuple(params)

    def matrix(self, backend=None):
        backend = _check_backend(backend)
        return backend.matrix_parametrized(self)


# ---------------------------------------------------------------------------
# single-qubit gates
# ---------------------------------------------------------------------------


class _OneQubit(Gate):
    _gname = None

    def __init__(self, q: int):
        super().__init__()
        self.name = self._gname
        self.draw_label = self.__class__.__name__
        self.target_qubits = (q,)
        self.init_args = [q]
        self.unitary = True


class H(_OneQubit):
    _gname = "h"


class X(_OneQubit):
    _gname = "x"

    def controlled_by(self, *q):
        """One control gives CNOT, two give TOFFOLI (gates.py:80-91)."""
        self._controls_guard()
        if len(q) == 1:
            return CNOT(q[0], self.target_qubits[0])
        if len(q) == 2:
            return TOFFOLI(q[0], q[1], self.target_qubits[0])
        return super().controlled_by(*q)


class Y(_OneQubit):
    _gname = "y"

    def controlled_by(self, *q):
        self._controls_guard()
        if len(q) == 1:
            return CY(q[0], self.target_qubits[0])
        return super().controlled_by(*q)


class Z(_OneQubit):
    _gname = "z"

    def controlled_by(self, *q):
        self._controls_guard()
        if len(q) == 1:
            return CZ(q[0], self.target_qubits[0])
        return super().controlled_by(*q)


class SX(_OneQubit):
    _gname = "sx"

    def _dagger(self):
        return SXDG(self.init_args[0])


class SXDG(_OneQubit):
    _gname = "sxdg"

    def _dagger(self):
        return SX(self.init_args[0])


class S(_OneQubit):
    _gname = "s"

    def _dagger(self):
        return SDG(*self.init_args)


class SDG(_OneQubit):
    _gname = "sdg"

    def _dagger(self):
        return S(*self.init_args)


class T(_OneQubit):
    _gname = "t"

    def _dagger(self):
        return TDG(*self.init_args)


class TDG(_OneQubit):
    _gname = "tdg"

    def _dagger(self):
        return T(*self.init_args)


class I(Gate):
    def __init__(self, *q: int):
        super().__init__()
        self.name = "id"
        self.target_qubits = tuple(q)
        self.init_args = list(q)
        self.unitary = True


class Align(ParametrizedGate):
    def __init__(self, q: int, delay: int = 0, trainable: bool = False):
        super().__init__(trainable)
        self.name = "align"
        self.target_qubits = (q,)
        self.init_args = [q]
        self.init_kwargs = {"delay": delay, "trainable": trainable}
        self.parameter_names = "delay"
        self.parameters = delay
        self.nparams = 1


class _Rn_(ParametrizedGate):
    """RX/RY/RZ: ``exp(-i theta P / 2)`` (gates.py:586-646)."""

    _gname = None
    _controlled = None

    def __init__(self, q: int, theta: float, trainable: bool = True):
        super().__init__(trainable)
        self.name = self._gname
        self.draw_label = self.__class__.__name__
        self.target_qubits = (q,)
        self.unitary = True
        self.init_args = [q]
        self.init_kwargs = {"theta": theta, "trainable": trainable}
        self.parameters = theta

    def _dagger(self):
        return self.__class__(self.target_qubits[0], -self.parameters[0])

    def controlled_by(self, *q):
        self._controls_guard()
        if len(q) == 1:
            return self._controlled()(q[0], self.target_qubits[0], **self.init_kwargs)
        return super().controlled_by(*q)

    def generator_eigenvalue(self):
        return 0.5


class RX(_Rn_):
    _gname = "rx"

    @staticmethod
    def _controlled():
        return CRX


class RY(_Rn_):
    _gname = "ry"

    @staticmethod
    def _controlled():
        return CRY


class RZ(_Rn_):
    _gname = "rz"

    @staticmethod
    def _controlled():
        return CRZ


class PRX(ParametrizedGate):
    def __init__(self, q, theta, phi, trainable=True):
        super().__init__(trainable)
        self.name = "prx"
        self.target_qubits = (q,)
        self.unitary = True
        self.parameter_names = ["theta", "phi"]
        self.nparams = 2
        self.init_args = [q]
        self.init_kwargs = {"theta": theta, "phi": phi, "trainable": trainable}
        self.parameters = theta, phi

    def _dagger(self):
        theta, phi = self.parameters
        return self.__class__(self.target_qubits[0], -theta, phi)


class GPI(ParametrizedGate):
    def __init__(self, q, phi, trainable=True):
        super().__init__(trainable)
        self.name = "gpi"
        self.target_qubits = (q,)
        self.unitary = True
        self.parameter_names = "phi"
        self.nparams = 1
        self.init_args = [q]
        self.init_kwargs = {"phi": phi, "trainable": trainable}
        self.parameters = phi


class GPI2(ParametrizedGate):
    def __init__(self, q, phi, trainable=True):
        super().__init__(trainable)
        self.name = "gpi2"
        self.target_qubits = (q,)
        self.unitary = True
        self.parameter_names = "phi"
        self.nparams = 1
        self.init_args = [q]
        self.init_kwargs = {"phi": phi, "trainable": trainable}
        self.parameters = phi

    def _dagger(self):
        return self.__class__(self.target_qubits[0], self.parameters[0] + math.pi)


class _Un_(ParametrizedGate):
    _controlled_name = None

    def __init__(self, q, trainable=True):
        super().__init__(trainable)
        self.nparams = 0
        self.target_qubits = (q,)
        self.init_args = [q]
        self.unitary = True
        self.init_kwargs = {"trainable": trainable}

    def controlled_by(self, *q):
        self._controls_guard()
        if len(q) == 1 and self._controlled_name is not None:
            cls = globals()[self._controlled_name]
            return cls(q[0], self.target_qubits[0], **self.init_kwargs)
        return super().controlled_by(*q)


class U1(_Un_):
    _controlled_name = "CU1"

    def __init__(self, q, theta, trainable=True):
        super().__init__(q, trainable=trainable)
        self.name = "u1"
        self.nparams = 1
        self.init_kwargs = {"theta": theta, "trainable": trainable}
        self.parameters = theta

    def _dagger(self):
        return self.__class__(self.target_qubits[0], -self.parameters[0])

    def generator_eigenvalue(self):
        return 0.5


class U2(_Un_):
    _controlled_name = "CU2"

    def __init__(self, q, phi, lam, trainable=True):
        super().__init__(q, trainable=trainable)
        self.name = "u2"
        self.nparams = 2
        self.parameter_names = ["phi", "lam"]
        self.init_kwargs = {"phi": phi, "lam": lam, "trainable": trainable}
        self.parameters = phi, lam

    def _dagger(self):
        phi, lam = self.parameters
        return self.__class__(self.target_qubits[0], math.pi - lam, -math.pi - phi)


class U3(_Un_):
    _controlled_name = "CU3"

    def __init__(self, q, theta, phi, lam, trainable=True):
        super().__init__(q, trainable=trainable)
        self.name = "u3"
        self.nparams = 3
        self.parameter_names = ["theta", "phi", "lam"]
        self.init_kwargs = {"theta": theta, "phi": phi, "lam": lam, "trainable": trainable}
        self.parameters = theta, phi, lam

    def _dagger(self):
        theta, lam, phi = tuple(-x for x in self.parameters)
        return self.__class__(self.target_qubits[0], theta, phi, lam)


class U1q(_Un_):
    def __init__(self, q, theta, phi, trainable=True):
        super().__init__(q, trainable=trainable)
        self.name = "u1q"
        self.nparams = 2
        self.parameter_names = ["theta", "phi"]
        self.init_kwargs = {"theta": theta, "phi": phi, "trainable": trainable}
        self.parameters = theta, phi

    def _dagger(self):
        theta, phi = self.parameters
        return self.__class__(self.init_args[0], -theta, phi)


# ---------------------------------------------------------------------------
# two- and three-qubit gates
# ---------------------------------------------------------------------------


class _Controlled(Gate):
    """Fixed gate with one (or two) built-in controls: dense matrix in qibo."""

    _gname = None
    _ncontrols = 1

    def __init__(self, *q):
        super().__init__()
        self.name = self._gname
        nc = self._ncontrols
        self.control_qubits = tuple(q[:nc])
        self.target_qubits = tuple(q[nc:])
        self.init_args = list(q)
        self.unitary = True


class CNOT(_Controlled):
    _gname = "cx"


class CY(_Controlled):
    _gname = "cy"


class CZ(_Controlled):
    _gname = "cz"


class CH(_Controlled):
    _gname = "ch"


class CSX(_Controlled):
    _gname = "csx"

    def _dagger(self):
        return CSXDG(*self.init_args)


class CSXDG(_Controlled):
    _gname = "csxdg"

    def _dagger(self):
        return CSX(*self.init_args)


class TOFFOLI(_Controlled):
    _gname = "ccx"
    _ncontrols = 2


class CCZ(_Controlled):
    _gname = "ccz"
    _ncontrols = 2


class _CRn_(ParametrizedGate):
    _gname = None

    def __init__(self, q0, q1, theta, trainable=True):
        super().__init__(trainable)
        self.name = self._gname
        self.control_qubits = (q0,)
        self.target_qubits = (q1,)
        self.unitary = True
        self.init_args = [q0, q1]
        self.init_kwargs = {"theta": theta, "trainable": trainable}
        self.parameters = theta

    def _dagger(self):
        return self.__class__(self.control_qubits[0], self.target_qubits[0], -self.parameters[0])

    def generator_eigenvalue(self):
        return 0.5


class CRX(_CRn_):
    _gname = "crx"


class CRY(_CRn_):
    _gname = "cry"


class CRZ(_CRn_):
    _gname = "crz"


class _CUn_(ParametrizedGate):
    def __init__(self, q0, q1, trainable=True):
        super().__init__(trainable)
        self.nparams = 0
        self.control_qubits = (q0,)
        self.target_qubits = (q1,)
        self.init_args = [q0, q1]
        self.unitary = True
        self.init_kwargs = {"trainable": trainable}


class CU1(_CUn_):
    def __init__(self, q0, q1, theta, trainable=True):
        super().__init__(q0, q1, trainable=trainable)
        self.name = "cu1"
        self.nparams = 1
        self.init_kwargs = {"theta": theta, "trainable": trainable}
        self.parameters = theta

    def _dagger(self):
        return self.__class__(self.control_qubits[0], self.target_qubits[0], -self.parameters[0])


class CU2(_CUn_):
    def __init__(self, q0, q1, phi, lam, trainable=True):
        super().__init__(q0, q1, trainable=trainable)
        self.name = "cu2"
        self.nparams = 2
        self.parameter_names = ["phi", "lam"]
        self.init_kwargs = {"phi": phi, "lam": lam, "trainable": trainable}
        self.parameters = phi, lam

    def _dagger(self):
        phi, lam = self.parameters
        return self.__class__(
            self.control_qubits[0], self.target_qubits[0], math.pi - lam, -math.pi - phi
        )


class CU3(_CUn_):
    def __init__(self, q0, q1, theta, phi, lam, trainable=True):
        super().__init__(q0, q1, trainable=trainable)
        self.name = "cu3"
        self.nparams = 3
        self.parameter_names = ["theta", "phi", "lam"]
        self.init_kwargs = {"theta": theta, "phi": phi, "lam": lam, "trainable": trainable}
        self.parameters = theta, phi, lam

    def _dagger(self):
        theta, lam, phi = tuple(-x for x in self.parameters)
        return self.__class__(self.control_qubits[0], self.target_qubits[0], theta, phi, lam)


class _TwoQubit(Gate):
    _gname = None

    def __init__(self, q0, q1):
        super().__init__()
        self.name = self._gname
        self.target_qubits = (q0, q1)
        self.init_args = [q0, q1]
        self.unitary = True


class SWAP(_TwoQubit):
    _gname = "swap"


class iSWAP(_TwoQubit):
    _gname = "iswap"


class SiSWAP(_TwoQubit):
    _gname = "siswap"

    def _dagger(self):
        return SiSWAPDG(*self.qubits)


class SiSWAPDG(_TwoQubit):
    _gname = "siswapdg"

    def _dagger(self):
        return SiSWAP(*self.qubits)


class FSWAP(_TwoQubit):
    _gname = "fswap"


class ECR(_TwoQubit):
    _gname = "ecr"


class SYC(_TwoQubit):
    _gname = "syc"


class fSim(ParametrizedGate):
    def __init__(self, q0, q1, theta, phi, trainable=True):
        super().__init__(trainable)
        self.name = "fsim"
        self.target_qubits = (q0, q1)
        self.unitary = True
        self.parameter_names = ["theta", "phi"]
        self.nparams = 2
        self.init_args = [q0, q1]
        self.init_kwargs = {"theta": theta, "phi": phi, "trainable": trainable}
        self.parameters = theta, phi

    def _dagger(self):
        q0, q1 = self.target_qubits
        return self.__class__(q0, q1, *(-x for x in self.parameters))


class GeneralizedfSim(ParametrizedGate):
    def __init__(self, q0, q1, unitary, phi, trainable=True):
        super().__init__(trainable)
        self.name = "generalizedfsim"
        self.target_qubits = (q0, q1)
        self.unitary = True
        self.parameter_names = ["unitary", "phi"]
        self.nparams = 5
        self.init_args = [q0, q1]
        self.init_kwargs = {"unitary": unitary, "phi": phi, "trainable": trainable}
        self._parameters = (np.asarray(unitary), phi)

    @Gate.parameters.setter
    def parameters(self, x):
        u, phi = x
        self._parameters = (np.reshape(u, (2, 2)), phi)

    def _dagger(self):
        u, phi = self.parameters
        q0, q1 = self.target_qubits
        return self.__class__(q0, q1, np.conj(np.transpose(u)), -phi)


class _Rnn_(ParametrizedGate):
    _gname = None

    def __init__(self, q0, q1, theta, trainable=True):
        super().__init__(trainable)
        self.name = self._gname
        self.target_qubits = (q0, q1)
        self.unitary = True
        self.init_args = [q0, q1]
        self.init_kwargs = {"theta": theta, "trainable": trainable}
        self.parameters = theta

    def _dagger(self):
        q0, q1 = self.target_qubits
        return self.__class__(q0, q1, -self.parameters[0])

    def generator_eigenvalue(self):
        return 0.5


class RXX(_Rnn_):
    _gname = "rxx"


class RYY(_Rnn_):
    _gname = "ryy"


class RZZ(_Rnn_):
    _gname = "rzz"


class RZX(_Rnn_):
    _gname = "rzx"


class RXXYY(_Rnn_):
    _gname = "rxxyy"


class MS(ParametrizedGate):
    def __init__(self, q0, q1, phi0, phi1, theta=math.pi / 2, trainable=True):
        super().__init__(trainable)
        self.name = "ms"
        self.target_qubits = (q0, q1)
        self.unitary = True
        self.parameter_names = ["phi0", "phi1", "theta"]
        self.nparams = 3
        self.init_args = [q0, q1]
        self.init_kwargs = {"phi0": phi0, "phi1": phi1, "theta": theta, "trainable": trainable}
        self.parameters = phi0, phi1, theta

    def _dagger(self):
        phi0, phi1, theta = self.parameters
        q0, q1 = self.target_qubits
        return self.__class__(q0, q1, phi0 + math.pi, phi1, theta)


class GIVENS(ParametrizedGate):
    def __init__(self, q0, q1, theta, trainable=True):
        super().__init__(trainable)
        self.name = "g"
        self.target_qubits = (q0, q1)
        self.unitary = True
        self.parameter_names = "theta"
        self.nparams = 1
        self.init_args = [q0, q1]
        self.init_kwargs = {"theta": theta, "trainable": trainable}
        self.parameters = theta

    def _dagger(self):
        q0, q1 = self.target_qubits
        return self.__class__(q0, q1, -self.parameters[0])


class RBS(GIVENS):
    def __init__(self, q0, q1, theta, trainable=True):
        super().__init__(q0, q1, theta, trainable)
        self.name = "rbs"


class DEUTSCH(ParametrizedGate):
    def __init__(self, q0, q1, q2, theta, trainable=True):
        super().__init__(trainable)
        self.name = "deutsch"
        self.control_qubits = (q0, q1)
        self.target_qubits = (q2,)
        self.unitary = True
        self.init_args = [q0, q1, q2]
        self.init_kwargs = {"theta": theta, "trainable": trainable}
        self.parameters = theta


class Unitary(ParametrizedGate):
    """Arbitrary matrix on ``q`` (gates.py:2774-2877)."""

    def __init__(self, unitary, *q, trainable=True, name=None, check_unitary=True):
        super().__init__(trainable)
        self.name = "Unitary" if name is None else name
        self.target_qubits = tuple(q)
        self.parameter_names = "u"
        u = np.asarray(unitary)
        dim = 2 ** len(self.target_qubits)
        if u.shape != (dim, dim):
            raise_error(ValueError, f"Unitary of shape {u.shape} does not act on {len(q)} qubits.")
        self._parameters = (u,)
        self.nparams = 4 ** len(self.target_qubits)
        self.init_args = [unitary] + list(q)
        self.init_kwargs = {"name": name, "check_unitary": check_unitary, "trainable": trainable}
        if check_unitary:
            product = np.conj(u).T @ u
            diag_ok = bool(np.all(np.abs(1 - np.diag(product)) < PRECISION_TOL))
            off_ok = bool(np.all(np.abs(product - np.diag(np.diag(product))) < PRECISION_TOL))
            self.unitary = diag_ok and off_ok

    @Gate.parameters.setter
    def parameters(self, x):
        shape = self._parameters[0].shape
        if isinstance(x, tuple):
            x = x[0]
        self._parameters = (np.reshape(np.asarray(x), shape),)

    def _dagger(self):
        ud = np.conj(self.parameters[0].T)
        return self.__class__(ud, *self.target_qubits, **self.init_kwargs)


CX = CNOT
CCX = TOFFOLI

# special and measurement gates live in their own modules; re-exported here
from .special import CallbackGate, FusedGate  # noqa: E402  (re-export)
from .measurements_gate import M  # noqa: E402  (re-export)
from .channels import (  # noqa: E402  (re-export)
    AmplitudeDampingChannel,
    Channel,
    DepolarizingChannel,
    KrausChannel,
    PauliNoiseChannel,
    PhaseDampingChannel,
    ReadoutErrorChannel,
    ResetChannel,
    ThermalRelaxationChannel,
    UnitaryChannel,
)
