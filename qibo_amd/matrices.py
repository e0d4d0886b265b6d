"""Host-side gate matrices, built with numpy on the CPU and uploaded as kernel
arguments (or, for fused blocks, as a small device buffer).

The values must be bit-identical to qibo's ``NumpyMatrices``
(``src/qibo/backends/npmatrices.py:12-577``): a gate matrix is the one input of
a gate pass that the host computes, so any last-bit difference here would show
up as a state difference that no kernel could remove. Each builder therefore
performs the same numpy operations, on the same dtypes and in the same order,
as the corresponding formula there (cited per entry). Attribute names equal the
gate class names because the backend resolves them with
``getattr(matrices, gate.__class__.__name__)`` exactly like
``Backend.matrix`` (``src/qibo/backends/abstract.py:2649-2678``).
"""

import cmath
import math
from functools import cached_property

import numpy as np

from .config import raise_error

_SQRT2 = math.sqrt(2)


class GateMatrices:
    """Matrix of every supported gate as a numpy array of ``dtype``."""

    def __init__(self, dtype):
        self.dtype = dtype
        self.engine = np

    # -- helpers ---------------------------------------------------------
    def _arr(self, rows):
        return np.array(rows, dtype=self.dtype)

    def _embed_controlled(self, block, total_dim):
        """``[[I, 0], [0, block]]`` as a Python list (one control qubit)."""
        d = len(block)
        rows = []
        for r in range(total_dim):
            if r < total_dim - d:
                rows.append([1 if c == r else 0 for c in range(total_dim)])
            else:
                rows.append([0] * (total_dim - d) + list(block[r - (total_dim - d)]))
        return rows

    # -- fixed single-qubit gates (npmatrices.py:26-75) ------------------
    @cached_property
    def H(self):
        return self._arr([[1, 1], [1, -1]]) / _SQRT2

    @cached_property
    def X(self):
        return self._arr([[0, 1], [1, 0]])

    @cached_property
    def Y(self):
        return self._arr([[0j, -1j], [1j, 0j]])

    @cached_property
    def Z(self):
        return self._arr([[1, 0], [0, -1]])

    @cached_property
    def SX(self):
        return self._arr([[1 + 1j, 1 - 1j], [1 - 1j, 1 + 1j]]) / 2

    @cached_property
    def SXDG(self):
        return self._arr([[1 - 1j, 1 + 1j], [1 + 1j, 1 - 1j]]) / 2

    @cached_property
    def S(self):
        return self._arr([[1 + 0j, 0j], [0j, 1j]])

    @cached_property
    def SDG(self):
        return self._arr([[1 + 0j, 0j], [0j, -1j]])

    @cached_property
    def T(self):
        return self._arr([[1 + 0j, 0], [0, cmath.exp(1j * math.pi / 4.0)]])

    @cached_property
    def TDG(self):
        return self._arr([[1 + 0j, 0], [0, cmath.exp(-1j * math.pi / 4.0)]])

    def I(self, n: int = 2):
        return np.eye(n, dtype=self.dtype)

    def Align(self, delay, n=2):
        return self.I(n)

    # -- parametrized single-qubit gates (npmatrices.py:80-135) -----------
    def RX(self, theta):
        c = np.cos(theta / 2.0) + 0j
        s = -1j * np.sin(theta / 2.0)
        return self._arr([[c, s], [s, c]])

    def RY(self, theta):
        c = np.cos(theta / 2.0) + 0j
        s = np.sin(theta / 2.0) + 0j
        return self._arr([[c, -s], [s, c]])

    def RZ(self, theta):
        ph = np.exp(0.5j * theta)
        return self._arr([[np.conj(ph), 0], [0, ph]])

    def PRX(self, theta, phi):
        c = np.cos(theta / 2)
        s = np.sin(theta / 2)
        e1 = -1.0j * np.exp(-1.0j * phi)
        e2 = -1.0j * np.exp(1.0j * phi)
        return self._arr([[c + 0j, e1 * s], [e2 * s, c + 0j]])

    def GPI(self, phi):
        ph = np.exp(1.0j * phi)
        return self._arr([[0, np.conj(ph)], [ph, 0]])

    def GPI2(self, phi):
        ph = np.exp(1.0j * phi)
        return self._arr([[1, -1.0j * np.conj(ph)], [-1.0j * ph, 1]]) / _SQRT2

    def U1(self, theta):
        ph = np.exp(1j * theta)
        return self._arr([[1, 0], [0, ph]])

    def U2(self, phi, lam):
        ep = np.exp(1j * (phi + lam) / 2.0)
        em = np.exp(1j * (phi - lam) / 2.0)
        return self._arr([[np.conj(ep), -np.conj(em)], [em, ep]]) / _SQRT2

    def U3(self, theta, phi, lam):
        ct = np.cos(theta / 2)
        st = np.sin(theta / 2)
        ep = np.exp(1j * (phi + lam) / 2.0)
        em = np.exp(1j * (phi - lam) / 2.0)
        return self._arr([[np.conj(ep) * ct, -np.conj(em) * st], [em * st, ep * ct]])

    def U1q(self, theta, phi):
        return self.U3(theta, phi - math.pi / 2, math.pi / 2 - phi).astype(self.dtype)

    # -- two-qubit gates (npmatrices.py:137-420) ---------------------------
    @cached_property
    def CNOT(self):
        return self._arr(self._embed_controlled([[0, 1], [1, 0]], 4))

    @cached_property
    def CY(self):
        return self._arr(
            [[1 + 0j, 0j, 0j, 0j], [0j, 1 + 0j, 0j, 0j], [0j, 0j, 0j, -1j], [0j, 0j, 1j, 0j]]
        )

    @cached_property
    def CZ(self):
        return self._arr(self._embed_controlled([[1, 0], [0, -1]], 4))

    @cached_property
    def CH(self):
        e = 1 / math.sqrt(2)
        return self._arr(self._embed_controlled([[e, e], [e, -e]], 4))

    @cached_property
    def CSX(self):
        a, b = (1 + 1j) / 2, (1 - 1j) / 2
        return self._arr([[1 + 0j, 0, 0, 0], [0, 1 + 0j, 0, 0], [0, 0, a, b], [0, 0, b, a]])

    @cached_property
    def CSXDG(self):
        a, b = (1 + 1j) / 2, (1 - 1j) / 2
        return self._arr([[1 + 0j, 0, 0, 0], [0, 1 + 0j, 0, 0], [0, 0, b, a], [0, 0, a, b]])

    def CRX(self, theta):
        c = np.cos(theta / 2.0) + 0j
        s = -1j * np.sin(theta / 2.0)
        return self._arr(self._embed_controlled([[c, s], [s, c]], 4))

    def CRY(self, theta):
        c = np.cos(theta / 2.0) + 0j
        s = np.sin(theta / 2.0) + 0j
        return self._arr(self._embed_controlled([[c, -s], [s, c]], 4))

    def CRZ(self, theta):
        ph = np.exp(0.5j * theta)
        return self._arr(self._embed_controlled([[np.conj(ph), 0], [0, ph]], 4))

    def CU1(self, theta):
        ph = np.exp(1j * theta)
        return self._arr(self._embed_controlled([[1, 0], [0, ph]], 4))

    def CU2(self, phi, lam):
        ep = np.exp(1j * (phi + lam) / 2.0) / math.sqrt(2)
        em = np.exp(1j * (phi - lam) / 2.0) / math.sqrt(2)
        return self._arr(self._embed_controlled([[np.conj(ep), -np.conj(em)], [em, ep]], 4))

    def CU3(self, theta, phi, lam):
        ct = np.cos(theta / 2)
        st = np.sin(theta / 2)
        ep = np.exp(1j * (phi + lam) / 2.0)
        em = np.exp(1j * (phi - lam) / 2.0)
        blk = [[np.conj(ep) * ct, -np.conj(em) * st], [em * st, ep * ct]]
        return self._arr(self._embed_controlled(blk, 4))

    @cached_property
    def SWAP(self):
        return self._arr([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]])

    @cached_property
    def iSWAP(self):
        return self._arr(
            [[1 + 0j, 0j, 0j, 0j], [0j, 0j, 1j, 0j], [0j, 1j, 0j, 0j], [0j, 0j, 0j, 1 + 0j]]
        )

    @cached_property
    def SiSWAP(self):
        r = 1 / math.sqrt(2)
        i = 1j / math.sqrt(2)
        return self._arr(
            [[1 + 0j, 0j, 0j, 0j], [0j, r + 0j, i, 0j], [0j, i, r + 0j, 0j], [0j, 0j, 0j, 1 + 0j]]
        )

    @cached_property
    def SiSWAPDG(self):
        r = 1 / math.sqrt(2)
        i = -1j / math.sqrt(2)
        return self._arr(
            [[1 + 0j, 0j, 0j, 0j], [0j, r + 0j, i, 0j], [0j, i, r + 0j, 0j], [0j, 0j, 0j, 1 + 0j]]
        )

    @cached_property
    def FSWAP(self):
        return self._arr([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, -1]])

    def fSim(self, theta, phi):
        c = np.cos(theta) + 0j
        s = -1j * np.sin(theta)
        ph = np.exp(-1j * phi)
        return self._arr([[1, 0, 0, 0], [0, c, s, 0], [0, s, c, 0], [0, 0, 0, ph]])

    @cached_property
    def SYC(self):
        c = math.cos(math.pi / 2) + 0j
        s = -1j * math.sin(math.pi / 2)
        ph = cmath.exp(-1j * math.pi / 6)
        return self._arr([[1 + 0j, 0, 0, 0], [0, c, s, 0], [0, s, c, 0], [0, 0, 0, ph]])

    def GeneralizedfSim(self, u, phi):
        ph = np.exp(-1j * phi)
        return self._arr(
            [
                [1, 0, 0, 0],
                [0, complex(u[0, 0]), complex(u[0, 1]), 0],
                [0, complex(u[1, 0]), complex(u[1, 1]), 0],
                [0, 0, 0, ph],
            ]
        )

    def RXX(self, theta):
        c = np.cos(theta / 2.0) + 0j
        s = -1j * np.sin(theta / 2.0)
        return self._arr([[c, 0, 0, s], [0, c, s, 0], [0, s, c, 0], [s, 0, 0, c]])

    def RYY(self, theta):
        c = np.cos(theta / 2.0) + 0j
        s = -1j * np.sin(theta / 2.0)
        return self._arr([[c, 0, 0, -s], [0, c, s, 0], [0, s, c, 0], [-s, 0, 0, c]])

    def RZZ(self, theta):
        ph = np.exp(0.5j * theta)
        cph = np.conj(ph)
        return self._arr([[cph, 0, 0, 0], [0, ph, 0, 0], [0, 0, ph, 0], [0, 0, 0, cph]])

    def RZX(self, theta):
        c, s = np.cos(theta / 2) + 0j, np.sin(theta / 2) + 0j
        return self._arr(
            [[c, -1j * s, 0, 0], [-1j * s, c, 0, 0], [0, 0, c, 1j * s], [0, 0, 1j * s, c]]
        )

    def RXXYY(self, theta):
        c, s = np.cos(theta / 2) + 0j, np.sin(theta / 2) + 0j
        return self._arr([[1, 0, 0, 0], [0, c, -1j * s, 0], [0, -1j * s, c, 0], [0, 0, 0, 1]])

    def MS(self, phi0, phi1, theta):
        plus = np.exp(1.0j * (phi0 + phi1))
        minus = np.exp(1.0j * (phi0 - phi1))
        c = np.cos(theta / 2) + 0j
        s = np.sin(theta / 2) + 0j
        return self._arr(
            [
                [c, 0, 0, -1.0j * np.conj(plus) * s],
                [0, c, -1.0j * np.conj(minus) * s, 0],
                [0, -1.0j * minus * s, c, 0],
                [-1.0j * plus * s, 0, 0, c],
            ]
        )

    def GIVENS(self, theta):
        return self._arr(
            [
                [1, 0, 0, 0],
                [0, np.cos(theta), -np.sin(theta), 0],
                [0, np.sin(theta), np.cos(theta), 0],
                [0, 0, 0, 1],
            ]
        )

    def RBS(self, theta):
        return self.GIVENS(-theta)

    @cached_property
    def ECR(self):
        return (
            self._arr(
                [
                    [0j, 0j, 1 + 0j, 1j],
                    [0j, 0j, 1j, 1 + 0j],
                    [1 + 0j, -1j, 0j, 0j],
                    [-1j, 1 + 0j, 0j, 0j],
                ]
            )
            / _SQRT2
        )

    # -- three-qubit gates (npmatrices.py:484-530) -------------------------
    @cached_property
    def TOFFOLI(self):
        m = np.eye(8, dtype=np.int64)
        m[6:, 6:] = [[0, 1], [1, 0]]
        return self._arr(m.tolist())

    @cached_property
    def CCZ(self):
        m = np.eye(8, dtype=np.int64)
        m[7, 7] = -1
        return self._arr(m.tolist())

    def DEUTSCH(self, theta):
        s = np.sin(theta) + 0j
        c = np.cos(theta) + 0j
        rows = [[1 if col == r else 0 for col in range(8)] for r in range(6)]
        rows.append([0] * 6 + [1j * c, s])
        rows.append([0] * 6 + [s, 1j * c])
        return self._arr(rows)

    def Unitary(self, u):
        return np.asarray(u, dtype=self.dtype)

    def M(self):  # pragma: no cover - measurement has no matrix
        raise_error(NotImplementedError)

    def FusedGate(self):  # pragma: no cover - built by the backend
        raise_error(NotImplementedError)
