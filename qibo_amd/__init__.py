"""qibo_amd: an MI355X (gfx950) state-vector backend for Qibo.

Loaded by qibo as a backend plugin (``qibo.set_backend("qibo_amd")`` imports
this package and calls ``MetaBackend.load``). The package also carries the
circuit, gate and result classes of the state-vector path so that it runs on
machines where qibo itself is not installed:

    from qibo_amd import Circuit, gates
    c = Circuit(2); c.add(gates.H(0)); c.add(gates.CNOT(0, 1)); c.add(gates.M(0, 1))
    result = c(nshots=100)

Every numerical step runs in ``libqsim_mi355x.so`` (HIP, gfx950); importing
the backend without the built library raises ``ImportError``.
"""

from . import config, gates, hamiltonians, models
from .backends import (
    MetaBackend,
    construct_backend,
    get_backend,
    get_dtype,
    list_available_backends,
    set_backend,
    set_dtype,
    set_seed,
)
from .circuit import Circuit
from .config import set_hbm_budget, set_max_qubits
from .result import CircuitResult, MeasurementOutcomes, QuantumState

__version__ = "0.1.0"

__all__ = [
    "Circuit",
    "CircuitResult",
    "MeasurementOutcomes",
    "MetaBackend",
    "QuantumState",
    "config",
    "construct_backend",
    "gates",
    "get_backend",
    "get_dtype",
    "hamiltonians",
    "list_available_backends",
    "models",
    "set_backend",
    "set_dtype",
    "set_hbm_budget",
    "set_max_qubits",
    "set_seed",
]
