"""Global constants, the error helper and the logger of the MI355X backend.

The names and defaults follow ``src/qibo/config.py`` of qibo 0.3.5:
``SHOT_BATCH_SIZE`` (:41), ``MAX_QUBITS`` (:22-28), ``PRECISION_TOL`` and
``raise_error`` (:104-112), so that code written against qibo's conventions
sees the same values and the same exception types here.
"""

import logging
import os

# Number of shots drawn per batch by ``sample_frequencies`` (qibo config.py:41).
# The batching decides how many uniforms each ``sample_shots`` call consumes,
# so it must match for seeded frequencies to agree.
SHOT_BATCH_SIZE = 2**18

# Largest state vector allowed; -1 disables the check (qibo config.py:22-28).
MAX_QUBITS = int(os.environ.get("QIBO_MAX_QUBITS", "-1"))

# Largest density matrix allowed; -1 disables the check (qibo config.py:26-28).
MAX_QUBITS_DM = int(os.environ.get("QIBO_MAX_QUBITS_DM", "-1"))

# Tolerance used by unitary checks (qibo config.py:31).
PRECISION_TOL = 1e-8

# Default device memory budget for state vectors, in bytes. ``None`` means the
# whole device. Forcing a small budget makes allocation raise ``MemoryError``
# on the host before any device call, which ``execute_circuit`` maps to the
# same ``RuntimeError`` qibo raises (abstract.py:2502-2511).
HBM_BUDGET_BYTES = (
    int(os.environ["QIBO_AMD_HBM_BUDGET"]) if "QIBO_AMD_HBM_BUDGET" in os.environ else None
)

_LOG_LEVEL = int(os.environ.get("QIBO_LOG_LEVEL", "1"))


class _CustomFormatter(logging.Formatter):
    def format(self, record):
        return f"[qibo_amd|{record.levelname}|{self.formatTime(record, '%Y-%m-%d %H:%M:%S')}]: {record.getMessage()}"


log = logging.getLogger("qibo_amd.config")
if not log.handlers:
    _handler = logging.StreamHandler()
    _handler.setFormatter(_CustomFormatter())
    log.addHandler(_handler)
log.setLevel(10 * _LOG_LEVEL)
log.propagate = False


def raise_error(exception, message=None):
    """Log ``message`` at error level, then raise ``exception(message)``.

    Mirrors ``qibo.config.raise_error`` (config.py:104-112).
    """
    if message:
        log.error(message)
        raise exception(message)
    raise exception


def set_max_qubits(n: int) -> None:
    """Set the largest allowed state vector (-1 disables the check)."""
    global MAX_QUBITS
    if not isinstance(n, int) or (n < 1 and n != -1):
        raise_error(ValueError, f"max qubits must be a positive integer or -1, got {n}.")
    MAX_QUBITS = n


def set_max_qubits_dm(n: int) -> None:
    """Set the largest allowed density matrix (-1 disables the check)."""
    global MAX_QUBITS_DM
    if not isinstance(n, int) or (n < 1 and n != -1):
        raise_error(ValueError, f"max qubits must be a positive integer or -1, got {n}.")
    MAX_QUBITS_DM = n


def set_hbm_budget(nbytes) -> None:
    """Set the device memory budget for state vectors (``None`` = whole device)."""
    global HBM_BUDGET_BYTES
    if nbytes is not None and (not isinstance(nbytes, int) or nbytes <= 0):
        raise_error(ValueError, f"HBM budget must be a positive int or None, got {nbytes}.")
    HBM_BUDGET_BYTES = nbytes
