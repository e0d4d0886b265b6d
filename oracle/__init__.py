"""CPU oracle of the state-vector hot path -- TEST INFRASTRUCTURE ONLY.

This package restates, in plain numpy/scipy, what qibo 0.3.5's NumpyBackend
does on the hot path (gate application, gate and fused matrices,
probabilities, shot sampling, frequencies, Pauli-sum expectations and the
gradient rules). It is the checker the MI355X backend is compared against and
the CPU baseline that ``bench.py`` times beside the GPU. Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg may import it;
the product package ``qibo_amd`` never does.

Every function cites the reference lines it restates. How each one is pinned
(golden vectors from the reference's tests, closed forms, brute force) is
listed in ``tests/test_oracle_pinning.py``.
"""

from .numpy_backend import OracleBackend, OracleMatrices  # noqa: F401
from .density import OracleDensity  # noqa: F401,E402
