"""Full-size checks at the benchmark configurations (needs an MI355X).

Size-independent properties where the oracle cannot run the whole circuit:
* the 30-qubit, 20-layer complex128 layered circuit (16 GiB state): the
  circuit followed by its inverse returns |0...0>, and the probabilities
  sum to 1;
* 100 000 shots drawn from that state's 2^30 probabilities equal, bit for
  bit, numpy's np.random.choice on the same probabilities with the same seed
  (the reference's sample_shots, abstract.py:2774-2781), run on the host;
* the 32-qubit complex64 layered circuit grouped by ``fuse(max_qubits=5)``
  (32 GiB state) followed by the unfused inverse returns |0...0> to complex64
  accuracy. On this circuit qibo's fusion forms 2- and 3-qubit groups only,
  which run as their gates inside blocked passes (GatePlanner.plan_ops); the
  dense 5-qubit block kernels at 32 qubits are covered by
  tests/test_gpu_configs.py.
"""

import numpy as np
import pytest
import torch

from qibo_amd import models

pytestmark = pytest.mark.gpu


def test_layered30_round_trip(backend):
    n = 30
    c = models.random_layered_circuit(n, 20, seed=1234, measure=False)
    state = backend.execute_circuit(c).state()
    p = backend.calculate_probabilities(state, list(range(n)), n)
    assert abs(float(p.sum()) - 1.0) < 1e-10
    del p
    back = backend.execute_circuit(c.invert(), initial_state=state).state()
    del state
    assert abs(complex(back[0].item()) - 1.0) < 1e-10
    del back
    torch.cuda.empty_cache()


def test_layered30_samples_bit_exact(backend):
    n = 30
    c = models.random_layered_circuit(n, 20, seed=1234, measure=False)
    state = backend.execute_circuit(c).state()
    probs = backend.calculate_probabilities(state, list(range(n)), n)
    del state
    torch.cuda.empty_cache()
    np.random.seed(2024)
    got = backend.sample_shots(probs, 100000).cpu().numpy()
    host = probs.cpu().numpy()
    del probs
    torch.cuda.empty_cache()
    np.random.seed(2024)
    want = np.random.choice(np.arange(host.size), size=100000, p=host)
    assert np.array_equal(got, want)


def test_fused32_c64_round_trip(backend64):
    n = 32
    c = models.random_layered_circuit(n, 6, seed=7, measure=False)
    fused = c.fuse(max_qubits=5)
    state = backend64.execute_circuit(fused).state()
    back = backend64.execute_circuit(c.invert(), initial_state=state).state()
    del state
    assert abs(complex(back[0].item()) - 1.0) < 1e-4
    del back
    torch.cuda.empty_cache()
