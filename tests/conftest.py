import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def golden():
    with open(os.path.join(GOLDEN, "reference_values.json")) as fh:
        return json.load(fh)


def _gpu_backend(dtype):
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    from qibo_amd.backend import MI355XBackend

    # no fallback: a missing library raises ImportError here and fails the test
    b = MI355XBackend(dtype=dtype)
    # run-time specialised passes (background hiprtc compiles) are exercised by
    # test_gpu_specialize.py on backends of their own; the shared backends
    # interpret every pass, so the other tests need no compile time
    b.specialize = False
    return b


@pytest.fixture(scope="session")
def backend():
    return _gpu_backend("complex128")


@pytest.fixture(scope="session")
def backend64():
    return _gpu_backend("complex64")


@pytest.fixture
def rng():
    return np.random.default_rng(20241202)
