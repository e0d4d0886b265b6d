"""The C-ABI library: loads, carries the ABI version, exports every entry point
that include/qsim_mi355x.h declares, and rejects bad arguments on the host
before any device work (no GPU needed: each call below fails validation)."""

import ctypes
import os
import re

import pytest

from qibo_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "qsim_mi355x.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(qsim_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared_symbols()
    assert len(names) >= 15
    for name in names:
        assert hasattr(lib, name), name
        assert name in _lib.SIGNATURES, f"{name} has no ctypes signature"
    assert set(_lib.SIGNATURES) == set(names)


def test_version_and_error_string():
    lib = _lib.load()
    assert lib.qsim_version() == _lib.ABI_VERSION
    assert isinstance(lib.qsim_last_error(), bytes)


def test_host_validation_rejects_bad_arguments():
    lib = _lib.load()
    fake = ctypes.c_void_p(16)  # never dereferenced: validation fails first
    mat = (ctypes.c_double * 8)()
    # qubit out of range
    rc = lib.qsim_apply_dense(fake, 4, _lib.C128, _lib.int_array([7]), 1, None, 0, mat, None)
    assert rc == _lib.EINVAL and b"out of range" in lib.qsim_last_error()
    # repeated qubit across targets and controls
    rc = lib.qsim_apply_dense(fake, 4, _lib.C128, _lib.int_array([1]), 1, _lib.int_array([1]), 1, mat, None)
    assert rc == _lib.EINVAL and b"twice" in lib.qsim_last_error()
    # bad dtype
    assert lib.qsim_apply_diag(fake, 4, 9, _lib.int_array([0]), 1, None, 0, mat, None) == _lib.EINVAL
    # too many qubits
    assert lib.qsim_zero_state(fake, 99, _lib.C128, None) == _lib.EINVAL
    # null pointers
    assert lib.qsim_probabilities(None, 3, _lib.C128, _lib.int_array([0]), 1, None, None, 0, None) == _lib.EINVAL
    # workspace too small
    n = 1 << 20
    assert lib.qsim_cumsum_workspace(n) > 0
    assert lib.qsim_cumsum_exact(fake, _lib.F64, n, fake, fake, 8, None) == _lib.EINVAL
    with pytest.raises(ValueError):
        _lib.check(_lib.EINVAL, lib)


def test_workspace_queries():
    lib = _lib.load()
    assert lib.qsim_probabilities_workspace(10, 10) == 0
    assert lib.qsim_probabilities_workspace(28, 2) > 0
    assert lib.qsim_expval_workspace(20, 39) >= 39 * 8
