"""ctypes binding of ``libqsim_mi355x.so`` (C ABI in include/qsim_mi355x.h).

The library is loaded from this package directory, where ``_build.py`` puts
it. There is no fallback: if the library is missing or its ABI version does
not match, importing the backend raises ``ImportError`` naming the package,
which qibo's ``construct_backend`` turns into ``MissingBackend``
(``src/qibo/backends/__init__.py:342-350``).
"""

import ctypes
import os
import threading

ABI_VERSION = 11
# QIBO_AMD_LIB names another build of the library (an A/B comparison of two
# builds on one device); by default the one next to this file
LIB_PATH = os.environ.get("QIBO_AMD_LIB") or os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                          "libqsim_mi355x.so")

C128, C64 = 0, 1
SPEC_PASS, SPEC_PASS_FROM_ZERO, SPEC_ADJOINT = 0, 1, 2
F64, F32 = 0, 1
OK, EINVAL, ENOMEM, EHIP = 0, -1, -2, -3

_c_int_p = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p
_i = ctypes.c_int
_i64 = ctypes.c_int64
_sz = ctypes.c_size_t

# name -> (restype, argtypes); every exported entry point of the header
SIGNATURES = {
    "qsim_version": (_i, []),
    "qsim_last_error": (ctypes.c_char_p, []),
    "qsim_zero_state": (_i, [_vp, _i, _i, _vp]),
    "qsim_zero_outside_support": (_i, [_vp, _i, _i, ctypes.c_uint64, _vp]),
    "qsim_collapse": (_i, [_vp, _i, _i, _c_int_p, _i, _i64, ctypes.c_double, _vp]),
    "qsim_prob_stats_workspace": (_sz, [_i64]),
    "qsim_prob_stats": (_i, [_vp, _i, _i64, _vp, _vp, _sz, _vp]),
    "qsim_apply_dense": (_i, [_vp, _i, _i, _c_int_p, _i, _c_int_p, _i, _vp, _vp]),
    "qsim_apply_diag": (_i, [_vp, _i, _i, _c_int_p, _i, _c_int_p, _i, _vp, _vp]),
    "qsim_apply_swap": (_i, [_vp, _i, _i, _i, _i, _c_int_p, _i, _vp]),
    "qsim_apply_fused": (_i, [_vp, _i, _i, _c_int_p, _i, _vp, _vp]),
    "qsim_probabilities_workspace": (_sz, [_i, _i]),
    "qsim_probabilities": (_i, [_vp, _i, _i, _c_int_p, _i, _vp, _vp, _sz, _vp]),
    "qsim_cumsum_workspace": (_sz, [_i64]),
    "qsim_cumsum_exact": (_i, [_vp, _i, _i64, _vp, _vp, _sz, _vp]),
    "qsim_np_sum_workspace": (_sz, [_i64]),
    "qsim_np_sum": (_i, [_vp, _i, _i64, _vp, _vp, _sz, _vp]),
    "qsim_divide_by": (_i, [_vp, _i, _i64, _vp, _vp, _vp]),
    "qsim_search_sorted": (_i, [_vp, _i64, _vp, _i64, _vp, _vp]),
    "qsim_sample_workspace": (_sz, [_i64]),
    "qsim_sample_exact": (_i, [_vp, _i, _i64, _vp, _vp, _i64, _vp, _vp, _vp, _sz, _vp]),
    "qsim_tile_stats_bytes": (_sz, [_i64]),
    "qsim_probabilities_tiles": (_i, [_vp, _i, _i, _vp, _vp, _sz, _vp]),
    "qsim_samples_to_binary": (_i, [_vp, _i64, _i, _vp, _vp]),
    "qsim_histogram": (_i, [_vp, _i64, _vp, _i64, _vp]),
    "qsim_expval_workspace": (_sz, [_i, _i]),
    "qsim_expval_pauli": (_i, [_vp, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _sz, _vp]),
    "qsim_expval_pauli_noalloc": (_i, [_vp, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp]),
    "qsim_expval_noalloc_max_groups": (_i, []),
    "qsim_apply_pauli_sum": (_i, [_vp, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp]),
    "qsim_inner_matrix": (_i, [_vp, _vp, _i, _i, _c_int_p, _i, _c_int_p, _i, _vp, _vp, _vp, _sz, _vp]),
    "qsim_blocked_header_bytes": (_sz, []),
    "qsim_apply_blocked": (_i, [_vp, _i, _i, _c_int_p, _vp, _sz, _i, _vp]),
    "qsim_apply_blocked_from_zero": (_i, [_vp, _i, _i, _c_int_p, _vp, _sz, _i, _vp]),
    "qsim_apply_blocked_batched": (_i, [_vp, _i, _i, _i, _c_int_p, _vp, _sz, _sz, _i, _vp]),
    "qsim_blocked_specialize": (_i, [_vp, _sz, _i, _i, _i, _vp]),
    "qsim_blocked_specialized_lookup": (_i, [_vp, _sz, _i, _i, _i, _vp, _vp]),
    "qsim_apply_blocked_specialized": (_i, [ctypes.c_uint64, _vp, _i, _i, _i, _c_int_p, ctypes.c_uint64, _vp, _sz,
                                            _sz, _i, _vp]),
    "qsim_apply_blocked_support": (_i, [_vp, _i, _i, _c_int_p, ctypes.c_uint64, _vp, _sz, _i, _vp]),
    "qsim_adjoint_blocked_specialized": (_i, [ctypes.c_uint64, _vp, _vp, _i, _i, _i, _c_int_p, _vp, _sz, _sz, _i,
                                              _i, _vp, _sz, _vp]),
    "qsim_blocked_specialized_source": (_sz, [_vp, _sz, _i, _i, _i, ctypes.c_char_p, _sz]),
    "qsim_blocked_compile_check": (_i, [_vp, _sz, _i, _i, _i, _vp]),
    "qsim_blocked_specialized_count": (_sz, []),
    "qsim_adjoint_partials_count": (_sz, [_i, _i, _i]),
    "qsim_adjoint_blocked": (_i, [_vp, _vp, _i, _i, _i, _c_int_p, _vp, _sz, _sz, _i, _i, _vp, _sz, _vp]),
    "qsim_adjoint_reduce": (_i, [_vp, _sz, _i, _i, _i, _vp, _i, _vp, _vp]),
    "qsim_dm_axpy": (_i, [_vp, _vp, ctypes.c_uint64, _i, ctypes.c_double, _i, _vp]),
    "qsim_dm_depolarize": (_i, [_vp, _i, _i, _c_int_p, _i, ctypes.c_double, _vp]),
    "qsim_dm_reset": (_i, [_vp, _i, _i, _i, ctypes.c_double, ctypes.c_double, _vp]),
    "qsim_dm_probabilities": (_i, [_vp, _i, _i, _c_int_p, _i, _vp, _vp]),
    "qsim_dm_expval_workspace": (_sz, [_i]),
    "qsim_dm_expval_pauli": (_i, [_vp, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _sz, _vp]),
    "qsim_plan_pass": (_i, [_i, _vp, _vp, _vp, _i, _i, _i, _i, _vp, _vp]),
    "qsim_compile_pass_capacity": (_sz, [_i, _i]),
    "qsim_fuse_plan": (_i, [_i, _i, _vp, _vp, _i, _vp, _vp, _vp]),
    "qsim_matrix_fused": (_i, [_i, _c_int_p, _i, _vp, _vp, _vp, _vp]),
    "qsim_compile_pass": (_i, [_i, _c_int_p, _i, _vp, _vp, _vp, _i, _i, _vp, _sz, _vp, _vp, _vp]),
    "qsim_apply_bitflips": (_i, [_vp, _i64, _i, _vp, _vp, _vp, _vp, _vp]),
    "qsim_compare_workspace": (_sz, [_i]),
    "qsim_compare_states": (_i, [_vp, _i, _vp, _i, _i, _vp, _vp, _sz, _vp]),
}

_lock = threading.Lock()
_lib = None


class QsimError(RuntimeError):
    """A C-ABI call returned an error code."""


def load(path: str = LIB_PATH):
    """Load the library once (thread-safe) and check its ABI version."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(path):
            raise ImportError(
                f"qibo_amd: native library {path} is missing; build it with "
                "`python -m qibo_amd._build` (hipcc, gfx950)"
            )
        lib = ctypes.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        ver = lib.qsim_version()
        if ver != ABI_VERSION:
            raise ImportError(f"qibo_amd: library ABI {ver} != expected {ABI_VERSION}")
        _lib = lib
        return lib


def check(rc: int, lib=None):
    """Raise on a negative return code with the library's thread-local message."""
    if rc == OK:
        return
    lib = lib or _lib
    msg = lib.qsim_last_error().decode() if lib is not None else ""
    if rc == ENOMEM:
        raise MemoryError(msg)
    if rc == EINVAL:
        raise ValueError(msg)
    raise QsimError(f"qsim error {rc}: {msg}")


def int_array(values):
    """ctypes int array (or None for an empty sequence)."""
    values = [int(v) for v in values]
    if not values:
        return None
    return (ctypes.c_int * len(values))(*values)
