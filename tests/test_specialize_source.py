"""Run-time specialised blocked passes, host side (no GPU): the kernel source
generated from a program's record sequence, and its hiprtc compile for
gfx950, for programs that hold every record kind in both precisions (the
fast forms and the ordered per-gate forms), from |0...0> and not. A
conditional mapping change is refused. The GPU side (bitwise equality with
the interpreted pass) is test_gpu_specialize.py."""

import ctypes
import struct

import numpy as np
import pytest

from qibo_amd import _lib

from tests.specialize_programs import programs, record_ops


@pytest.fixture(scope="module")
def lib():
    return _lib.load()


def _source(lib, prog, count, code, from_zero):
    n = lib.qsim_blocked_specialized_source(prog, len(prog), count, code, from_zero, None, 0)
    assert n > 0, lib.qsim_last_error()
    buf = ctypes.create_string_buffer(n)
    assert lib.qsim_blocked_specialized_source(prog, len(prog), count, code, from_zero, buf, n) == n
    return buf.value.decode()


@pytest.mark.parametrize("cdtype,code", [(np.complex128, _lib.C128), (np.complex64, _lib.C64)])
def test_source_lists_every_record(lib, cdtype, code):
    for _, prog, count in programs(14, cdtype, seed=2, depth=2):
        src = _source(lib, prog, count, code, 0)
        recs = record_ops(prog, cdtype)
        # one block per record, in order, at the record's offset
        assert src.count("const Hdr h = load_hdr(prog + ") == len(recs)
        assert "switch" not in src
        assert ("cplx<double>" in src) == (code == _lib.C128)
        zsrc = _source(lib, prog, count, code, 1)
        assert "load_lane(a, psi" not in zsrc and "load_lane(a, psi" in src


@pytest.mark.parametrize("cdtype,code,ordered", [(np.complex128, _lib.C128, False), (np.complex64, _lib.C64, False),
                                                 (np.complex128, _lib.C128, True)])
def test_sources_compile_for_gfx950(lib, cdtype, code, ordered):
    """hiprtc compiles every pass of the mixed circuit (all record kinds)."""
    kinds = set()
    progs = programs(14, cdtype, seed=3, depth=2, ordered=ordered)
    for _, prog, count in progs[:3]:
        kinds |= {op for op, _, _ in record_ops(prog, cdtype)}
        size = ctypes.c_size_t(0)
        rc = lib.qsim_blocked_compile_check(prog, len(prog), count, code, 0, ctypes.byref(size))
        assert rc == 0, lib.qsim_last_error().decode()[-2000:]
        assert size.value > 1000
    want = {0, 1, 2, 3} if ordered else {1, 2, 3, 4, 5, 7, 10, 11, 12}
    assert want <= kinds


def test_conditional_mapping_change_is_refused(lib):
    _, prog, count = programs(14, np.complex128, seed=4, depth=1)[0]
    prog = bytearray(prog)
    off = 0
    for op, _, _ in record_ops(bytes(prog), np.complex128):
        if op == 2:  # MAP: give it a tile condition
            struct.pack_into("<Q", prog, off + 24, 1 << 13)
            break
        (nent,) = struct.unpack_from("<I", prog, off + 4)
        off += 32 + ((min(nent, 64) * 16 + 15) & ~15)
    else:
        pytest.skip("no mapping change in the first pass")
    prog = bytes(prog)
    assert lib.qsim_blocked_specialized_source(prog, len(prog), count, _lib.C128, 0, None, 0) == 0
    assert b"conditional mapping change" in lib.qsim_last_error()
    assert lib.qsim_blocked_compile_check(prog, len(prog), count, _lib.C128, 0, None) == _lib.EINVAL


def test_adjoint_sweep_sources_compile(lib):
    """The adjoint sweep's kernels (both states, INNER records at fixed slots)
    for the forward and reverse passes of a hardware-efficient ansatz."""
    from qibo_amd import models
    from qibo_amd.adjoint import SweepTemplate
    from qibo_amd.backend import GatePlanner

    tmpl = SweepTemplate(GatePlanner("complex64"), models.hardware_efficient_ansatz(13, 1), np.complex64)
    p = tmpl.rev[0]
    prog = p.row.tobytes()
    ops = {op for op, _, _ in record_ops(prog, np.complex64)}
    assert {8, 9} <= ops  # INNER1 and INNER_DIAG
    src = _source(lib, prog, p.count, _lib.C64, _lib.SPEC_ADJOINT)
    assert src.count("wave_sum(acc)") == len(p.slots)
    size = ctypes.c_size_t(0)
    assert lib.qsim_blocked_compile_check(prog, len(prog), p.count, _lib.C64, _lib.SPEC_ADJOINT,
                                          ctypes.byref(size)) == 0, lib.qsim_last_error().decode()[-2000:]
    assert lib.qsim_blocked_specialized_source(prog, len(prog), p.count, _lib.C64, 3, None, 0) == 0
