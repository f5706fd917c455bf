"""The C ABI without a GPU: the library loads, exports every entry point that
include/qtx.h declares, and validates arguments before touching a device."""
import ctypes

import pytest

from quicktex_amd.capi import BC1Params
from tests._qtx import LIB_PATH, Qtx, declared_symbols


@pytest.fixture(scope='module')
def lib():
    return Qtx().lib


def test_exports_every_declared_symbol(lib):
    syms = declared_symbols()
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_params_init_matches_oracle(lib, oracle):
    for interp in range(4):
        for mode in (4, 3, 0x13):
            for level in range(20):
                p = BC1Params()
                assert lib.qtx_bc1_params_init(ctypes.byref(p), level, mode, interp) == 0
                assert p.as_tuple() == oracle.params(level, mode, interp).as_tuple(), (level, mode, interp)


def test_params_errors(lib):
    lib.qtx_last_error.restype = ctypes.c_char_p
    p = BC1Params()
    assert lib.qtx_bc1_params_init(ctypes.byref(p), 20, 4, 0) == 1  # QTX_EINVAL
    assert b'Level out of range' in lib.qtx_last_error()
    assert lib.qtx_bc1_params_init(ctypes.byref(p), 5, 7, 0) == 1
    assert b'color mode' in lib.qtx_last_error()
    assert lib.qtx_bc1_params_init(ctypes.byref(p), 5, 4, 9) == 1
    # clamps of SetOrderings / SetPowerIterations (BC1Encoder.cpp:251-258)
    assert lib.qtx_bc1_params_init(ctypes.byref(p), 5, 4, 0) == 0
    p.orderings4, p.orderings3, p.power_iterations = 0, 1000, 99
    assert lib.qtx_bc1_params_clamp(ctypes.byref(p)) == 0
    assert (p.orderings4, p.orderings3, p.power_iterations) == (1, 32, 10)
    p.power_iterations = 1
    assert lib.qtx_bc1_params_clamp(ctypes.byref(p)) == 0
    assert p.power_iterations == 4


def test_entry_points_reject_bad_arguments(lib):
    buf = (ctypes.c_uint8 * 64)()
    p = BC1Params()
    lib.qtx_bc1_params_init(ctypes.byref(p), 5, 4, 0)
    u8 = ctypes.cast(buf, ctypes.POINTER(ctypes.c_uint8))
    assert lib.qtx_bc1_encode(None, ctypes.byref(p), u8, 4, 4, u8, 8) == 1  # null context
    assert lib.qtx_bc4_encode(None, 0, u8, 0, 4, u8, 8) == 1  # width 0
    lib.qtx_last_error.restype = ctypes.c_char_p
    assert b'width must be greater than 0' in lib.qtx_last_error()
