"""CPU checks of the C-ABI library: it loads and exports every symbol include/fxg.h declares,
and the ctypes structure sizes agree with the header."""
import ctypes
import re
from pathlib import Path

import pytest

from difx_amd import fxg_abi as abi

ROOT = Path(__file__).resolve().parent.parent


def _lib():
    p = ROOT / "difx_amd" / "libfxg.so"
    if not p.exists():
        pytest.skip("libfxg.so not built")
    return ctypes.CDLL(str(p))


def test_exports_match_header():
    lib = _lib()
    hdr = (ROOT / "include" / "fxg.h").read_text()
    declared = set(re.findall(r"^\s*(?:int|const char\*)\s+(fxg_\w+)\s*\(", hdr, re.M))
    assert declared == set(abi.EXPORTED_SYMBOLS), declared ^ set(abi.EXPORTED_SYMBOLS)
    for name in declared:
        assert hasattr(lib, name), name


def test_strerror_no_device_needed():
    lib = _lib()
    lib.fxg_strerror.restype = ctypes.c_char_p
    assert lib.fxg_strerror(0) == b"ok"
    assert b"invalid" in lib.fxg_strerror(-1)


def test_struct_sizes():
    # fxg_freq: 2 doubles + 4 int32
    assert ctypes.sizeof(abi.fxg_freq) == 32
    assert ctypes.sizeof(abi.fxg_baseline) == 4 * (3 + 2 * 32 + 2 * 32 * 4)
    assert ctypes.sizeof(abi.fxg_layout) == 24


def test_create_rejects_null_config():
    lib = _lib()
    h = ctypes.c_void_p()
    lib.fxg_create.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]
    assert lib.fxg_create(None, ctypes.byref(h)) == abi.FXG_E_INVALID
