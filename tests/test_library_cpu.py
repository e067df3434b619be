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
    declared = set(re.findall(r"^\s*(?:int|int64_t|const char\*)\s+(fxg_\w+)\s*\(", hdr, re.M))
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
    # 4 int32, threadmap[32], int64, 2 int32, int64
    assert ctypes.sizeof(abi.fxg_vdif_mux_params) == 16 + 4 * abi.FXG_MUX_MAX_THREADS + 24


def test_create_rejects_null_config():
    lib = _lib()
    h = ctypes.c_void_p()
    lib.fxg_create.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]
    assert lib.fxg_create(None, ctypes.byref(h)) == abi.FXG_E_INVALID


# ------------------------------------------------ sizing and validation without a device
def test_footprint_without_device():
    """fxg_footprint validates and sizes a configuration with no device present; the result
    length agrees with the oracle's populateResultLengths restatement."""
    _lib()
    import fxoracle as fo
    from difx_amd.core import footprint
    from fxcases import case_cfg1, case_cfg3
    for cfg in (case_cfg1(B=32, nstations=4, nfreq=2), case_cfg3(B=32, xengine="fp16")):
        dev, crl, slot = footprint(cfg)
        assert crl == fo.result_layout(cfg).coreresultlength
        assert dev > 0 and slot > 0
    a, _, _ = footprint(case_cfg1(B=64, nstations=4, nfreq=2, chunk_ffts=16))
    b, _, _ = footprint(case_cfg1(B=64, nstations=4, nfreq=2, chunk_ffts=64))
    assert b > a                                   # spectra chunk buffers scale with the chunk


@pytest.mark.parametrize("mutate,ok", [
    (lambda c: setattr(c, "fringe_rot_order", 3), False),
    (lambda c: setattr(c, "fringe_rot_order", 0), True),
    (lambda c: setattr(c, "fringe_rot_order", 2), True),
    (lambda c: setattr(c, "subint_ns", c.subint_ns + 1000), False),     # not a whole FFT count
    (lambda c: (setattr(c.freqs[0], "nchan", 16384), setattr(c, "subint_ns", 2 * 512000)), True),   # L 32768
    (lambda c: (setattr(c.freqs[0], "nchan", 32768), setattr(c, "subint_ns", 2 * 1024000)), False),  # L > 32768
    (lambda c: setattr(c.freqs[0], "chanavg", 3), False),                # does not divide nchan
    # table indices out of range are refused before any lookup through them
    (lambda c: setattr(c.datastreams[0], "recfreqs", [99]), False),
    (lambda c: setattr(c.datastreams[1], "recfreqs", [-1]), False),
    (lambda c: setattr(c.datastreams[0], "recbands", [(5, "R")]), False),
    (lambda c: (setattr(c.datastreams[0], "zoomfreqs", [(99, 0, 0)]),
                setattr(c.datastreams[0], "zoombands", [(0, "R")])), False),
    (lambda c: (setattr(c.datastreams[0], "zoomfreqs", [(0, 0, 0)]),
                setattr(c.datastreams[0], "zoombands", [(3, "R")])), False),
    (lambda c: setattr(c.datastreams[0], "phase_offset_deg", [float("nan")]), False),
    (lambda c: setattr(c.datastreams[0], "clock_offset_delta_us", [float("inf")]), False),
    (lambda c: setattr(c.datastreams[0], "phase_offset_deg", [30.0]), True),
])
def test_validation(mutate, ok):
    _lib()
    from difx_amd.core import FxgError, footprint
    from fxcases import case_cfg0
    cfg = case_cfg0(B=32)
    mutate(cfg)
    if ok:
        footprint(cfg)
    else:
        with pytest.raises(FxgError):
            footprint(cfg)


def test_create_failure_message_without_a_context():
    """A failed create leaves no context; fxg_last_error(NULL) gives its message."""
    import ctypes as C
    from difx_amd import fxg_abi as abi
    from difx_amd.core import load_library
    from fxcases import case_cfg0
    lib = load_library()
    cfg = case_cfg0(B=8)
    fc, keep = cfg.to_abi()
    fc.fringe_rot_order = 7
    h = C.c_void_p()
    assert lib.fxg_create_layout(C.byref(fc), C.byref(h)) == abi.FXG_E_INVALID
    assert not h.value
    assert b"fringe rotation order" in lib.fxg_last_error(None)
    fc.fringe_rot_order = 1
    assert lib.fxg_create_layout(C.byref(fc), C.byref(h)) == abi.FXG_OK
    assert lib.fxg_last_error(None) == b"null context"
    lib.fxg_destroy(h)
