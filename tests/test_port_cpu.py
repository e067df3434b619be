"""The C restatement (oracle/fxport.c) against the FP64 oracle on the same inputs (CPU)."""
import numpy as np
import pytest

import fxoracle as fo
from fxcases import (case_cfg0, case_cfg1, case_cfg2, case_cfg3, case_fourstep_ffts, case_long_ffts, case_mark5b,
                     case_offsets, case_real8, damaged_slot, record_errors, slot, channel_errors)


def _port():
    import fxport
    if not fxport.available():
        pytest.skip("libfxport.so not built")
    return fxport


@pytest.mark.parametrize("name", ["cfg0", "cfg0_lsb", "cfg1", "cfg2", "cfg3", "long", "fourstep"])
def test_port_matches_fp64_oracle(name):
    fxport = _port()
    cfg = {"cfg0": lambda: case_cfg0(B=64), "cfg0_lsb": lambda: case_cfg0(B=32, lsb=True),
           "cfg1": lambda: case_cfg1(B=16, nstations=4, nfreq=4),
           "cfg2": lambda: case_cfg2(B=8, xengine="fp32", nstations=6),
           "cfg3": lambda: case_cfg3(B=16), "long": lambda: case_long_ffts(B=4),
           "fourstep": lambda: case_fourstep_ffts(B=2)}[name]()
    data, ctrl = slot(cfg)
    got, w = fxport.process_subint(cfg, [0, 1, 0], data, ctrl, nthreads=3)
    ref, st = fo.process_subint(cfg, [0, 1, 0], data, ctrl)
    e = record_errors(cfg, fo.result_layout(cfg), got.astype(np.complex128), ref)
    assert max(e["cross"]) < 1e-5 and max(e["auto"]) < 1e-5
    np.testing.assert_array_equal(w, np.array([s.weight for s in st], dtype=np.float32))


def test_channel_errors_of_the_port():
    """The per-channel error measure of the cfg2 GPU test, on the FP32 C restatement: zero
    for the oracle itself, FP32-sized for the port, and it sees a planted 1e-3 error."""
    fxport = _port()
    cfg = case_cfg2(B=8, xengine="fp32", nstations=4)
    lay = fo.result_layout(cfg)
    data, ctrl = slot(cfg)
    got, _ = fxport.process_subint(cfg, [0, 1, 0], data, ctrl, nthreads=2)
    ref, _ = fo.process_subint(cfg, [0, 1, 0], data, ctrl)
    n0, p0 = channel_errors(cfg, lay, ref, ref)
    assert np.max(n0) == 0 and np.max(p0) < 1e-12
    n1, p1 = channel_errors(cfg, lay, got.astype(np.complex128), ref)
    assert np.max(n1) < 1e-6 and np.median(p1) < 1e-2
    bad = ref.copy()
    off = next(iter(lay.baseline_off.values()))
    bad[off: off + 8] *= 1.001
    n2, _ = channel_errors(cfg, lay, bad, ref)
    assert np.max(n2) > 0


def test_port_thread_count_invariance():
    fxport = _port()
    cfg = case_cfg0(B=96)
    data, ctrl = slot(cfg)
    a, _ = fxport.process_subint(cfg, [0, 1, 0], data, ctrl, nthreads=1)
    b, _ = fxport.process_subint(cfg, [0, 1, 0], data, ctrl, nthreads=5)
    rel = np.mean(np.abs(a - b)) / np.mean(np.abs(a))
    assert rel < 1e-6


@pytest.mark.parametrize("order", [0, 2])
@pytest.mark.parametrize("lsb", [False, True])
def test_port_rotation_orders_match_oracle(order, lsb):
    """Orders 0 (post-F, real FFT with the LSB conj-flip) and 2 (quadratic rotator built by
    repeated FP32 products) in the C restatement against the FP64 oracle."""
    fxport = _port()
    cfg = case_cfg0(B=32, lsb=lsb, fringe_rot_order=order)
    data, ctrl = slot(cfg)
    got, _ = fxport.process_subint(cfg, [0, 1, 0], data, ctrl, nthreads=2)
    ref, _ = fo.process_subint(cfg, [0, 1, 0], data, ctrl)
    e = record_errors(cfg, fo.result_layout(cfg), got.astype(np.complex128), ref)
    assert max(e["cross"]) < 1e-5 and max(e["auto"]) < 1e-5


@pytest.mark.parametrize("cs,lsb", [(1, True), (2, False), (2, True)])
def test_port_complex_sidebands_match_oracle(cs, lsb):
    fxport = _port()
    cfg = case_cfg2(B=8, xengine="fp32", nstations=4, complex_sampling=cs, lsb=lsb)
    data, ctrl = slot(cfg)
    got, _ = fxport.process_subint(cfg, [0, 1, 0], data, ctrl, nthreads=2)
    ref, _ = fo.process_subint(cfg, [0, 1, 0], data, ctrl)
    e = record_errors(cfg, fo.result_layout(cfg), got.astype(np.complex128), ref)
    assert max(e["cross"]) < 1e-5 and max(e["auto"]) < 1e-5


@pytest.mark.parametrize("name", ["offsets", "mark5b", "real8", "damaged"])
def test_port_more_cases_match_oracle(name):
    """LO / clock offsets with channel averaging, Mark5B, 8-bit real data, and a slot with
    flagged FFTs, an invalid frame and a fill-pattern frame."""
    fxport = _port()
    if name == "offsets":
        cfg = case_offsets(B=16)
    elif name == "mark5b":
        cfg = case_mark5b(B=16)
    elif name == "real8":
        cfg = case_real8(B=16)
    else:
        cfg = case_cfg1(B=32, nstations=3, nfreq=2)
    data, ctrl = damaged_slot(cfg) if name == "damaged" else slot(cfg)
    got, w = fxport.process_subint(cfg, [0, 1, 0], data, ctrl, nthreads=2)
    ref, st = fo.process_subint(cfg, [0, 1, 0], data, ctrl)
    e = record_errors(cfg, fo.result_layout(cfg), got.astype(np.complex128), ref)
    assert max(e["cross"]) < 1e-5 and max(e["auto"]) < 1e-5
    np.testing.assert_array_equal(w, np.array([s.weight for s in st], dtype=np.float32))
    if name == "damaged":
        assert (w[0] == 0).sum() >= 4 and (w[1] < 1).sum() >= 1
