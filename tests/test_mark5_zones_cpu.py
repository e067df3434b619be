"""Mark5B fill-pattern zones (blanker_mark5.c:43-172) in the oracle and the C port (CPU).

The expected byte ranges below are worked by hand from the blanker's halving search over the
1250 64-bit words of a 10000-byte Mark5B payload: probes start at word 625 and move by
312, 156, 78, 39, 19, 9, 4, 2, 1; findfirstinvalid keeps the lowest probe holding the fill
pattern (1249 if none) and subtracts 4 bytes when the 32-bit word just before it is fill;
findfirstvalid keeps the lowest probe not holding it and adds 4 bytes when the 32-bit word two
before it (the low half of the previous 64-bit word) is fill.
"""
from types import SimpleNamespace

import numpy as np
import pytest

import fxoracle as fo
from fxcases import case_mark5b, mark5_fill, mark5b_partial_slot, record_errors

ALL = 1 << 30


def _payload(seed=1):
    rng = np.random.default_rng(seed)
    return rng.integers(0, 256, 10000, dtype=np.uint8)


@pytest.mark.parametrize("words,half,expect", [
    # no fill; fill only between the end words: the whole payload decodes
    ([], -1, (0, ALL)),
    (range(100, 300), -1, (0, ALL)),
    # both end words fill: nothing decodes
    (list(range(0, 3)) + list(range(1240, 1250)), -1, (ALL, 0)),
    # tail from 625: the first probe hits, every later probe (313, 469, ..., 620) is valid
    (range(625, 1250), -1, (0, 5000)),
    # tail from 313: probes 625 and 313 hit, 157 ... 309 are valid
    (range(313, 1250), -1, (0, 2504)),
    # all but word 0: every probe hits down to 6 (the last step moves to 5 without a test),
    # 48 bytes, and the upper half of word 5 is fill: 44
    (range(1, 1250), -1, (0, 44)),
    # tail from 625 with the upper half of word 624 filled: 5000 - 4
    (range(625, 1250), 624, (0, 4996)),
    # head to word 624: probe 625 is valid, every later probe (313 ... 621) is fill; the low
    # half of word 624 is fill: 5000 + 4
    (range(0, 625), -1, (5004, ALL)),
    # word 0 alone: every probe is valid down to 6; the low half of word 5 is data: 48
    (range(0, 1), -1, (48, ALL)),
])
def test_zone_hand_worked(words, half, expect):
    p = _payload()
    mark5_fill(p, words, half)
    assert fo.mark5_zone(p) == expect


def test_zone_low_half_quirk():
    """Head fill to word 623 and word 624 = (fill, data): word 624 is not a fill word, probe
    625 is the lowest valid probe, and the reference's straggler test reads the low half of
    word 624 (fill): 5004.  A test of the high half (data) would give 5000."""
    p = _payload(2)
    mark5_fill(p, range(0, 624))
    p[8 * 624: 8 * 624 + 4] = np.frombuffer(np.uint32(0x11223344).tobytes(), np.uint8)
    assert fo.mark5_zone(p) == (5004, ALL)


@pytest.mark.parametrize("nbands,words,expect", [
    # 8 bands x 2 bits = 2 bytes per time sample: byte 2n < 5000
    (8, range(625, 1250), lambda n: n < 2500),
    # 1 band: byte n/4 < 5000
    (1, range(625, 1250), lambda n: n < 20000),
    # 16 bands: 4 bytes per sample, byte 4n >= 5004
    (16, range(0, 625), lambda n: n >= 1251),
])
def test_sample_validity_in_zone(nbands, words, expect):
    """A time sample decodes when the byte where its fields begin lies in the zone
    (mark5_format_mark5b.c:241-243, 1295-1343, 1868-1932); frames on either side decode whole
    and a frame with the invalid flag decodes nothing."""
    ds = SimpleNamespace(format="MARK5B", framebytes=10016, headerbytes=16, nbits=2, complex_sampling=0,
                         recbands=[0] * nbands)
    spf = fo.samples_per_frame(ds)
    rng = np.random.default_rng(3)
    buf = rng.integers(0, 256, 4 * 10016, dtype=np.uint8)
    for f in range(4):
        buf[f * 10016 + 5] &= 0x7F
    mark5_fill(buf[10016 + 16: 2 * 10016], words)
    buf[3 * 10016 + 5] |= 0x80
    ok = fo.sample_valid(buf, ds)
    n = np.arange(spf)
    assert ok[:spf].all() and ok[2 * spf: 3 * spf].all() and not ok[3 * spf:].any()
    np.testing.assert_array_equal(ok[spf: 2 * spf], expect(n))
    vals, valid = fo.decode_stream(buf, ds)
    np.testing.assert_array_equal(valid, ok)
    assert np.all(vals[:, ~ok] == 0)


def test_vdif_fill_blanks_whole_frames():
    """VDIF keeps blanker_vdif's rule: fill at either end blanks the whole frame, fill between
    the ends blanks nothing."""
    ds = SimpleNamespace(format="VDIF", framebytes=8032, headerbytes=32, nbits=2, complex_sampling=0,
                         recbands=[0])
    buf = np.zeros(3 * 8032, np.uint8)
    rng = np.random.default_rng(4)
    for f in range(3):
        buf[f * 8032 + 32: (f + 1) * 8032] = rng.integers(0, 256, 8000, dtype=np.uint8)
        buf[f * 8032 + 8: f * 8032 + 12] = [1, 0, 0, 0]            # word 2 nonzero
        buf[f * 8032 + 3] &= 0x7F
    mark5_fill(buf[32: 8032], range(990, 1000))                    # tail fill: frame blanked
    mark5_fill(buf[8032 + 32: 2 * 8032], range(10, 20))            # middle fill: frame decodes
    ok = fo.sample_valid(buf, ds)
    spf = fo.samples_per_frame(ds)
    assert not ok[:spf].any() and ok[spf:].all()


def test_port_matches_oracle_on_zones():
    """The C restatement's zones (its own search code) against the oracle's on a Mark5B slot
    with fill-pattern zones: equal per-FFT weights and FP32-sized visibility differences; the
    weights differ from what whole-frame blanking would give."""
    import fxport
    if not fxport.available():
        pytest.skip("libfxport.so not built")
    cfg = case_mark5b(B=32)
    data, ctrl = mark5b_partial_slot(cfg)
    got, w = fxport.process_subint(cfg, [0, 1, 0], data, ctrl, nthreads=3)
    ref, st = fo.process_subint(cfg, [0, 1, 0], data, ctrl)
    e = record_errors(cfg, fo.result_layout(cfg), got.astype(np.complex128), ref)
    assert max(e["cross"]) < 1e-5 and max(e["auto"]) < 1e-5
    np.testing.assert_array_equal(w, np.array([s.weight for s in st], dtype=np.float32))
    # frame-level blanking would leave only weights made of whole frames
    ds = cfg.datastreams[1]
    spf = fo.samples_per_frame(ds)
    ok = fo.sample_valid(data[1], ds)
    per_frame = ok[: (len(ok) // spf) * spf].reshape(-1, spf).sum(axis=1)
    assert np.any((per_frame > 0) & (per_frame < spf))
