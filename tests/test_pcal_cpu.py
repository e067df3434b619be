"""Pins of the phase-cal restatement (oracle/pcal.py): the tone table of Configuration
(configuration.cpp:1765-1835) worked by hand, the extractors' tone definition (pcal.cpp) on
planted tones whose expected values follow in closed form, the FFT selection against the FP64
F-engine restatement's data weights, and the period folding the device path uses."""
import math

import numpy as np
import pytest

import fxoracle as fo
import pcal as pc
from difx_amd import config as cf
from fxcases import case_cfg0, case_real8, damaged_slot, pcal_tone_slot, slot


@pytest.mark.parametrize("lo,bw,lsb,interval,base,want", [
    # USB: 8400.25 MHz edge, 5 MHz comb: tones 8405, 8410, 8415 MHz; offset 8405 - 8400.25 MHz
    (8400.25, 16.0, False, 5_000_000, 0, (3, 4_750_000, [8405e6, 8410e6, 8415e6])),
    # LSB: tones 8400, 8395, 8390, 8385 MHz down to 8384.25; offset 8400.25 - 8400 MHz
    (8400.25, 16.0, True, 5_000_000, 0, (4, 250_000, [8400e6, 8395e6, 8390e6, 8385e6])),
    # a tone on the band edge itself is skipped (USB: strictly above the edge), one on the far
    # edge is kept: 8401 ... 8416 MHz
    (8400.0, 16.0, False, 1_000_000, 0, (16, 1_000_000, [8401e6 + k * 1e6 for k in range(16)])),
    (8400.0, 16.0, True, 1_000_000, 0, (16, 1_000_000, [8399e6 - k * 1e6 for k in range(16)])),
    # no tone in a 0.5 MHz band: one placeholder tone, offset flagged -1
    (8400.25, 0.5, False, 5_000_000, 0, (1, -1, [0.0])),
    # a 1 MHz base moves the comb (8401 ... 8416 MHz) but not the offset, which turns negative
    (8400.25, 16.0, False, 5_000_000, 1_000_000, (4, -250_000, [8401e6, 8406e6, 8411e6, 8416e6])),
])
def test_tone_table_by_hand(lo, bw, lsb, interval, base, want):
    n, off, sky = pc.tone_table(lo, bw, lsb, interval, base)
    assert (n, off) == want[:2]
    assert sky == want[2]


def test_baseband_frequencies_and_period():
    """A zero offset drops the DC bin (tones at 1, 2, 3 intervals); the tone period is the
    shortest M with f M / fs whole for every tone, checked by brute force."""
    assert pc.baseband_freqs(0, 1_000_000, 3) == [1_000_000, 2_000_000, 3_000_000]
    assert pc.baseband_freqs(250_000, 5_000_000, 3) == [250_000, 5_250_000, 10_250_000]
    for fs, off, iv in [(32_000_000, 4_750_000, 5_000_000), (64_000_000, 1_000_000, 1_000_000),
                        (32_000_000, 10_000, 1_000_000), (16_000_000, 0, 1_000_000)]:
        M = pc.tone_period(fs, off, iv)
        fk = pc.baseband_freqs(off, iv, 4)
        brute = next(p for p in range(1, fs + 1) if all((f * p) % fs == 0 for f in fk + [iv]))
        assert M == brute
    assert pc.tone_period(32_000_000, 4_750_000, 5_000_000) == 128


def test_folding_by_period_equals_the_direct_sum():
    """Summing samples into M bins by sample number mod M and taking an M-point DFT at the
    tone bins (what the device path does) equals the direct sum over samples."""
    rng = np.random.default_rng(3)
    fs, off, iv, nt = 32_000_000, 4_750_000, 5_000_000, 3
    M = pc.tone_period(fs, off, iv)
    fk = np.array(pc.baseband_freqs(off, iv, nt))
    x = rng.standard_normal(5000)
    m = 123_456_789 + np.arange(5000) + np.repeat(np.arange(10) * 17, 500)   # gapped runs
    direct = np.array([(x * np.exp(-2j * np.pi * ((f * (m % fs)) % fs) / fs)).sum() for f in fk])
    bins = np.zeros(M)
    np.add.at(bins, m % M, x)
    j = np.arange(M)
    folded = np.array([(bins * np.exp(-2j * np.pi * ((f * j) % fs) / fs)).sum() for f in fk])
    np.testing.assert_allclose(folded, direct, rtol=1e-10, atol=1e-9)


@pytest.mark.parametrize("make", [lambda: case_real8(B=32), lambda: case_cfg0(B=32)])
def test_fft_selection_matches_the_fengine_weights(make):
    """The FFTs that feed the extractor are those the FP64 F-engine restatement gives a
    nonzero data weight, on a slot with flagged FFTs and invalid and fill-pattern frames."""
    cfg = make()
    offsets = [0, 1, 0]
    data, ctrl = damaged_slot(cfg, offsets)
    for d in range(2):
        st = fo.process_station(cfg, d, offsets, data[d], ctrl[d])
        got, _ = pc.valid_ffts(cfg, d, offsets, data[d], ctrl[d])
        assert [i for i, _ in got] == [int(i) for i in np.nonzero(st.weight > 0)[0]]
        assert len(got) < cfg.blockspersend()


def tone_case(lsb=False, B=32):
    """One station pair, 8-bit real, one 16 MHz band at 1400 MHz, 256 channels (L = 512)."""
    return cf.standard_config(2, 1, [1400.0], 16.0, 256, subint_ns=B * 16000, nbits=8, lsb=lsb,
                              max_delay_us=50.0, max_rate_us_per_s=0.01, seed=21)


@pytest.mark.parametrize("lsb", [False, True])
def test_planted_tones_closed_form(lsb):
    """Two planted tones, at 8 MHz (fs/4) and 1 MHz (fs/32), on a 1 MHz comb with offset
    1 MHz: the quantised signal repeats every P = 32 samples and every FFT covers whole periods
    (L = 512), so tone k = (N / P) sum_{m<P} Q(x_m) exp(-2 pi i f_k m / fs) with N the samples
    summed (B FFTs of L); before quantisation that is N A_j / 2 exp(i phi_j) at the planted
    tones and 0 elsewhere."""
    cfg = tone_case(lsb)
    B, L, fs, P = cfg.blockspersend(), cfg.fftchannels(0), 32_000_000, 32
    planted = [(8_000_000, 30.0, 0.7), (1_000_000, 12.0, -2.1)]
    offsets = [0, 1, 0]
    data, ctrl, decoded = pcal_tone_slot(cfg, planted, offsets)
    n, off, _ = pc.tone_table(1400.0, 16.0, lsb, 1_000_000)
    assert (n, off) == (16, 1_000_000)
    fk = np.array(pc.baseband_freqs(off, 1_000_000, n))
    for d in range(2):
        ffts, _ = pc.valid_ffts(cfg, d, offsets, data[d], ctrl[d])
        assert len(ffts) == B
        got = pc.extract(cfg, d, offsets, data[d], ctrl[d], 1_000_000)[0]
        mp = np.arange(P)
        q = decoded[d](mp).astype(np.float32).astype(np.float64)
        want = np.array([(B * L / P) * (q * np.exp(-2j * np.pi * ((f * mp) % fs) / fs)).sum() for f in fk])
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-9 * B * L * 30.0)
        ideal = np.zeros(n, complex)
        for f, a, ph in planted:
            ideal[int(f // 1_000_000) - 1] = B * L * a / 2 * np.exp(1j * ph)
        assert np.max(np.abs(got - ideal)) < 0.01 * B * L * 12.0 / 2


def test_sample_numbering_moves_the_phase():
    """The tone phase follows the sample number since the scan start: the same signal read a
    subint later gives the same phases (the signal is a function of absolute time)."""
    cfg = tone_case()
    planted = [(3_000_000, 20.0, 0.3)]
    got = []
    for offsets in ([0, 1, 0], [0, 1, 512000]):
        data, ctrl, _ = pcal_tone_slot(cfg, planted, offsets)
        got.append(pc.extract(cfg, 0, offsets, data[0], ctrl[0], 1_000_000)[0][2])
    for g in got:
        assert abs(np.angle(g) - 0.3) < 0.01
    assert math.isclose(abs(got[0]), abs(got[1]), rel_tol=0.01)


def test_flagged_offset_gives_zero_tones():
    """A base that turns the offset negative selects the dummy extractor: zero tones.  With
    a 6 MHz base the comb's first tone above 1400 MHz is 1401 MHz = 279 x 5 + 6, and the offset
    279 x 5 MHz - 1400 MHz = -5 MHz."""
    cfg = tone_case()
    data, ctrl = slot(cfg)
    out = pc.extract(cfg, 0, [0, 1, 0], data[0], ctrl[0], 5_000_000, 6_000_000)[0]
    n, off, _ = pc.tone_table(1400.0, 16.0, False, 5_000_000, 6_000_000)
    assert (n, off) == (4, -5_000_000)
    assert out.shape == (n,) and not np.any(out)


def _lib():
    from pathlib import Path
    if not (Path(__file__).resolve().parent.parent / "difx_amd" / "libfxg.so").exists():
        pytest.skip("libfxg.so not built")
    from difx_amd.core import load_library
    return load_library()


@pytest.mark.parametrize("lsb", [False, True])
@pytest.mark.parametrize("interval,base", [(1_000_000, 0), (5_000_000, 0), (500_000, 0),
                                           (5_000_000, 1_000_000), (5_000_000, 6_000_000),
                                           (10_000_000, 0)])
def test_library_tone_table_equals_the_restatement(lsb, interval, base):
    """fxg_set_pcal / fxg_pcal_info on a host-only context: the library's tone count, offset
    and tone period of every recorded band equal the restatement's, over band edges on whole,
    quarter, hundredth and fractional MHz, and 8, 16 and 32 MHz bands."""
    import ctypes as C
    lib = _lib()
    for freqs, bw in [([8400.25, 8400.01, 4800.37, 1400.0], 16.0), ([4800.0, 4832.5], 32.0),
                      ([1400.0, 1408.0], 8.0)]:
        cfg = cf.standard_config(2, 1, freqs, bw, 256, subint_ns=8 * (256 * 1000 // int(bw)),
                                 nbits=2, lsb=lsb, seed=5)
        fa, keep = cfg.to_abi()
        h = C.c_void_p()
        assert lib.fxg_create_layout(C.byref(fa), C.byref(h)) == 0
        try:
            fs = int(2e6 * bw)
            want = [pc.tone_table(f, bw, lsb, interval, base)[:2] for f in freqs]
            periods = [pc.tone_period(fs, off, interval) if off >= 0 else 0 for _, off in want]
            rc = lib.fxg_set_pcal(h, 0, interval, base)
            if max(periods) > 16384:                   # rejected: the device folds in LDS
                assert rc != 0
                continue
            assert rc == 0
            for b, ((n, off), M) in enumerate(zip(want, periods)):
                nt, of, per = C.c_int32(), C.c_int64(), C.c_int32()
                assert lib.fxg_pcal_info(h, 0, b, C.byref(nt), C.byref(of), C.byref(per)) == 0
                assert (nt.value, of.value, per.value) == (n, off, M), (freqs[b], bw)
            # phase-cal stays off on the other datastream
            nt, of, per = C.c_int32(), C.c_int64(), C.c_int32()
            assert lib.fxg_pcal_info(h, 1, 0, C.byref(nt), C.byref(of), C.byref(per)) != 0
        finally:
            lib.fxg_destroy(h)


def test_library_rejects_long_tone_periods():
    """An 8400.0123 MHz band edge with a 1 MHz comb puts the first tone 987.7 kHz in: the tone
    period is 32 MHz / gcd(32 MHz, 987.7 kHz, 1 MHz) = 320000 samples, beyond the 16384 the
    device folds in LDS, so fxg_set_pcal refuses it and keeps the previous setting (off)."""
    import ctypes as C
    lib = _lib()
    n, off, _ = pc.tone_table(8400.0123, 16.0, False, 1_000_000)
    assert pc.tone_period(32_000_000, off, 1_000_000) == 320000
    cfg = cf.standard_config(2, 1, [8400.0123], 16.0, 256, subint_ns=8 * 16000, nbits=2, seed=5)
    fa, keep = cfg.to_abi()
    h = C.c_void_p()
    assert lib.fxg_create_layout(C.byref(fa), C.byref(h)) == 0
    try:
        assert lib.fxg_set_pcal(h, 0, 1_000_000, 0) == -1               # FXG_E_INVALID
        nt, of, per = C.c_int32(), C.c_int64(), C.c_int32()
        assert lib.fxg_pcal_info(h, 0, 0, C.byref(nt), C.byref(of), C.byref(per)) == -4   # off
    finally:
        lib.fxg_destroy(h)
