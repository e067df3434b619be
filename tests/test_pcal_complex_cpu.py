"""Pins of the complex-sampled phase-cal restatement (oracle/pcal.py, extractor_tones) and of
the results' phase-cal block, CPU only.

- The reference's two complex extractors, restated step by step from pcal.cpp (folding into
  a 2N buffer from the sample offset, the mixing rotator of the shifting extractor, the N-point
  DFT, the double-sideband reversal, the bin read-out with its early stop), give the same tones
  as the restatement's direct sums at its rational frequencies, for single and double sideband,
  with and without the doubling of N to the 1 MHz resolution.
- Complex tones planted at the comb frequencies come back as N A exp(i phi) (N the samples
  summed); for double-sideband data the tone a bin reads sits at -(b + 1) fs / N.
- The phase-cal block of the results (configuration.cpp:2611-2619) counted by hand, and the
  library's layout equal to the restatement's.
- The block's writedata scaling (visibility.cpp:650-675) in closed form.
"""
import math

import numpy as np
import pytest

import fxoracle as fo
import pcal as pc
import visout as vo
from difx_amd import config as cf
from fxcases import pcal_complex_tone_slot, slot


def literal_complex_extractor(fs, offset, interval, ntones, dsb, chunks):
    """PCalExtractorImplicitShiftComplex / PCalExtractorShiftingComplex as written
    (pcal.cpp:808-1020, 1419-1590): chunks = [(sample offset since the scan start, samples)]."""
    no = fs // math.gcd(fs, offset)
    np_ = fs // math.gcd(fs, interval)
    implicit = no % np_ == 0
    n = no if implicit else np_
    while fs // (2 * n) > 1_000_000:
        n *= 2
    acc = np.zeros(2 * n, complex)
    if not implicit:
        rlen = fs // math.gcd(fs, offset)
        rlen = n * rlen // math.gcd(rlen, n)
        rot = np.exp(1j * (-2 * np.pi * offset / fs) * np.arange(2 * rlen))
    for so, x in chunks:
        pidx = so % n
        if implicit:
            tail = len(x) % n
            end = len(x) - tail
            for k in range(0, end, n):
                acc[pidx:pidx + n] += x[k:k + n]
            if tail:
                acc[pidx:pidx + tail] += x[end:]
        else:
            ridx = so % rlen
            tail = len(x) % rlen
            end = len(x) - tail
            for k in range(0, end, rlen):
                rotated = x[k:k + rlen] * rot[ridx:ridx + rlen]
                for p in range(rlen // n):
                    acc[pidx:pidx + n] += rotated[p * n:(p + 1) * n]
            if tail:
                rotated = x[end:] * rot[ridx:ridx + tail]
                t2 = tail % n
                for p in range((tail - t2) // n):
                    acc[pidx:pidx + n] += rotated[p * n:(p + 1) * n]
                acc[pidx:pidx + t2] += rotated[tail - t2:]
    folded = acc[:n] + acc[n:]
    dft = np.fft.fft(folded)                       # forward, unnormalised
    if dsb:
        dft = dft[::-1].copy()
    step = int(math.floor(n * interval / fs))
    obin = int(math.floor(n * offset / fs)) if implicit else 0
    out = np.zeros(ntones, complex)
    for k in range(ntones):
        b = obin + k * step
        if b >= n:
            break
        out[k] = dft[b]
    return out, implicit, n


def direct_sums(tones, chunks, ntones):
    out = np.zeros(ntones, complex)
    for k, t in enumerate(tones):
        if t is None:
            continue
        num, den = t
        for so, x in chunks:
            m = so + np.arange(len(x), dtype=np.int64)
            out[k] += (x * np.exp(-2j * np.pi * ((num * (m % den)) % den) / den)).sum()
    return out


@pytest.mark.parametrize("fs,offset,interval,ntones,kind", [
    (16_000_000, 1_000_000, 1_000_000, 15, "implicit, N = 16"),
    (16_000_000, 2_000_000, 1_000_000, 14, "shifting, N = 16"),
    (16_000_000, 250_000, 1_000_000, 16, "implicit, N = 64"),
    (64_000_000, 8_000_000, 8_000_000, 7, "implicit, N doubled 8 -> 32"),
    (64_000_000, 16_000_000, 8_000_000, 6, "shifting, N doubled 8 -> 32"),
    (32_000_000, 5_000_000, 5_000_000, 6, "implicit, offset = interval"),
    (32_000_000, 640_000, 5_000_000, 7, "shifting, N = 32"),
])
@pytest.mark.parametrize("dsb", [False, True])
def test_literal_extractors_equal_the_direct_sums(fs, offset, interval, ntones, kind, dsb):
    rng = np.random.default_rng(3)
    # FFT-sized chunks at scattered sample offsets, as Mode hands them over
    chunks = []
    for so in (0, 4096, 9000, 123457, 2_000_001):
        n = int(rng.integers(300, 2000))
        chunks.append((so, rng.standard_normal(n) + 1j * rng.standard_normal(n)))
    want, implicit, n = literal_complex_extractor(fs, offset, interval, ntones, dsb, chunks)
    assert implicit == kind.startswith("implicit")
    assert str(n) in kind.split("N ")[-1] or "offset = interval" in kind
    tones = pc.extractor_tones(ntones, offset, interval, fs, True, dsb)
    got = direct_sums(tones, chunks, ntones)
    scale = np.sqrt(np.mean(np.abs(want) ** 2))
    assert np.max(np.abs(got - want)) < 1e-9 * scale * 100
    # single sideband: the tones the comb defines, offset + k interval
    if not dsb:
        for k, t in enumerate(tones):
            if t is not None:
                f = offset + k * interval
                g = math.gcd(f % fs, fs)
                assert t == ((f % fs) // g, fs // g)


def test_bins_past_the_dft_length_stay_zero():
    """A tone on the upper band edge reads bin N, past the DFT: the extractor stops there and
    the tone stays zero (pcal.cpp:1012-1016, 1586-1588)."""
    tones = pc.extractor_tones(16, 1_000_000, 1_000_000, 16_000_000, True, False)
    assert tones[-1] is None and all(t is not None for t in tones[:-1])


def _complex_case(lsb=False, dsb=False, B=8):
    cfg = cf.standard_config(2, 1, [1400.0], 16.0, 256, subint_ns=B * 16000, nbits=8,
                             complex_sampling=2 if dsb else 1, lsb=lsb, max_delay_us=50.0,
                             max_rate_us_per_s=0.01, seed=22)
    return cfg


@pytest.mark.parametrize("lsb", [False, True])
def test_complex_planted_tones_closed_form(lsb):
    """Complex tones planted at 3 and 7 MHz on a 1 MHz comb (offset 1 MHz, fs 16 MHz, a 16-sample
    period, FFTs of 256 samples): tone k = (N / 16) sum_{m<16} Q(x_m) exp(-2 pi i f_k m / fs),
    and before quantisation N A exp(i phi) at the planted tones, 0 elsewhere."""
    cfg = _complex_case(lsb)
    B, L, fs, P = cfg.blockspersend(), cfg.fftchannels(0), 16_000_000, 16
    planted = [(3_000_000, 11.0, 0.4), (7_000_000, 6.0, -1.2)]
    offsets = [0, 1, 0]
    data, ctrl, decoded = pcal_complex_tone_slot(cfg, planted, offsets)
    n, off, _ = pc.tone_table(1400.0, 16.0, lsb, 1_000_000)
    assert (n, off) == (16, 1_000_000)
    for d in range(2):
        ffts, _ = pc.valid_ffts(cfg, d, offsets, data[d], ctrl[d])
        assert len(ffts) == B
        got = pc.extract(cfg, d, offsets, data[d], ctrl[d], 1_000_000)[0]
        mp = np.arange(P)
        q = decoded[d](mp)
        q = (q.real.astype(np.float32) + 1j * q.imag.astype(np.float32)).astype(np.complex128)
        fk = [off + k * 1_000_000 for k in range(n)]
        want = np.array([(B * L / P) * (q * np.exp(-2j * np.pi * ((f * mp) % fs) / fs)).sum() for f in fk])
        want[-1] = 0                                  # the 16 MHz tone reads bin N: left at zero
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-9 * B * L * 11.0)
        ideal = np.zeros(n, complex)
        for f, a, ph in planted:
            ideal[f // 1_000_000 - 1] = B * L * a * np.exp(1j * ph)
        assert np.max(np.abs(got - ideal)) < 0.01 * B * L * 11.0


def test_double_sideband_reads_the_reversed_bins():
    """Double sideband: the implicit-shift extractor (N = 16) reads bin b_k = 1 + k of the
    reversed DFT, i.e. frequency -(b_k + 1) fs / N = -(k + 2) MHz (modulo fs = 16 MHz).  A tone
    planted at -4 MHz lands in tone k = 2 with N A exp(i phi); one at +4 MHz = -12 MHz in k = 10."""
    cfg = _complex_case(dsb=True)
    B, L = cfg.blockspersend(), cfg.fftchannels(0)
    offsets = [0, 1, 0]
    tones = pc.band_tones(cfg, 0, 0, 1_000_000)[2]
    assert tones[:3] == [(7, 8), (13, 16), (3, 4)]             # 14/16, 13/16, 12/16 reduced
    for f, k in ((-4_000_000, 2), (4_000_000, 10)):
        data, ctrl, _ = pcal_complex_tone_slot(cfg, [(f, 30.0, 0.9)], offsets)
        got = pc.extract(cfg, 0, offsets, data[0], ctrl[0], 1_000_000)[0]
        if k is None:
            assert np.max(np.abs(got)) < 0.01 * B * L * 30.0
        else:
            assert abs(got[k] - B * L * 30.0 * np.exp(0.9j)) < 0.01 * B * L * 30.0
            assert np.max(np.abs(np.delete(got, k))) < 0.01 * B * L * 30.0


def _pcal_config():
    """Two stations, two 16 MHz freqs dual pol (real 2-bit), a 1 MHz comb on station 0 only."""
    cfg = cf.standard_config(2, 2, [1400.0, 1416.37], 16.0, 256, subint_ns=8 * 16000, seed=23,
                             max_delay_us=50.0)
    cfg.datastreams[0].phasecal_interval_hz = 1_000_000
    return cfg


def test_pcal_block_layout_by_hand():
    """The block follows the autocorrelation weights: per datastream with a comb, per recorded
    freq its pols times its tones.  1400 MHz: tones 1401-1416 MHz (16, the last on the band
    edge); 1416.37 MHz: 1417-1432 MHz (16)."""
    cfg = _pcal_config()
    lay = fo.result_layout(cfg)
    plain = cf.standard_config(2, 2, [1400.0, 1416.37], 16.0, 256, subint_ns=8 * 16000, seed=23)
    lay0 = fo.result_layout(plain)
    assert lay.pcal_off == [lay0.coreresultlength, lay0.coreresultlength + 2 * 16 + 2 * 16]
    assert lay.coreresultlength == lay0.coreresultlength + 64
    assert [pc.tone_table(f, 16.0, False, 1_000_000)[0] for f in (1400.0, 1416.37)] == [16, 16]


def test_library_layout_has_the_block():
    import ctypes as C
    from pathlib import Path
    if not (Path(__file__).resolve().parent.parent / "difx_amd" / "libfxg.so").exists():
        pytest.skip("libfxg.so not built")
    from difx_amd import fxg_abi as abi
    from difx_amd.core import load_library
    lib = load_library()
    for cfg in (_pcal_config(), _complex_case(dsb=True)):
        if cfg.datastreams[0].phasecal_interval_hz == 0:
            for d in cfg.datastreams:
                d.phasecal_interval_hz = 1_000_000
        lay = fo.result_layout(cfg)
        fa, keep = cfg.to_abi()
        h = C.c_void_p()
        assert lib.fxg_create_layout(C.byref(fa), C.byref(h)) == 0
        try:
            lo = abi.fxg_layout()
            assert lib.fxg_get_layout(h, C.byref(lo)) == 0
            assert lo.coreresultlength == lay.coreresultlength
            p = (C.c_int32 * len(cfg.datastreams))()
            assert lib.fxg_result_pcal_offsets(h, p) == 0
            assert list(p) == lay.pcal_off
            # the tone period the device folds with is a multiple of every tone's denominator
            for d, ds in enumerate(cfg.datastreams):
                if ds.phasecal_interval_hz <= 0:
                    continue
                for b in range(len(ds.recbands)):
                    nt, of, per = C.c_int32(), C.c_int64(), C.c_int32()
                    assert lib.fxg_pcal_info(h, d, b, C.byref(nt), C.byref(of), C.byref(per)) == 0
                    n, off, tones = pc.band_tones(cfg, d, b, ds.phasecal_interval_hz)
                    assert (nt.value, of.value) == (n, off)
                    assert per.value > 0 and all(per.value % t[1] == 0 for t in tones if t is not None)
        finally:
            lib.fxg_destroy(h)


def test_library_refuses_a_changed_comb_on_a_configured_datastream():
    import ctypes as C
    from pathlib import Path
    if not (Path(__file__).resolve().parent.parent / "difx_amd" / "libfxg.so").exists():
        pytest.skip("libfxg.so not built")
    from difx_amd.core import load_library
    lib = load_library()
    cfg = _pcal_config()
    fa, keep = cfg.to_abi()
    h = C.c_void_p()
    assert lib.fxg_create_layout(C.byref(fa), C.byref(h)) == 0
    try:
        assert lib.fxg_set_pcal(h, 0, 1_000_000, 0) == 0          # the configured comb: accepted
        assert lib.fxg_set_pcal(h, 0, 5_000_000, 0) != 0          # another: the layout fixes it
        assert lib.fxg_set_pcal(h, 1, 5_000_000, 0) == 0          # a datastream without a block
    finally:
        lib.fxg_destroy(h)


def test_pcal_block_scaling_closed_form():
    """writedata scales each band's tones by 1 / (acw meansubints (float)(B 2 nchan)), acw the
    band's parallel autocorrelation weight over the FFTs of the integration
    (visibility.cpp:650-675, 448-459)."""
    cfg = _pcal_config()
    data, ctrl = slot(cfg, [0, 1, 0], seed=5)
    res, _ = fo.process_subint(cfg, [0, 1, 0], data, ctrl)
    lay = fo.result_layout(cfg)
    out = vo.normalise(cfg, res.astype(np.complex64), lay, 1.0)
    B = cfg.blockspersend()
    fres = res.view(np.float64)
    w = lay.acweight_off[0] * 2
    o = lay.pcal_off[0]
    for k in range(4):                             # bands R, L of freq 0, then of freq 1
        acw = np.float32(fres[w + k] / B)
        sc = np.float32(1.0 / (np.float64(acw) * 1.0 * np.float64(np.float32(B * 2 * 256))))
        t = res[o: o + 16].astype(np.complex64)
        np.testing.assert_array_equal(out[o: o + 16], (t.view(np.float32) * sc).view(np.complex64))
        assert np.any(t)
        o += 16
