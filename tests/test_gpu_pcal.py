"""Phase-cal extraction on the device (fxg_set_pcal / fxg_get_pcal) against the restatement
(oracle/pcal.py): 8-bit and 2-bit VDIF, USB and LSB bands, a damaged slot (flagged FFTs,
invalid and fill-pattern frames), two subints in a row (the sums restart every subint), a
flagged offset; and planted tones whose values follow in closed form.
Bar: |d tone| <= 1e-4 of the band's rms tone (FP32 folding of up to 1e5 samples per bin
against FP64 sums)."""
import numpy as np
import pytest

import fxoracle as fo
import pcal as pc
import visout as vo
from difx_amd import config as cf
from fxcases import (case_cfg0, case_cfg1, case_cfg2, case_cfg3, case_dsb_fraclo, case_mark5b, case_offsets,
                     case_real8, damaged_slot, pcal_complex_tone_slot, pcal_tone_slot, record_errors, slot)

pytestmark = pytest.mark.gpu

TOL = 1e-4


def check_against_oracle(core, cfg, offsets, data, ctrl, interval, base=0):
    worst = 0.0
    for d, ds in enumerate(cfg.datastreams):
        want = pc.extract(cfg, d, offsets, data[d], ctrl[d], interval, base)
        for b in range(len(ds.recbands)):
            got = core.getPcal(d, b)
            assert got.shape == want[b].shape
            if not np.any(want[b]):
                assert not np.any(got)
                continue
            scale = np.sqrt(np.mean(np.abs(want[b]) ** 2))
            worst = max(worst, float(np.max(np.abs(got - want[b]))) / scale)
    print("  max |d tone| / rms tone", worst, flush=True)
    assert worst <= TOL


@pytest.mark.parametrize("name,make,interval", [
    ("8-bit real, 2 freqs dual pol", lambda: case_real8(B=32), 1_000_000),
    ("2-bit VDIF USB, 5 MHz comb", lambda: case_cfg0(B=32), 5_000_000),
    ("2-bit VDIF LSB", lambda: case_cfg0(B=32, lsb=True), 1_000_000),
    ("2-bit VDIF 16 bands", lambda: case_cfg1(B=16, nstations=2, nfreq=4), 1_000_000),
    ("2-bit Mark5B 16 bands", lambda: case_mark5b(B=16), 1_000_000),
    # 4800.37 MHz: offset 630 kHz, a 6400-sample period (per-sample LDS adds) beside a
    # 64-sample one (register sums) in the same datastream
    ("fractional-MHz band edge", lambda: case_offsets(B=32), 1_000_000),
])
def test_pcal_matches_the_oracle(gpu, name, make, interval):
    from difx_amd.core import FxCore
    print("pcal", name, flush=True)
    cfg = make()
    core = FxCore(cfg)
    for d in range(len(cfg.datastreams)):
        core.set_pcal(d, interval)
    offsets = [0, 1, 0]
    data, ctrl = damaged_slot(cfg, offsets)
    core.processdata(offsets, data, ctrl)
    check_against_oracle(core, cfg, offsets, data, ctrl, interval)
    # the next subint starts its sums afresh
    offsets2 = [0, 1, cfg.subint_ns]
    data2, ctrl2 = slot(cfg, offsets2, seed=31)
    core.processdata(offsets2, data2, ctrl2)
    check_against_oracle(core, cfg, offsets2, data2, ctrl2, interval)
    core.close()


def test_flagged_offset_and_off(gpu):
    """A 6 MHz base on a 5 MHz comb at 1400 MHz flags the offset (-5 MHz): zero tones; a
    datastream with phase-cal off reports a state error."""
    from difx_amd.core import FxCore, FxgError
    print("pcal flagged offset", flush=True)
    cfg = cf.standard_config(2, 1, [1400.0], 16.0, 256, subint_ns=32 * 16000, nbits=8,
                             max_delay_us=50.0, max_rate_us_per_s=0.01, seed=21)
    core = FxCore(cfg)
    core.set_pcal(0, 5_000_000, 6_000_000)
    data, ctrl = slot(cfg)
    core.processdata([0, 1, 0], data, ctrl)
    got = core.getPcal(0, 0)
    assert got.shape == (4,) and not np.any(got)
    with pytest.raises(FxgError):
        core.getPcal(1, 0)
    core.close()


@pytest.mark.parametrize("lsb", [False, True])
def test_planted_tones(gpu, lsb):
    """Tones planted at 8 and 1 MHz on a 1 MHz comb: the device returns N A / 2 exp(i phi)
    at them (within the 8-bit quantisation, 1%) and the exact quantised-period values
    (N / P) sum_{m<P} Q(x_m) exp(-2 pi i f m / fs), P = 32, within the bar."""
    from difx_amd.core import FxCore
    print("pcal planted tones lsb", lsb, flush=True)
    B = 32
    cfg = cf.standard_config(2, 1, [1400.0], 16.0, 256, subint_ns=B * 16000, nbits=8, lsb=lsb,
                             max_delay_us=50.0, max_rate_us_per_s=0.01, seed=21)
    L, fs, P = cfg.fftchannels(0), 32_000_000, 32
    planted = [(8_000_000, 30.0, 0.7), (1_000_000, 12.0, -2.1)]
    offsets = [0, 1, 0]
    data, ctrl, decoded = pcal_tone_slot(cfg, planted, offsets)
    core = FxCore(cfg)
    for d in range(2):
        core.set_pcal(d, 1_000_000)
    core.processdata(offsets, data, ctrl)
    fk = np.arange(1, 17) * 1_000_000
    mp = np.arange(P)
    for d in range(2):
        got = core.getPcal(d, 0)
        q = decoded[d](mp).astype(np.float32).astype(np.float64)
        want = np.array([(B * L / P) * (q * np.exp(-2j * np.pi * ((f * mp) % fs) / fs)).sum() for f in fk])
        err = float(np.max(np.abs(got - want))) / (B * L * 30.0 / 2)
        print(f"  ds {d}: tone 8 MHz {got[7]:.6g}, max err / tone {err:.3g}", flush=True)
        assert err <= TOL
        for f, a, ph in planted:
            k = f // 1_000_000 - 1
            assert abs(abs(got[k]) / (B * L * a / 2) - 1) < 0.01
            assert abs(np.angle(got[k] * np.exp(-1j * ph))) < 0.01
    core.close()


def test_buffer_before_the_scan_start(gpu):
    """offsets (0, 0, 0): station 0's buffer starts 5.6 ms before the scan start (datasec -1),
    so every sample number of its subint is negative; the bins take m mod M in [0, M)."""
    from difx_amd.core import FxCore
    print("pcal buffer before the scan start", flush=True)
    cfg = case_real8(B=32)
    offsets = [0, 0, 0]
    data, ctrl = slot(cfg, offsets)
    assert int(ctrl[0][1]) == -1
    core = FxCore(cfg)
    for d in range(len(cfg.datastreams)):
        core.set_pcal(d, 1_000_000)
    core.processdata(offsets, data, ctrl)
    check_against_oracle(core, cfg, offsets, data, ctrl, 1_000_000)
    core.close()


@pytest.mark.parametrize("mode", [1, 2])
def test_uneven_bands_both_fengine_forms(gpu, mode):
    """Uneven bands (8, 32 and 64 MHz datastreams with zooms; tone periods 16, 64 and 128
    samples) with the per-length F-engine launches (1) and the persistent one (2): the
    phase-cal pass reads each chunk's per-FFT delays from either."""
    from difx_amd.core import FxCore
    print("pcal uneven bands, F-engine form", mode, flush=True)
    cfg = case_cfg3(B=32, fengine_mode=mode)
    offsets = [0, 1, 0]
    data, ctrl = damaged_slot(cfg, offsets)
    core = FxCore(cfg)
    for d in range(len(cfg.datastreams)):
        core.set_pcal(d, 1_000_000)
    core.processdata(offsets, data, ctrl)
    check_against_oracle(core, cfg, offsets, data, ctrl, 1_000_000)
    core.close()


def _cfg2_lsb():
    cfg = case_cfg2(B=16, nstations=4, xengine="fp32")
    cfg.freqs[0].lsb = True
    return cfg


@pytest.mark.parametrize("name,make,interval", [
    ("complex 8-bit USB, 16 MHz (implicit-shift extractor)", lambda: case_cfg2(B=16, nstations=4, xengine="fp32"), 1_000_000),
    ("complex 8-bit LSB, 16 MHz", _cfg2_lsb, 1_000_000),
    ("complex 8-bit USB, 3 MHz comb, offset 2 MHz (shifting extractor)",
     lambda: cf.standard_config(4, 1, [1399.0], 16.0, 4096, subint_ns=16 * 256000, nbits=8, complex_sampling=1,
                                xengine="fp32", seed=3), 3_000_000),
    ("complex double sideband, USB and LSB freqs, 5 MHz", lambda: case_dsb_fraclo(B=16), 1_000_000),
])
def test_complex_pcal_matches_the_oracle(gpu, name, make, interval):
    """Complex-sampled data: the real and imaginary fields folded separately on the device,
    the tones at the extractor's rational frequencies (implicit-shift or shifting complex
    extractor, double-sideband reversal) against the restatement, on a damaged slot and on
    the next subint."""
    from difx_amd.core import FxCore
    print("pcal", name, flush=True)
    cfg = make()
    core = FxCore(cfg)
    for d in range(len(cfg.datastreams)):
        core.set_pcal(d, interval)
    offsets = [0, 1, 0]
    data, ctrl = damaged_slot(cfg, offsets)
    core.processdata(offsets, data, ctrl)
    check_against_oracle(core, cfg, offsets, data, ctrl, interval)
    offsets2 = [0, 1, cfg.subint_ns]
    data2, ctrl2 = slot(cfg, offsets2, seed=33)
    core.processdata(offsets2, data2, ctrl2)
    check_against_oracle(core, cfg, offsets2, data2, ctrl2, interval)
    core.close()


@pytest.mark.parametrize("dsb", [False, True])
def test_complex_planted_tones(gpu, dsb):
    """Complex tones planted at +3 and -4 MHz on a 1 MHz comb (16 MHz complex band): the
    device returns N A exp(i phi) in the tone that reads each frequency (single sideband: the
    +3 MHz tone; double sideband: the reversed bins, -4 MHz in tone 2 and +3 MHz = -13 MHz in
    tone 11)."""
    from difx_amd.core import FxCore
    print("pcal complex planted tones, dsb", dsb, flush=True)
    cfg = cf.standard_config(2, 1, [1400.0], 16.0, 256, subint_ns=8 * 16000, nbits=8,
                             complex_sampling=2 if dsb else 1, max_delay_us=50.0, max_rate_us_per_s=0.01,
                             seed=22)
    B, L = cfg.blockspersend(), cfg.fftchannels(0)
    planted = [(3_000_000, 24.0, 0.4), (-4_000_000, 12.0, -1.0)]   # peak 36 x 3.3 < 127: no clipping
    offsets = [0, 1, 0]
    data, ctrl, _ = pcal_complex_tone_slot(cfg, planted, offsets)
    core = FxCore(cfg)
    for d in range(2):
        core.set_pcal(d, 1_000_000)
    core.processdata(offsets, data, ctrl)
    where = {3_000_000: 11 if dsb else 2, -4_000_000: 2 if dsb else None}
    for d in range(2):
        got = core.getPcal(d, 0)
        want = pc.extract(cfg, d, offsets, data[d], ctrl[d], 1_000_000)[0]
        err = float(np.max(np.abs(got - want))) / (B * L * 24.0)
        print(f"  ds {d}: max err / tone {err:.3g}", flush=True)
        assert err <= TOL
        for f, a, ph in planted:                    # within the 8-bit quantisation
            k = where[f]
            if k is None:
                continue
            assert abs(abs(got[k]) / (B * L * a) - 1) < 0.015
            assert abs(np.angle(got[k] * np.exp(-1j * ph))) < 0.015
    core.close()


@pytest.mark.parametrize("name,make", [
    ("2-bit VDIF dual pol, two freqs", lambda: case_cfg1(B=16, nstations=3, nfreq=2)),
    ("complex double sideband", lambda: case_dsb_fraclo(B=16)),
])
def test_results_pcal_block(gpu, name, make):
    """A phase-cal interval in the configuration: every subint's tones go into the results'
    phase-cal block (Core::copyPCalTones), the rest of the results unchanged, against the
    restatement's process_subint; then the block's writedata scaling on the device, bit for bit
    against the restatement's."""
    import torch
    from difx_amd.core import FxCore
    print("pcal results block:", name, flush=True)
    cfg = make()
    for k, ds in enumerate(cfg.datastreams):
        ds.phasecal_interval_hz = 1_000_000 if k != 1 else 0       # datastream 1 without a comb
    lay = fo.result_layout(cfg)
    offsets = [0, 1, 0]
    data, ctrl = damaged_slot(cfg, offsets)
    core = FxCore(cfg)
    assert core.coreresultlength == lay.coreresultlength
    assert list(core.result_pcal_offsets()) == lay.pcal_off
    got = core.processdata(offsets, data, ctrl)
    ref, _ = fo.process_subint(cfg, offsets, data, ctrl)
    e = record_errors(cfg, lay, got.astype(np.complex128), ref)
    assert max(e["cross"]) < 1e-5 and max(e["auto"]) < 1e-5
    worst = 0.0
    for d, ds in enumerate(cfg.datastreams):
        o0 = lay.pcal_off[d]
        o1 = lay.pcal_off[d + 1] if d + 1 < len(cfg.datastreams) else lay.coreresultlength
        if ds.phasecal_interval_hz <= 0:
            assert o1 == o0
            continue
        g, w = got[o0:o1].astype(np.complex128), ref[o0:o1]
        assert np.any(w)
        worst = max(worst, float(np.max(np.abs(g - w)) / np.sqrt(np.mean(np.abs(w) ** 2))))
    print("  block max |d| / rms", worst, flush=True)
    assert worst <= TOL
    # the device scaling of the block against the restatement's, on the device results
    dres = torch.from_numpy(got.view(np.float32).copy()).to(gpu)
    core.normalise_device(dres.data_ptr(), 1.0)
    scaled = dres.cpu().numpy().view(np.complex64)
    want = vo.normalise(cfg, got, lay, 1.0)
    np.testing.assert_array_equal(scaled, want)
    o = lay.pcal_off[0]
    assert not np.array_equal(scaled[o:o + 4], got[o:o + 4])
    core.close()
