"""Phase-cal extraction on the device (fxg_set_pcal / fxg_get_pcal) against the restatement
(oracle/pcal.py): 8-bit and 2-bit VDIF, USB and LSB bands, a damaged slot (flagged FFTs,
invalid and fill-pattern frames), two subints in a row (the sums restart every subint), a
flagged offset; and planted tones whose values follow in closed form.
Bar: |d tone| <= 1e-4 of the band's rms tone (FP32 folding of up to 1e5 samples per bin
against FP64 sums)."""
import numpy as np
import pytest

import pcal as pc
from difx_amd import config as cf
from fxcases import (case_cfg0, case_cfg1, case_cfg3, case_mark5b, case_offsets, case_real8, damaged_slot,
                     pcal_tone_slot, slot)

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
