"""Spectral kurtosis on the device (fxg_set_kurtosis / fxg_get_kurtosis) against the
restatement (oracle/kurtosis.py) applied to the FP64 oracle's spectra, for the FP32 and the
FP16 spectra layouts, full resolution and averaged; and a planted tone: a constant-modulus
signal drives its channel's estimator to -1 while the noise channels stay near 0.
Bars: |d sk| <= 1e-4 (FP32 spectra), <= 2e-2 (FP16 spectra)."""
import sys

import numpy as np
import pytest

import fxoracle as fo
import kurtosis as ku
from difx_amd import synth
from fxcases import case_cfg1, slot

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("xengine,tol", [("fp32", 1e-4), ("fp16", 2e-2)])
def test_kurtosis_matches_the_oracle(gpu, xengine, tol):
    from difx_amd.core import FxCore
    print("kurtosis", xengine, flush=True)
    cfg = case_cfg1(B=32, nstations=3, nfreq=2, xengine=xengine)
    data, ctrl = slot(cfg, seed=51)
    core = FxCore(cfg)
    core.set_kurtosis(True)
    core.processdata([0, 1, 0], data, ctrl)
    _, stations = fo.process_subint(cfg, [0, 1, 0], data, ctrl)
    B = cfg.blockspersend()
    worst = 0.0
    for d, ds in enumerate(cfg.datastreams):
        for b in range(len(ds.recbands)):
            s1, s2 = ku.sums(stations[d].spectra[b])
            C = s1.size
            for mc in (C, C // 8):
                got = core.getKurtosis(d, b, mc)
                want = ku.estimator(s1, s2, B, mc)
                assert got.shape == want.shape
                worst = max(worst, float(np.max(np.abs(got - want))))
    print("  max |d sk|", worst, flush=True)
    assert worst <= tol
    core.close()


def test_planted_tone(gpu):
    from difx_amd import config as cf
    from difx_amd.core import FxCore
    print("kurtosis planted tone", flush=True)
    B, nchan, bw, k = 256, 256, 32.0, 100
    cfg = cf.standard_config(2, 1, [8400.0], bw, nchan, subint_ns=B * 8000, max_delay_us=50.0,
                             max_rate_us_per_s=0.01, seed=52)
    offsets = [0, 1, 0]
    rng = np.random.default_rng(53)
    f = (k + 0.5) * bw / nchan * 1e6                       # centre of channel k, Hz
    data, ctrl = [], []
    for d, ds in enumerate(cfg.datastreams):
        start_ns = synth.buffer_start_ns(cfg, d, offsets)
        nbytes = synth.slot_bytes(cfg, d, offsets, start_ns)
        nframes = nbytes // ds.framebytes
        nsamp = nframes * (ds.framebytes - ds.headerbytes) * 4
        t = start_ns * 1e-9 + np.arange(nsamp) / (2 * bw * 1e6)
        x = rng.standard_normal(nsamp) + 3.0 * np.cos(2 * np.pi * f * t)
        payload = synth.pack_2bit(synth.quantize_2bit(x).reshape(-1, 1))
        frame_ns = int(round(synth.samples_per_frame(ds) * cfg.sampletime_us(d) * 1000.0))
        data.append(synth.frame_payloads(payload, ds, int(start_ns // frame_ns)))
        ctrl.append(synth.control_words(cfg, d, offsets, start_ns))
    core = FxCore(cfg)
    core.set_kurtosis(True)
    core.processdata(offsets, data, ctrl)
    for d in range(2):
        sk = core.getKurtosis(d, 0, nchan)
        others = np.delete(sk, [k])
        print(f"  ds {d}: sk[tone] {sk[k]:.3f}, median |sk| elsewhere {np.median(np.abs(others)):.3f}", flush=True)
        assert sk[k] < -0.9
        assert np.median(np.abs(others)) < 0.2
    core.close()
