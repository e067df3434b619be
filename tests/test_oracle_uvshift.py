"""Pins of the oracle's multi-phase-centre uv shift (Core::uvshiftAndAverageBaselineFreq,
core.cpp:1605-1938; layout configuration.cpp:2464-2566), CPU only.

The pins: a phase centre at the pointing centre changes nothing; a phase centre offset by a
known delay on one station turns every channel by exactly 2 pi delay x its sky frequency;
a common signal that sits at a phase centre, correlated with the pointing-centre model, is
brought to flat phase and full coherence by the shift to that centre; the interval
partition and the decorrelation factors follow their definitions on hand-worked cases."""
import math

import numpy as np
import pytest

import fxoracle as fo
from difx_amd import config as cf
from difx_amd import synth

OFFSETS = [0, 1, 0]


def _const_models(cfg, delays):
    for d, ds in enumerate(cfg.datastreams):
        ds.model = np.zeros_like(ds.model)
        ds.model[:, 0] = delays[d]


def _with_centres(cfg, centre_delays, **kw):
    """Phase centres given as per-datastream constant delays: centre_delays[s][d]."""
    cfg.nphasecentres = len(centre_delays)
    for k, v in kw.items():
        setattr(cfg, k, v)
    for d, ds in enumerate(cfg.datastreams):
        ms = []
        for cd in centre_delays:
            m = np.zeros_like(ds.model)
            m[:, 0] = cd[d]
            ms.append(m)
        ds.phasecentre_models = ms
    return cfg


def test_intervals_hand_worked():
    cfg = cf.standard_config(2, 1, [4800.0], 32.0, 256, subint_ns=40 * 8000)
    cfg.num_buffered_ffts = 10
    cfg.max_ns_between_xc_avg = 15 * 8000 + 100        # 15 FFTs -> 10 (multiple of 10)
    cfg.process_threads = 2
    iv = fo.xc_intervals(cfg, 40)
    assert [(a, n) for a, n, _, _ in iv] == [(0, 10), (10, 10), (20, 10), (30, 10)]
    assert [c for _, _, c, _ in iv] == [5 * 8000, 15 * 8000, 25 * 8000, 35 * 8000]
    assert all(w == 10 * 8000 for _, _, _, w in iv)
    cfg.subint_ns = 41 * 8000
    iv = fo.xc_intervals(cfg, 41)
    assert [(a, n) for a, n, _, _ in iv] == [(0, 10), (10, 10), (20, 1), (21, 10), (31, 10)]
    assert iv[2][2] == (20 + 0.5) * 8000 and iv[2][3] == 8000
    cfg.max_ns_between_xc_avg = 5 * 8000                # below numBufferedFFTs: one buffer each
    iv = fo.xc_intervals(cfg, 41)
    assert [n for _, n, _, _ in iv] == [10, 10, 1, 10, 10]


def test_layout_with_phase_centres_hand_count():
    # 3 stations single pol, one freq of 256 channels, 2 phase centres: 3 baselines x 2 x 256
    # crosses, 3 x 1 weight words, 3 x 1 shift-decorrelation words, 3 x 256 autos, 3 acweights
    cfg = cf.standard_config(3, 1, [4800.0], 32.0, 256, subint_ns=8 * 8000)
    cfg.nphasecentres = 2
    lay = fo.result_layout(cfg)
    assert lay.baseline_off[(0, 0)] == 0 and lay.baseline_off[(0, 1)] == 512 and lay.baseline_off[(0, 2)] == 1024
    assert lay.bweight_off[(0, 0)] == 1536 and lay.bweight_off[(0, 2)] == 1538
    assert lay.shiftdecorr_off[(0, 0)] == 1539 and lay.shiftdecorr_off[(0, 2)] == 1541
    assert lay.autocorr_off == [1542, 1542 + 256, 1542 + 512]
    assert lay.coreresultlength == 1542 + 768 + 3
    cfg.nphasecentres = 3                               # 3 floats need 2 cf32 words
    lay = fo.result_layout(cfg)
    assert lay.shiftdecorr_off[(0, 1)] - lay.shiftdecorr_off[(0, 0)] == 2


def _run(cfg, data, ctrl):
    res, _ = fo.process_subint(cfg, OFFSETS, data, ctrl)
    return res, fo.result_layout(cfg)


def _block(res, lay, cfg, f, j, s, p=0):
    fr = cfg.freqs[f]
    nprod = len(dict(cfg.baselines[j].freqs)[f])
    nout = (fr.nchan + fr.chanavg - 1) // fr.chanavg
    o = lay.baseline_off[(f, j)] + (s * nprod + p) * nout
    return res[o: o + nout]


@pytest.mark.parametrize("lsb", [False, True])
def test_pointing_centre_is_identity_and_offset_is_linear_phase(lsb):
    cfg1 = cf.standard_config(3, 1, [4800.0], 32.0, 256, subint_ns=32 * 8000, seed=31, lsb=lsb)
    _const_models(cfg1, [12.5, -3.25, 7.0])
    data, ctrl = synth.noise_slot(cfg1, OFFSETS, seed=31)
    base, lay1 = _run(cfg1, data, ctrl)
    delta = 0.0123                                      # us, on station 2 (datastream 1) only
    cfg = _with_centres(cf.standard_config(3, 1, [4800.0], 32.0, 256, subint_ns=32 * 8000, seed=31, lsb=lsb),
                        [[12.5, -3.25, 7.0], [12.5, -3.25 + delta, 7.0]],
                        max_ns_between_xc_avg=8 * 8000, num_buffered_ffts=4)
    _const_models(cfg, [12.5, -3.25, 7.0])
    res, lay = _run(cfg, data, ctrl)
    fr = cfg.freqs[0]
    c = np.arange(fr.nchan)
    chbw = fr.bw_mhz / fr.nchan
    fsky = fr.freq_mhz - (fr.nchan - 1 - c) * chbw if lsb else fr.freq_mhz + c * chbw
    for j, bl in enumerate(cfg.baselines):
        b0 = _block(base, lay1, cfg1, 0, j, 0)
        # centre 0 is the pointing centre: the unshifted visibilities
        np.testing.assert_allclose(_block(res, lay, cfg, 0, j, 0), b0, rtol=0, atol=1e-9 * np.max(np.abs(b0)))
        # centre 1 moves station 2 by delta: baseline (0, 1) turns by +delta, (1, 2) by -delta
        dd = delta if bl.ds2 == 1 else (-delta if bl.ds1 == 1 else 0.0)
        want = b0 * np.exp(2j * np.pi * dd * fsky)
        np.testing.assert_allclose(_block(res, lay, cfg, 0, j, 1), want, rtol=0, atol=1e-8 * np.max(np.abs(b0)))
        # shift decorrelation: no rate difference, delay decorrelation 1 - |dd| / (nchan / bw)
        sd = res.view(np.float64)[lay.shiftdecorr_off[(0, j)] * 2: lay.shiftdecorr_off[(0, j)] * 2 + 2]
        assert sd[0] == pytest.approx(cfg.subint_ns, rel=1e-12)
        assert sd[1] == pytest.approx(cfg.subint_ns * (1 - abs(dd) / (fr.nchan / fr.bw_mhz)), rel=1e-12)
    # autocorrelations and weights are those of the single-centre run
    np.testing.assert_array_equal(res[lay.autocorr_off[0]:], base[lay1.autocorr_off[0]:])


def test_shift_to_the_source_flattens_the_fringe():
    """A common signal placed at a phase centre (station delays tau_k + dtau_k) and correlated
    with the pointing-centre model tau_k shows a residual fringe across the band; shifting to
    the centre where the signal sits restores flat phase and the coherence of a zero-offset
    correlation."""
    B = 256
    taus = [3.0, 5.5, -2.25]
    dtaus = [0.0, 0.037, -0.021]                       # us
    mk = lambda: cf.standard_config(3, 1, [4800.0], 32.0, 256, subint_ns=B * 8000, seed=33)
    src = mk()
    _const_models(src, [t + d for t, d in zip(taus, dtaus)])
    data, ctrl = synth.correlated_slot(src, OFFSETS, seed=33, amp=1.0)
    # the slot buffers were placed for the source delays; their control words stay valid
    # centre 0 where the signal is, centre 1 the pointing centre itself
    cfg = _with_centres(mk(), [[t + d for t, d in zip(taus, dtaus)], taus], max_ns_between_xc_avg=16 * 8000)
    _const_models(cfg, taus)
    res, lay = _run(cfg, data, ctrl)
    ref, _ = _run(src, data, ctrl)                      # correlated with the source's own model
    lay_ref = fo.result_layout(src)
    for j in range(len(cfg.baselines)):
        v_src = _block(res, lay, cfg, 0, j, 0)
        v_pt = _block(res, lay, cfg, 0, j, 1)
        v_ref = _block(ref, lay_ref, src, 0, j, 0)
        band = slice(16, 240)
        ph = lambda v: np.angle(np.sum(v[band]))
        slope = lambda v: np.angle(np.sum(v[band][1:] * np.conj(v[band][:-1])))
        # the shifted visibilities have the phase and slope of the source-model correlation
        assert abs(np.angle(np.exp(1j * (ph(v_src) - ph(v_ref))))) < 0.02
        assert abs(ph(v_src)) < 0.02
        assert abs(slope(v_src)) < 0.005
        # the pointing-centre block keeps the residual delay: a phase slope across channels
        ddl = dtaus[cfg.baselines[j].ds2] - dtaus[cfg.baselines[j].ds1]
        assert slope(v_pt) == pytest.approx(-2 * np.pi * ddl * 32.0 / 256, abs=3e-3)
        assert abs(slope(v_pt)) > 0.015
        assert abs(np.mean(np.abs(v_src[band]))) > 0.95 * abs(np.mean(np.abs(v_ref[band])))


def test_time_smearing_decorrelation_closed_form():
    fr = cf.Freq(4800.0, 32.0, 256)
    rate = 1e-9                                        # us per us (1 ns/s) differential rate
    w = 1e8                                            # 0.1 s in ns
    m = 2 * np.pi * rate * (w / 1000.0) * 4800.0
    assert fo.shift_decorrelation(fr, np.array([rate, 0.0]), w) == pytest.approx(w * math.sin(m / 2) / (m / 2))
    # past the first null the interval decorrelates completely
    assert fo.shift_decorrelation(fr, np.array([1e-6, 0.0]), w) == 0.0
    assert fo.shift_decorrelation(fr, np.array([0.0, 100.0]), w) == 0.0    # delay beyond the window


def test_shift_pin_catches_a_sign_mistake(monkeypatch):
    """The shift with the opposite sign leaves twice the residual slope: the flat-fringe pin
    above fails for it."""
    real = fo.uvshift_phasor
    monkeypatch.setattr(fo, "uvshift_phasor", lambda fr, d: np.conj(real(fr, d)))
    with pytest.raises(AssertionError):
        test_shift_to_the_source_flattens_the_fringe()
