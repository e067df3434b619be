"""Pins of the oracle's per-freq correction terms (CPU only): the per-polarisation
clock-delta phasors (mode.cpp:326-341, 1030-1122, 1257-1261), the polarisation phase offset
(mode.cpp:421-447, 1318-1330), the clock offset, the LO offset with its fractional-Hz term
(mode.cpp:910-931) and the fractional-LO test of double-sideband data (mode.cpp:83-101).

Each is checked against what the terms mean rather than against their formula: a phase that a
physical quantity fixes (a cross-hand phase, a fringe phase at a wall-clock time), or an
equivalence between two configurations.  Planted mistakes (a single phasor for both pols, a
dropped fractional-Hz term, the raw band edge in the fractional-LO test) fail them."""
import math

import numpy as np
import pytest

import fxoracle as fo
from difx_amd import config as cf
from difx_amd import synth

OFFSETS = [0, 1, 0]


def _dual_pol_station(B=4, nchan=256, lofreq=4800.0, pols="RL", seed=21):
    """Two stations, one freq with two bands; station 0's two bands carry identical samples."""
    cfg = cf.standard_config(2, 2, [lofreq], 32.0, nchan, subint_ns=B * 1000 * nchan // 32, seed=seed)
    for d in cfg.datastreams:
        d.recbands = [(0, pols[0]), (0, pols[1])]
    data, ctrl = synth.noise_slot(cfg, OFFSETS, seed=seed)
    ds = cfg.datastreams[0]
    nframes = len(data[0]) // ds.framebytes
    pay = ds.framebytes - ds.headerbytes
    rng = np.random.default_rng(seed + 1)
    codes = synth.random_2bit_codes(rng, nframes * pay * 2)          # one code per time sample
    payload = synth.pack_2bit(np.repeat(codes.reshape(-1, 1), 2, axis=1))   # both bands equal
    start = synth.buffer_start_ns(cfg, 0, OFFSETS)
    frame_ns = int(round(synth.samples_per_frame(ds) * cfg.sampletime_us(0) * 1000.0))
    data[0] = synth.frame_payloads(payload, ds, start // frame_ns)
    return cfg, data, ctrl


def _chan_rad_per_us(cfg, d=0):
    """Baseband channel angular frequency (rad/us) of each USB channel."""
    f = cfg.freqs[cfg.datastreams[d].recfreqs[0]]
    return fo.TWO_PI * np.arange(f.nchan) * f.bw_mhz / f.nchan


# ------------------------------------------------ per-pol clock delta (fracsamprotatorA / B)
def _crosshand_phase_error(delta_us):
    cfg, data, ctrl = _dual_pol_station()
    cfg.datastreams[0].clock_offset_delta_us = [delta_us]
    st = fo.process_station(cfg, 0, OFFSETS, data[0], ctrl[0])
    # R carries +delta/2 and L -delta/2 of the same spectra: R x conj(L) = |X|^2 exp(i w delta)
    want = _chan_rad_per_us(cfg) * delta_us
    rl = st.autos[1, 0]
    ok = np.abs(rl) > 0
    return float(np.max(np.abs(np.angle(rl[ok] * np.exp(-1j * want[ok]))))), st


def test_pol_delta_crosshand_phase_is_chanfreq_times_delta():
    err, st = _crosshand_phase_error(0.004)        # 4 ns: 0.8 rad at the band edge
    assert err < 1e-9
    # the parallel-hand autocorrelations are untouched, and L x conj(R) is the conjugate
    np.testing.assert_allclose(st.autos[0, 0], st.autos[0, 1], rtol=1e-12)
    np.testing.assert_allclose(st.autos[1, 1], np.conj(st.autos[1, 0]), rtol=1e-12)


def test_pol_delta_pin_catches_one_phasor_for_both_pols(monkeypatch):
    # the mistake of applying +delta/2 to every band: cross-hand phase stays zero
    monkeypatch.setattr(fo, "pol_halfdelta", lambda pol, hd: hd)
    err, _ = _crosshand_phase_error(0.004)
    assert err > 0.5


def test_pol_delta_sign_follows_R_not_band_order():
    """With the L band recorded first the R band still takes +delta/2 (mode.cpp:1257)."""
    cfg, data, ctrl = _dual_pol_station(pols="LR")
    cfg.datastreams[0].clock_offset_delta_us = [0.004]
    st = fo.process_station(cfg, 0, OFFSETS, data[0], ctrl[0])
    # band 0 is L, band 1 is R: L x conj(R) = exp(-i w delta)
    lr = st.autos[1, 0]
    want = -_chan_rad_per_us(cfg) * 0.004
    ok = np.abs(lr) > 0
    assert np.max(np.abs(np.angle(lr[ok] * np.exp(-1j * want[ok])))) < 1e-9


# ------------------------------------------------------------------- pol phase offset
@pytest.mark.parametrize("pols", ["RL", "LR"])
def test_pol_phase_offset_turns_the_second_pol(pols):
    """phasecorrA = exp(+i phi/2) on the L spectra (mode.cpp:421-447, 1318-1330): with equal
    R and L samples R x conj(L) has phase -phi/2 in every channel."""
    phi = 50.0
    cfg, data, ctrl = _dual_pol_station(pols=pols)
    cfg.datastreams[0].phase_offset_deg = [phi]
    st = fo.process_station(cfg, 0, OFFSETS, data[0], ctrl[0])
    r, l = (0, 1) if pols == "RL" else (1, 0)
    rl = st.autos[1, r]
    ok = np.abs(rl) > 0
    assert np.max(np.abs(np.angle(rl[ok] * np.exp(1j * math.radians(phi / 2))))) < 1e-9
    np.testing.assert_allclose(st.autos[0, 0], st.autos[0, 1], rtol=1e-12)
    # the R spectra are those without the offset
    cfg.datastreams[0].phase_offset_deg = [0.0]
    st0 = fo.process_station(cfg, 0, OFFSETS, data[0], ctrl[0])
    np.testing.assert_array_equal(st.spectra[r], st0.spectra[r])
    np.testing.assert_allclose(st.spectra[l], st0.spectra[l] * np.exp(1j * math.radians(phi / 2)),
                               rtol=1e-12, atol=1e-12 * np.max(np.abs(st0.spectra[l])))


def test_pol_phase_offset_needs_both_pols():
    """A freq with one recorded band is not turned (count > 1, mode.cpp:1299)."""
    cfg = cf.standard_config(2, 1, [4800.0], 32.0, 256, subint_ns=4 * 8000, seed=22)
    data, ctrl = synth.noise_slot(cfg, OFFSETS, seed=22)
    st0 = fo.process_station(cfg, 0, OFFSETS, data[0], ctrl[0])
    cfg.datastreams[0].recbands = [(0, "L")]
    cfg.datastreams[0].phase_offset_deg = [50.0]
    st1 = fo.process_station(cfg, 0, OFFSETS, data[0], ctrl[0])
    np.testing.assert_array_equal(st0.spectra, st1.spectra)


# ------------------------------------------------------------------------ clock offset
def test_clock_offset_equals_delay_model_shift():
    """A clock offset c shifts only the baseband delay (the fractional-sample phasors), a model
    delay change shifts the fringe phase as well, by c x lofreq turns.  With c lofreq a whole
    number of turns, and FFTs whose nearest sample does not move, the two give the same
    spectra to FP64 rounding; a sign mistake in either term breaks the equality."""
    lofreq = 4000.0
    c = 1.0 / lofreq                          # 0.25 ns: one turn at the LO
    cfg = cf.standard_config(2, 1, [lofreq], 32.0, 256, subint_ns=16 * 8000, seed=23)
    for d in cfg.datastreams:            # a small delay keeps the FP64 phase rounding small
        d.model[:] = 0.0
        d.model[0, :2] = [12.3456, 0.4]
    data, ctrl = synth.noise_slot(cfg, OFFSETS, seed=23)
    B = cfg.blockspersend()
    cfg.datastreams[0].clock_offset_us = [c]
    st_clk = fo.process_station(cfg, 0, OFFSETS, data[0], ctrl[0])
    cfg.datastreams[0].clock_offset_us = [0.0]
    st_base = fo.process_station(cfg, 0, OFFSETS, data[0], ctrl[0])
    cfg.datastreams[0].model = cfg.datastreams[0].model.copy()
    cfg.datastreams[0].model[:, 0] += c
    st_mod = fo.process_station(cfg, 0, OFFSETS, data[0], ctrl[0])

    def nearest(model_shift):
        ds = cfg.datastreams[0]
        f0 = cfg.freqs[0]
        interp = fo.delay_interpolator(ds.model, ds.model_order, cfg.model_inc_secs, 1.0,
                                       B * f0.nchan / f0.bw_mhz / 1e6, B).copy()
        interp[2] -= model_shift
        return fo.fft_params(cfg, 0, OFFSETS, ctrl[0], interp, B)[0]

    same = nearest(0.0) == nearest(c)
    assert same.sum() >= B // 2
    scale = np.max(np.abs(st_mod.spectra))
    np.testing.assert_allclose(st_clk.spectra[0, same], st_mod.spectra[0, same], rtol=0, atol=1e-9 * scale)
    # and the offset does change the spectra: a phase slope of -w c across the band
    assert np.max(np.abs(st_clk.spectra[0, same] - st_base.spectra[0, same])) > 1e-3 * scale
    ratio = st_clk.spectra[0, same][:, 1:] / st_base.spectra[0, same][:, 1:]
    np.testing.assert_allclose(np.angle(ratio), np.broadcast_to(-_chan_rad_per_us(cfg)[1:] * c, ratio.shape),
                               atol=1e-9)


# ------------------------------------------------------------------------- LO offset
def _lo_offset_fft_phases(dlo_hz, scan_start=1000.0):
    cfg = cf.standard_config(2, 1, [4800.0], 32.0, 256, subint_ns=8 * 8000, seed=24,
                             scan_start_sec=scan_start)
    data, ctrl = synth.noise_slot(cfg, OFFSETS, seed=24)
    cfg.datastreams[1] = cfg.datastreams[0]                  # the same station twice
    data[1], ctrl[1] = data[0], ctrl[0]
    st0 = fo.process_station(cfg, 0, OFFSETS, data[0], ctrl[0])
    ds1 = cf.Datastream(**{**cfg.datastreams[0].__dict__})
    ds1.lo_offset_hz = [dlo_hz]
    cfg.datastreams[1] = ds1
    st1 = fo.process_station(cfg, 1, OFFSETS, data[1], ctrl[1])
    L, dt = cfg.fftchannels(0), cfg.sampletime_us(0)
    B = cfg.blockspersend()
    got = np.angle(np.sum(st0.spectra[0] * np.conj(st1.spectra[0]), axis=1))
    # wall-clock time of each FFT's centre (s)
    t_mid = cfg.scan_start_sec + OFFSETS[1] + OFFSETS[2] / 1e9 + (np.arange(B) * L + L / 2) * dt / 1e6
    return got, t_mid


@pytest.mark.parametrize("dlo", [37.25, -2.5, 10.0])
def test_lo_offset_fringe_phase_follows_walltime(dlo):
    """An LO offset of df Hz on one of two identical stations gives a cross phase of
    -2 pi df t at wall-clock time t (integer, fractional-second and fractional-Hz terms of
    mode.cpp:910-931 together); at t ~ 1001 s the fractional-Hz term alone is 0.25 turn for
    37.25 Hz, and the phase advances by -2 pi df dt between FFTs."""
    got, t_mid = _lo_offset_fft_phases(dlo)
    want = -fo.TWO_PI * ((dlo * t_mid) % 1.0)
    err = np.angle(np.exp(1j * (got - want)))
    assert np.max(np.abs(err)) < 5e-3


def test_lo_offset_pin_catches_a_dropped_fractional_hz_term(monkeypatch):
    real = fo.rotator_phase

    def no_frac_hz(order, L, S, a, b, c, lofreq, intdelay, fractional_lo, loff, st, intwall, fracwall):
        ph = real(order, L, S, a, b, c, lofreq, intdelay, fractional_lo, loff, st, intwall, fracwall)
        fraclo = math.copysign(abs(loff) - int(abs(loff)), loff)
        return ph + fraclo * intwall                    # undo the fractional-Hz term
    monkeypatch.setattr(fo, "rotator_phase", no_frac_hz)
    got, t_mid = _lo_offset_fft_phases(37.25)
    want = -fo.TWO_PI * ((37.25 * t_mid) % 1.0)
    assert np.min(np.abs(np.angle(np.exp(1j * (got - want))))) > 1.0


# ------------------------------------------- fractional LO of double-sideband data
def test_fractional_lo_hand_cases():
    def ds_of(edge, bw, cplx, lsb=False):
        cfg = cf.standard_config(1, 1, [edge], bw, 256, subint_ns=int(256 / bw * 1000) * 4, nbits=8,
                                 complex_sampling=cplx, lsb=lsb)
        return cfg, cfg.datastreams[0]
    # real and single-sideband complex: the band edge
    assert fo.is_fractional_lo(*ds_of(4800.5, 16.0, 0))
    assert not fo.is_fractional_lo(*ds_of(4800.0, 5.0, 1))
    # double sideband: the centre, edge + BW/2 (USB) or edge - BW/2 (LSB)
    assert fo.is_fractional_lo(*ds_of(4800.0, 5.0, 2))            # centre 4802.5
    assert not fo.is_fractional_lo(*ds_of(4800.5, 5.0, 2))        # centre 4803
    assert not fo.is_fractional_lo(*ds_of(4800.5, 5.0, 2, lsb=True))   # centre 4798
    assert not fo.is_fractional_lo(*ds_of(4800.0, 16.0, 2))


def test_dsb_rotator_phase_is_total_delay_times_lo():
    """The rotator phase of every sample is -(total delay) x lofreq turns modulo 1, including
    the integer microseconds split off as integerdelay (the fractional-LO term of mode.cpp:
    900-903).  For a double-sideband band with a whole edge and a half-MHz centre the raw-edge
    test would drop a half turn per odd integerdelay."""
    cfg = cf.standard_config(2, 1, [4800.0], 5.0, 256, subint_ns=int(256 / 5.0 * 1000) * 4, nbits=8,
                             complex_sampling=2, seed=25)
    ds = cfg.datastreams[0]
    ds.model[:] = 0.0
    ds.model[0, 0] = 1234.6                      # 1235.3 us at t = 1 s: odd integerdelay
    ds.model[0, 1] = 0.7
    B = cfg.blockspersend()
    L, S = cfg.fftchannels(0), ds.arraystridelen or fo.calc_stride(256)
    interp = fo.delay_interpolator(ds.model, ds.model_order, cfg.model_inc_secs, 1.0,
                                   B * 256 / 5.0 / 1e6, B)
    ctrl = np.array([0, 1, 0], np.int32)
    ns, st, a, b, idl, _, c = fo.fft_params(cfg, 0, OFFSETS, ctrl, interp, B)
    lofreq = 4800.0 + 5.0 / 2.0
    frac_lo = fo.is_fractional_lo(cfg, ds)
    assert frac_lo
    for i in range(B):
        assert idl[i] % 2 == 1
        ph = fo.rotator_phase(1, L, S, a[i], b[i], c[i], lofreq, int(idl[i]), frac_lo, 0.0,
                              cfg.sampletime_us(0), 1, 0.0)
        total = (a[i] * np.arange(L) / L + b[i] + idl[i]) * lofreq
        d = (ph - total) % 1.0
        d = np.minimum(d, 1.0 - d)
        assert np.max(d) < 1e-7
        bad = fo.rotator_phase(1, L, S, a[i], b[i], c[i], lofreq, int(idl[i]), False, 0.0,
                               cfg.sampletime_us(0), 1, 0.0)
        d = (bad - total) % 1.0
        assert np.min(np.minimum(d, 1.0 - d)) > 0.4
