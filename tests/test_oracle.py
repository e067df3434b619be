"""Pins of the FP64 oracle (CPU only): each check is chosen so that a plausible mistake in the
restatement (a dropped term, a wrong sign or index, a transposed operand) fails one of them."""
import math

import numpy as np
import pytest

import fxoracle as fo
from difx_amd import config as cf
from difx_amd import synth


# ---------------------------------------------------------------- LUTs (format_vdif.c:66-145)
def test_vdif_lut_known_bytes():
    lut = fo.vdif_lut2bit()
    H = fo.HIMAG
    assert list(lut[0x00]) == [-H] * 4 or np.allclose(lut[0x00], -H)
    assert np.allclose(lut[0xFF], H)
    # 0xE4 = 11 10 01 00: fields LSB first are 0,1,2,3 -> -H, -1, +1, +H
    np.testing.assert_array_equal(lut[0xE4], np.float32([-H, -1, 1, H]))
    # 0x1B = 00 01 10 11: fields 3,2,1,0
    np.testing.assert_array_equal(lut[0x1B], np.float32([H, 1, -1, -H]))


def test_lut_complement_symmetry():
    # complementing every bit maps field l -> 3-l, which negates both LUTs
    for lut in (fo.vdif_lut2bit(), fo.mark5b_lut2bit()):
        for b in range(256):
            np.testing.assert_array_equal(lut[b ^ 0xFF], -lut[b])


def test_mark5b_lut_sign_magnitude():
    lut = fo.mark5b_lut2bit()
    H = np.float32(fo.HIMAG)
    # sample 0 = bits (sign=bit0, mag=bit1): sign 1 -> positive; mag 0 -> high magnitude
    assert lut[0b00][0] == -H and lut[0b01][0] == 1.0 and lut[0b10][0] == -1.0 and lut[0b11][0] == H
    # sample 3 sits in bits 6,7
    assert lut[0b01000000][3] == 1.0 and lut[0b10000000][3] == -1.0


def test_lut8():
    lut = fo.lut8bit()
    assert lut[128] == 0.0 and lut[161] == np.float32(33 / 3.3) and lut[0] == np.float32(-128 / 3.3)


# ------------------------------------------------------------ frames and unpack (mark5access)
def _ds(nbands=1, fmt="VDIF", nbits=2, cplx=False, payload=64):
    recb = [(0, "R")] * nbands
    hb = 32 if fmt == "VDIF" else 16
    return cf.Datastream(recfreqs=[0], recbands=recb, model=np.zeros((1, 6)), format=fmt,
                         nbits=nbits, complex_sampling=cplx, payloadbytes=payload, headerbytes=hb)


def test_frame_skip_and_blanking():
    ds = _ds()
    spf = synth.samples_per_frame(ds)           # 64 bytes * 4 samples
    assert spf == 256
    pay = np.arange(3 * 64, dtype=np.uint8)
    buf = synth.frame_payloads(pay, ds, 0)
    vals, valid = fo.decode_stream(buf, ds)
    lut = fo.vdif_lut2bit()
    # sample 300 = frame 1, byte 11, field 0: payload byte 64+11 = 75
    assert vals[0, 300] == lut[75][0]
    assert vals[0, 256 + 4 * 63 + 3] == lut[127][3]
    # invalid bit on frame 1 and fill pattern at the end of frame 2
    b2 = buf.copy()
    b2[96 + 3] |= 0x80
    b2[2 * 96 + 96 - 8: 2 * 96 + 96] = np.frombuffer(np.uint64(fo.FILL_WORD64).tobytes(), np.uint8)
    v2, ok = fo.decode_stream(b2, ds)
    assert ok[:256].all() and not ok[256:].any()
    assert np.all(v2[0, 256:] == 0) and not np.any(np.signbit(v2[0, 256:]))
    np.testing.assert_array_equal(v2[0, :256], vals[0, :256])


def test_multiband_field_order():
    ds = _ds(nbands=16, payload=64)
    pay = np.zeros(64, np.uint8)
    pay[0] = 0b11100100          # time 0: bands 0..3 = fields 0,1,2,3
    pay[7] = 0b11000000          # time 1 (bytes 4..7): band 15 = field 3
    buf = synth.frame_payloads(pay, ds, 0)
    vals, _ = fo.decode_stream(buf, ds)
    H = np.float32(fo.HIMAG)
    np.testing.assert_array_equal(vals[:4, 0], [-H, -1, 1, H])
    assert vals[15, 1] == H and vals[14, 1] == -H


def test_pack_roundtrip():
    rng = np.random.default_rng(0)
    codes = rng.integers(0, 4, size=(1024, 4)).astype(np.uint8)
    ds = _ds(nbands=4, payload=1024)
    buf = synth.frame_payloads(synth.pack_2bit(codes), ds, 0)
    vals, _ = fo.decode_stream(buf, ds)
    lev = np.float32([-fo.HIMAG, -1, 1, fo.HIMAG])
    np.testing.assert_array_equal(vals, lev[codes].T)


# --------------------------------------------------------- delay model (model.cpp:219-289)
def test_interpolator_exact_for_quadratics():
    rng = np.random.default_rng(1)
    for _ in range(5):
        c = np.zeros((1, 6))
        c[0, :3] = rng.uniform(-1e3, 1e3), rng.uniform(-1, 1), rng.uniform(-1e-3, 1e-3)
        t0, span, n = 7.25, 0.5, 125
        q = fo.delay_interpolator(c, 5, 120.0, t0, span, n)
        for x in (0.0, 0.5, 17.3, 124.0, 125.0):
            t = t0 + x * span / n
            exact = c[0, 0] + c[0, 1] * t + c[0, 2] * t * t
            assert abs(q[0] * x * x + q[1] * x + q[2] - exact) < 1e-9


def test_interpolator_out_of_model():
    assert fo.delay_interpolator(np.zeros((2, 6)), 5, 120.0, 300.0, 0.5, 10) is None


# ---------------------------------------------------- per-FFT delay chain (mode.cpp:671-788)
def _cfg_small(B=16, nchan=256, delay=0.0, lofreq=4800.0):
    cfg = cf.standard_config(2, 1, [lofreq], 32.0, nchan, subint_ns=B * 1000 * nchan // 32)
    for d in cfg.datastreams:
        d.model[:] = 0.0
        d.model[0, 0] = delay
    return cfg


def test_zero_delay_chain():
    cfg = _cfg_small()
    ctrl = np.array([0, 1, 0], np.int32)
    interp = np.zeros(3)
    ns, st, a, b, idl, _, _ = fo.fft_params(cfg, 0, [0, 1, 0], ctrl, interp, 16)
    L = cfg.fftchannels(0)
    np.testing.assert_array_equal(ns, np.arange(16) * L)
    np.testing.assert_allclose(st - ns * cfg.sampletime_us(0), 0, atol=1e-12)
    assert np.all(a == 0) and np.all(b == 0) and np.all(idl == 0)


def test_integer_sample_delay_chain():
    cfg = _cfg_small()
    dt = cfg.sampletime_us(0)
    interp = np.array([0.0, 0.0, 37 * dt])        # 37 samples of delay
    ctrl = np.array([0, 0, 999000000], np.int32)  # buffer starts 1 ms before the subint
    ns, st, a, b, idl, _, _ = fo.fft_params(cfg, 0, [0, 1, 0], ctrl, interp, 8)
    L = cfg.fftchannels(0)
    n1ms = int(round(1000.0 / dt))
    np.testing.assert_array_equal(ns, n1ms + np.arange(8) * L - 37)
    np.testing.assert_allclose(st - ns * dt, 0.0, atol=1e-9)
    # int(x + 0.5) truncates toward zero: -37 samples before the buffer start gives -36
    ctrl0 = np.array([0, 1, 0], np.int32)
    ns0, *_ = fo.fft_params(cfg, 0, [0, 1, 0], ctrl0, interp, 1)
    assert ns0[0] == -36


def test_rotator_unit_phase_cases():
    # integer-microsecond delay with an integer-MHz LO leaves the fringe phase unchanged:
    # the spectra of a delayed copy equal those of the undelayed one shifted in samples
    for delay_us in (0.0, 3.0):
        cfg = _cfg_small(B=4, delay=delay_us)
        offsets = [0, 1, 0]
        data, ctrl = synth.noise_slot(cfg, offsets, seed=4)
        st = fo.process_station(cfg, 0, offsets, data[0], ctrl[0], 4)
        # recompute with the delay and the phase terms removed: the rotator must be identity
        interp = np.array([0.0, 0.0, delay_us])
        ns, stt, a, b, idl, _, _ = fo.fft_params(cfg, 0, offsets, ctrl[0], interp, 4)
        assert np.all(b == 0) and np.all(idl == int(delay_us))
        vals, _ = fo.decode_stream(data[0], cfg.datastreams[0])
        L, C = cfg.fftchannels(0), 256
        for i in range(4):
            X = np.fft.fft(vals[0, ns[i]: ns[i] + L].astype(np.complex128))[:C]
            np.testing.assert_allclose(st.spectra[0, i], X, rtol=1e-9, atol=1e-9)


def test_order2_rotator_tracks_quadratic_delay():
    """Order 2 (mode.cpp:800-825, 956-1026) follows the delay q(i + n/L) of a curved model to
    within the bound of its sub/step split, a (S/L)^2 / 4 delay units (the sub term uses
    subxoff * S/L where the exact square has subxoff^2); order 1's straight line misses by
    far more."""
    L, S, lofreq = 2048, 32, 4800.0
    q = np.array([3.0e-7, 2.0e-3, 1234.5678])                 # strongly curved, FFT-index units
    for i in (0, 17, 500):
        x = i + np.arange(L) / L
        truth = q[0] * x * x + q[1] * x + q[2]
        a2, b2, c2 = q[0], q[1] + 2 * i * q[0], q[2] + i * q[1] + i * i * q[0]
        idl2 = int(c2)
        ph2 = fo.rotator_phase(2, L, S, a2, b2, c2 - idl2, lofreq, idl2, False, 0.0, 1 / 64, 0, 0.0)
        err2 = np.max(np.abs(ph2 - lofreq * (truth - idl2)))
        bound = q[0] * (S / L) ** 2 / 4 * lofreq
        rounding = 8 * np.finfo(float).eps * lofreq * abs(truth).max()   # FP64 of ~6e6 turns
        assert err2 <= bound + rounding
        d0, d1, d2 = [q[0] * y * y + q[1] * y + q[2] for y in (i, i + 0.5, i + 1.0)]
        a1 = d2 - d0
        b1 = d0 + (d1 - (a1 * 0.5 + d0)) / 3.0
        idl1 = int(b1)
        ph1 = fo.rotator_phase(1, L, S, a1, b1 - idl1, 0.0, lofreq, idl1, False, 0.0, 1 / 64, 0, 0.0)
        err1 = np.max(np.abs(ph1 - lofreq * (truth - idl1)))
        assert err1 > 100 * bound


def _constant_delay_cfg(order, lsb, delay_us=2.3456, lofreq=4800.37):
    cfg = cf.standard_config(2, 1, [lofreq], 32.0, 256, subint_ns=8 * 8000, seed=14, lsb=lsb,
                             fringe_rot_order=order)
    for k, d in enumerate(cfg.datastreams):
        d.model[:] = 0.0
        d.model[0, 0] = delay_us + 0.5 * k
    return cfg


@pytest.mark.parametrize("lsb", [False, True])
def test_order0_equals_order1_for_constant_delay(lsb):
    """A constant delay makes the order-1 rotator a constant phase (b lofreq plus the
    fractional-LO term), which order 0 applies after the FFT instead (mode.cpp:1040-1049);
    for real data the LSB conj-flip of bins 1..C (mode.cpp:1139-1156) equals the upper half
    of the complex FFT, so both orders give the same spectra."""
    offsets = [0, 1, 0]
    res = {}
    for order in (0, 1):
        cfg = _constant_delay_cfg(order, lsb)
        data, ctrl = synth.noise_slot(cfg, offsets, seed=15)
        res[order], _ = fo.process_subint(cfg, offsets, data, ctrl)
    np.testing.assert_allclose(res[0], res[1], rtol=0, atol=1e-9 * np.max(np.abs(res[1])))


def test_order2_equals_order1_for_linear_delay():
    """With a delay linear across the subint the quadratic coefficient vanishes and order 2
    (c + b x) and order 1 (b + a x) describe the same line."""
    offsets = [0, 1, 0]
    res = {}
    for order in (1, 2):
        cfg = cf.standard_config(2, 2, [4800.37, 4832.0], 32.0, 256, subint_ns=8 * 8000, seed=16,
                                 fringe_rot_order=order)
        for k, d in enumerate(cfg.datastreams):
            d.model[:, 2:] = 0.0
        data, ctrl = synth.noise_slot(cfg, offsets, seed=17)
        res[order], _ = fo.process_subint(cfg, offsets, data, ctrl)
    # order 1 takes its slope as d2 - d0 of delays near 1.2e3 us: FP64 cancellation leaves a
    # relative slope error near 3e-8, hence a 1e-7 bound
    np.testing.assert_allclose(res[2], res[1], rtol=0, atol=1e-7 * np.max(np.abs(res[1])))


def _complex_tone_station(sideband: str, f_bb_mhz: float, B: int = 4, conj: bool = False):
    """One 16 MHz complex 8-bit datastream (nchan 256) with zero delay carrying the baseband
    tone exp(2 pi i f t) (or its conjugate); returns the oracle's station spectra."""
    lsb = sideband in ("lsb", "dlsb")
    cfg = cf.standard_config(1, 1, [1400.0], 16.0, 256, subint_ns=B * 16000, nbits=8,
                             complex_sampling=2 if sideband in ("dusb", "dlsb") else 1, lsb=lsb)
    cfg.datastreams[0].model[:] = 0.0
    offsets = [0, 1, 0]
    ds = cfg.datastreams[0]
    start = synth.buffer_start_ns(cfg, 0, offsets)
    nbytes = synth.slot_bytes(cfg, 0, offsets, start)
    nsamp = (nbytes // ds.framebytes) * synth.samples_per_frame(ds)
    t_us = start / 1000.0 + np.arange(nsamp) / 16.0
    z = np.exp(2j * np.pi * f_bb_mhz * t_us)
    if conj:
        z = np.conj(z)
    pay = np.empty(2 * nsamp, np.uint8)
    pay[0::2] = np.round(128 + 40 * z.real).astype(np.uint8)
    pay[1::2] = np.round(128 + 40 * z.imag).astype(np.uint8)
    spf = synth.samples_per_frame(ds)
    data = synth.frame_payloads(pay, ds, start // int(round(spf * 1000.0 / 16.0)))
    ctrl = synth.control_words(cfg, 0, offsets, start)
    return fo.process_station(cfg, 0, offsets, data, ctrl).spectra[0]


@pytest.mark.parametrize("sideband,f_bb,chan", [
    ("usb", 4.0, 64),          # sky LO + 4 MHz, channel 0 at the LO
    ("lsb", 4.0, 192),         # sky LO - 4 MHz, channel 0 at LO - 16 MHz
    ("dusb", 4.0, 192),        # sky LO + 4 MHz, channel 0 at LO - 8 MHz
    ("dusb", -4.0, 64),
    ("dlsb", 4.0, 64),         # sky LO - 4 MHz, channel 0 at LO - 8 MHz
    ("dlsb", -4.0, 192)])
def test_complex_sideband_channel_order(sideband, f_bb, chan):
    """Channel 0 is the lowest sky frequency for every complex sideband type
    (mode.cpp:1224-1235): a baseband tone lands in the channel its sky frequency names."""
    sp = _complex_tone_station(sideband, f_bb)
    peak = np.argmax(np.abs(sp).sum(0))
    assert peak == chan
    assert np.abs(sp[:, chan]).min() > 50 * np.median(np.abs(sp))


def test_complex_lsb_is_usb_of_the_conjugate():
    """LSB processing conj-flips all channels but DC (mode.cpp:1199-1200), so the LSB spectra
    of conj(x) equal the USB spectra of x except channel 0, which is conjugated."""
    for f in (1.3, -5.1):
        u = _complex_tone_station("usb", f)
        l = _complex_tone_station("lsb", f, conj=True)
        np.testing.assert_allclose(l[:, 1:], u[:, 1:], rtol=0, atol=1e-9 * np.abs(u).max())
        np.testing.assert_allclose(l[:, 0], np.conj(u[:, 0]), rtol=0, atol=1e-9 * np.abs(u).max())


# ------------------------------------------------------------------------------ FFT
@pytest.mark.parametrize("L", [64, 256, 512])
def test_fft_against_bruteforce_dft(L):
    rng = np.random.default_rng(L)
    x = rng.standard_normal(L) + 1j * rng.standard_normal(L)
    n = np.arange(L)
    D = np.exp(-2j * np.pi * np.outer(n, n) / L)
    np.testing.assert_allclose(np.fft.fft(x), D @ x, rtol=1e-10, atol=1e-9)


# ------------------------------------------------------------- layout (configuration.cpp)
def test_layout_single_pol_hand_count():
    cfg = cf.standard_config(2, 1, [4800.0], 32.0, 8, subint_ns=8 * 250)
    lay = fo.result_layout(cfg)
    assert lay.baseline_off == {(0, 0): 0}
    assert lay.bweight_off == {(0, 0): 8}
    assert lay.autocorr_off == [9, 17] and lay.acweight_off == [25, 26]
    assert lay.coreresultlength == 27 and lay.bandsperautocorr == 1


def test_layout_dual_pol_hand_count():
    cfg = cf.standard_config(2, 2, [4800.0], 32.0, 8, subint_ns=8 * 250)
    lay = fo.result_layout(cfg)
    assert lay.baseline_off == {(0, 0): 0}
    assert lay.bweight_off == {(0, 0): 32}
    assert lay.autocorr_off == [34, 66] and lay.acweight_off == [98, 100]
    assert lay.coreresultlength == 102 and lay.bandsperautocorr == 2


def test_layout_zoom_hand_count():
    """One 8-channel parent (4 MHz channels) with a 2-channel zoom at channel offset 2, both
    correlated, single pol, 2 stations (configuration.cpp:2464-2621): crosses 8 + 2,
    bweights 1 + 1, autos (8 + 2) per station, acweights (2 + 1) // 2 = 1 per station."""
    cfg = cf.batched_config(2, [(32.0, 8, [4800.0])], [(0, 0, 8.0, 8.0)], subint_ns=8 * 250, npol=1)
    assert cfg.freqs[1].nchan == 2 and cfg.datastreams[0].zoomfreqs == [(1, 0, 2)]
    lay = fo.result_layout(cfg)
    assert lay.baseline_off == {(0, 0): 0, (1, 0): 8}
    assert lay.bweight_off == {(0, 0): 10, (1, 0): 11}
    assert lay.autocorr_off == [12, 22] and lay.acweight_off == [32, 33]
    assert lay.coreresultlength == 34


def test_zoom_records_are_parent_windows():
    """A zoom band is a window of its parent's channels (mode.cpp:184-196): every zoom
    record, cross and auto, equals the matching channels of the parent record."""
    cfg = cf.batched_config(2, [(32.0, 256, [4800.0]), (8.0, 64, [4900.0])],
                            [(0, 0, 4.0, 2.0), (1, 0, 1.0, 1.0)], subint_ns=8 * 8000, seed=12)
    offsets = [0, 1, 0]
    data, ctrl = synth.noise_slot(cfg, offsets, seed=13)
    res, _ = fo.process_subint(cfg, offsets, data, ctrl)
    lay = fo.result_layout(cfg)
    checked = 0
    for j, bl in enumerate(cfg.baselines):
        d1 = cfg.datastreams[bl.ds1]
        for zfi, plf, choff in d1.zoomfreqs:
            pfi = d1.recfreqs[plf]
            n, pn = cfg.freqs[zfi].nchan, cfg.freqs[pfi].nchan
            for p in range(4):
                z = res[lay.baseline_off[(zfi, j)] + p * n:][:n]
                par = res[lay.baseline_off[(pfi, j)] + p * pn + choff:][:n]
                np.testing.assert_array_equal(z, par)
                checked += 1
    # autos: datastream 0 (group 0) holds RR, LL of the parent then of the zoom, then cross-pol
    d0 = cfg.datastreams[0]
    a0 = lay.autocorr_off[0]
    _, _, choff = d0.zoomfreqs[0]
    n = cfg.freqs[d0.zoomfreqs[0][0]].nchan
    for w in range(2):                     # parallel block, cross-pol block
        base = a0 + w * (2 * 256 + 2 * n)
        for pol in range(2):
            np.testing.assert_array_equal(res[base + 2 * 256 + pol * n:][:n], res[base + pol * 256 + choff:][:n])
    assert checked == 8


# ---------------------------------------------------------- X-engine / accumulation invariants
def test_xmac_invariants():
    cfg = cf.standard_config(3, 2, [4800.0], 32.0, 256, subint_ns=8 * 8000, seed=5)
    offsets = [0, 1, 0]
    data, ctrl = synth.noise_slot(cfg, offsets, seed=6)
    # station 2 records exactly what station 0 records (same model, same bytes)
    cfg.datastreams[2].model = cfg.datastreams[0].model.copy()
    data[2], ctrl[2] = data[0], ctrl[0]
    res, st = fo.process_subint(cfg, offsets, data, ctrl)
    lay = fo.result_layout(cfg)
    # autos real and non-negative
    for d in range(3):
        a = res[lay.autocorr_off[d]: lay.autocorr_off[d] + 2 * 256]
        assert np.all(a.real >= 0) and np.max(np.abs(a.imag)) < 1e-9 * np.max(a.real)
    # baseline (0,2) of identical stations = station 0's autocorrelation (RR and LL)
    j = [k for k, b in enumerate(cfg.baselines) if (b.ds1, b.ds2) == (0, 2)][0]
    off = lay.baseline_off[(0, j)]
    np.testing.assert_allclose(res[off: off + 256], res[lay.autocorr_off[0]: lay.autocorr_off[0] + 256], rtol=1e-12)
    # RL product of (0,2) = station 0's cross-pol autocorrelation of band R
    np.testing.assert_allclose(res[off + 2 * 256: off + 3 * 256],
                               res[lay.autocorr_off[0] + 2 * 256: lay.autocorr_off[0] + 3 * 256], rtol=1e-12)
    # Hermitian: LR of (0,1) equals conj of RL with the stations swapped, i.e. the spectra give
    # V_ij = conj(V_ji)
    s0, s1 = st[0], st[1]
    v01 = np.einsum("tc,tc->c", s0.spectra[0], np.conj(s1.spectra[0]))
    v10 = np.einsum("tc,tc->c", s1.spectra[0], np.conj(s0.spectra[0]))
    np.testing.assert_allclose(v01, np.conj(v10), rtol=1e-12)


def test_flagged_ffts_have_zero_weight_and_spectra():
    cfg = cf.standard_config(2, 1, [4800.0], 32.0, 256, subint_ns=16 * 8000, seed=7)
    offsets = [0, 1, 0]
    data, ctrl = synth.noise_slot(cfg, offsets, seed=8)
    start = synth.buffer_start_ns(cfg, 0, offsets)
    ctrl[0] = synth.control_words(cfg, 0, offsets, start, invalid_ffts=[3, 7])
    st = fo.process_station(cfg, 0, offsets, data[0], ctrl[0])
    assert st.weight[3] == 0 and st.weight[7] == 0 and st.weight[0] == 1.0
    assert np.all(st.spectra[0, 3] == 0) and np.all(st.spectra[0, 7] == 0)
    # scan mismatch torches the whole subint (mode.cpp:1511-1515)
    st2 = fo.process_station(cfg, 0, [5, 1, 0], data[0], ctrl[0])
    assert np.all(st2.weight == 0)


def test_chanavg_means():
    cfg = cf.standard_config(2, 1, [4800.0], 32.0, 256, subint_ns=8 * 8000, seed=9)
    offsets = [0, 1, 0]
    data, ctrl = synth.noise_slot(cfg, offsets, seed=10)
    full, _ = fo.process_subint(cfg, offsets, data, ctrl)
    cfg.freqs[0].chanavg = 4
    avg, _ = fo.process_subint(cfg, offsets, data, ctrl)
    l1, l4 = fo.result_layout(cfg), None
    np.testing.assert_allclose(avg[0:64], full[0:256].reshape(64, 4).mean(1), rtol=1e-12)


# --------------------------------------------- end to end against the 2-bit closed form
def test_end_to_end_coherence_closed_form():
    """A common signal of relative amplitude a in unit noise, 2-bit quantised with +-0.95
    sigma thresholds, correlates at the coefficient of the quantised time series; the FX
    path must recover it flat in phase once the model delays are removed."""
    amp = 0.3
    cfg = cf.standard_config(3, 1, [4800.0], 32.0, 256, subint_ns=1500 * 8000, seed=3)
    rng = np.random.default_rng(5)
    for d in cfg.datastreams:
        d.model[:] = 0.0
        d.model[0, 0] = rng.uniform(-30.0, 30.0)
    offsets = [0, 1, 0]
    data, ctrl = synth.correlated_slot(cfg, offsets, seed=7, amp=amp)
    res, _ = fo.process_subint(cfg, offsets, data, ctrl)
    lay = fo.result_layout(cfg)
    # closed form: correlation coefficient of the quantised time series (no FX machinery)
    r2 = np.random.default_rng(9)
    n = 2_000_000
    s = r2.standard_normal(n)
    sd = math.sqrt(1 + amp * amp)
    lev = np.array([-fo.HIMAG, -1.0, 1.0, fo.HIMAG])
    q1 = lev[synth.quantize_2bit((r2.standard_normal(n) + amp * s) / sd)]
    q2 = lev[synth.quantize_2bit((r2.standard_normal(n) + amp * s) / sd)]
    rho_q = float(np.corrcoef(q1, q2)[0, 1])
    for j, b in enumerate(cfg.baselines):
        off = lay.baseline_off[(0, j)]
        v = res[off: off + 256][8:-8]
        a1 = res[lay.autocorr_off[b.ds1]: lay.autocorr_off[b.ds1] + 256].real[8:-8]
        a2 = res[lay.autocorr_off[b.ds2]: lay.autocorr_off[b.ds2] + 256].real[8:-8]
        coh = v / np.sqrt(a1 * a2)
        m = coh.mean()
        assert abs(np.angle(m)) < 0.05, np.angle(m)                  # flat, zero phase
        assert abs(abs(m) - rho_q) < 0.1 * rho_q, (abs(m), rho_q)     # amplitude of the closed form
        # no residual delay: phase slope across the band below 0.02 rad/channel-band
        ph = np.angle(coh * np.conj(m))
        slope = np.polyfit(np.arange(len(ph)), ph, 1)[0]
        assert abs(slope) * len(ph) < 0.2


def test_end_to_end_wrong_sign_fails_closed_form():
    """Planted mistake in the oracle's data model: without the LO phase the fringe rotation
    leaves a station-dependent phase, so the zero-phase check above would fail."""
    amp = 0.3
    cfg = cf.standard_config(2, 1, [4800.0], 32.0, 256, subint_ns=600 * 8000, seed=3)
    for d, tau in zip(cfg.datastreams, (12.345, -7.891)):
        d.model[:] = 0.0
        d.model[0, 0] = tau
    offsets = [0, 1, 0]
    data, ctrl = synth.correlated_slot(cfg, offsets, seed=7, amp=amp, lo_phase=False)
    res, _ = fo.process_subint(cfg, offsets, data, ctrl)
    lay = fo.result_layout(cfg)
    v = res[lay.baseline_off[(0, 0)]: lay.baseline_off[(0, 0)] + 256][8:-8]
    assert abs(np.angle(v.mean())) > 0.1
