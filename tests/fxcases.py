"""Small seeded cases shaped like the benchmark configurations, and parity metrics."""
from __future__ import annotations

import numpy as np

from difx_amd import config as cf
from difx_amd import synth


def case_cfg0(B: int = 128, xengine: str = "fp32", lsb: bool = False, seed: int = 1, **kw):
    """4 stations, single pol, 1 x 32 MHz real 2-bit VDIF, nchan 256 (cfg0 shape)."""
    return cf.standard_config(4, 1, [4800.0], 32.0, 256, subint_ns=B * 8000, xengine=xengine,
                              lsb=lsb, seed=seed, **kw)


def case_cfg1(B: int = 32, xengine: str = "fp32", nstations: int = 10, nfreq: int = 8, seed: int = 2, **kw):
    """10 stations, dual pol, 8 x 32 MHz, 16 bands per 32-bit VDIF word, nchan 1024."""
    freqs = [4800.0 + 40.0 * k for k in range(nfreq)]
    return cf.standard_config(nstations, 2, freqs, 32.0, 1024, subint_ns=B * 32000,
                              xengine=xengine, seed=seed, **kw)


def case_cfg2(B: int = 16, xengine: str = "fp16", nstations: int = 64, seed: int = 3,
              complex_sampling: int = 1, **kw):
    """64 stations, single pol, 16 MHz complex 8-bit, nchan 4096 (complex_sampling 2: double
    sideband)."""
    return cf.standard_config(nstations, 1, [1400.0], 16.0, 4096, subint_ns=B * 256000,
                              nbits=8, complex_sampling=complex_sampling, xengine=xengine, seed=seed, **kw)


def case_cfg3(B: int = 32, xengine: str = "fp32", nstations: int = 3, seed: int = 4, **kw):
    """Uneven bands (cfg3 shape, reduced): per station a 32 MHz / 512-channel, an 8 MHz /
    128-channel and a 64 MHz / 1024-channel datastream (FFT lengths 1024, 256, 2048, one
    16 us FFT duration), dual pol, plus zoom windows of 2, 1, 0.5 and 2 MHz."""
    groups = [(32.0, 512, [4800.0, 4832.0]), (8.0, 128, [4900.0, 4908.0]), (64.0, 1024, [5000.0])]
    zooms = [(0, 1, 4.0, 2.0), (0, 0, 10.0, 1.0), (1, 1, 2.0, 0.5), (2, 0, 20.0, 2.0)]
    return cf.batched_config(nstations, groups, zooms, subint_ns=B * 16000, xengine=xengine,
                             seed=seed, **kw)


def case_cfg3_full(B: int = 2, xengine: str = "fp32", nstations: int = 16, seed: int = 20261203, **kw):
    """The uneven-band configuration as the benchmark runs it, in one configuration: per
    station a 4 x 64 MHz (nchan 16384, L 32768), a 4 x 32 MHz (nchan 8192, L 16384) and an
    8 x 8 MHz (nchan 2048, L 4096) datastream, dual pol, one 256 us FFT duration, plus six zoom
    windows of 0.5-2 MHz (128-512 channels)."""
    groups = [(64.0, 16384, [4800.0 + 64.0 * k for k in range(4)]),
              (32.0, 8192, [5200.0 + 32.0 * k for k in range(4)]),
              (8.0, 2048, [5400.0 + 8.0 * k for k in range(8)])]
    zooms = [(0, 0, 10.0, 2.0), (0, 2, 30.0, 0.5), (1, 1, 4.0, 1.0), (1, 3, 20.0, 2.0),
             (2, 0, 1.0, 0.5), (2, 5, 3.0, 2.0)]
    return cf.batched_config(nstations, groups, zooms, subint_ns=B * 256000, xengine=xengine,
                             seed=seed, **kw)


def case_target16(B: int = 32, xengine: str = "fp32", seed: int = 20261205, **kw):
    """The MFMA-vs-plain target shape: 16 stations dual pol (32 X-engine inputs, one full
    32-row MFMA tile per freq), 8 x 32 MHz 2-bit VDIF, nchan 1024."""
    freqs = [4800.0 + 40.0 * k for k in range(8)]
    return cf.standard_config(16, 2, freqs, 32.0, 1024, subint_ns=B * 32000, xengine=xengine, seed=seed, **kw)


def case_long_ffts(B: int = 8, xengine: str = "fp32", nstations: int = 3, seed: int = 6, **kw):
    """The longest fused FFTs: 64 MHz / 4096 channels (L = 8192) beside 32 MHz / 2048
    channels (L = 4096), single pol, with a 1 MHz zoom in each."""
    groups = [(64.0, 4096, [6000.0]), (32.0, 2048, [6100.0])]
    zooms = [(0, 0, 5.0, 1.0), (1, 0, 3.0, 1.0)]
    return cf.batched_config(nstations, groups, zooms, subint_ns=B * 64000, npol=1, xengine=xengine,
                             seed=seed, **kw)


def case_fourstep_ffts(B: int = 4, xengine: str = "fp32", nstations: int = 3, seed: int = 7, **kw):
    """FFT lengths beyond one workgroup: 64 MHz / 16384 channels (L = 32768) beside
    32 MHz / 8192 channels (L = 16384), single pol, with a 2 MHz zoom (512 channels)."""
    groups = [(64.0, 16384, [6000.0]), (32.0, 8192, [6100.0])]
    zooms = [(0, 0, 10.0, 2.0), (1, 0, 3.0, 0.5)]
    return cf.batched_config(nstations, groups, zooms, subint_ns=B * 256000, npol=1, xengine=xengine,
                             seed=seed, **kw)


def case_complex_fourstep(B: int = 4, xengine: str = "fp32", nstations: int = 4, seed: int = 8, **kw):
    """Complex 8-bit 16 MHz with nchan 16384 (a 16384-point complex FFT)."""
    # delays within +-200 us: a 4-FFT subint is 4.1 ms, shorter than the default +-10 ms delays,
    # and the slot buffer may hold only getDataBytes plus four frames
    return cf.standard_config(nstations, 1, [1400.0], 16.0, 16384, subint_ns=B * 1024000,
                              nbits=8, complex_sampling=1, xengine=xengine, seed=seed,
                              max_delay_us=200.0, **kw)


def case_offsets(B: int = 32, xengine: str = "fp32", seed: int = 9, **kw):
    """4 stations dual pol, 2 x 32 MHz with a fractional-MHz band edge, per-freq LO offsets
    (integer and fractional Hz), clock offsets, pol clock deltas and pol phase offsets,
    channel averaging 4.  Station 1 records L before R, so the per-pol terms follow the pol
    and not the band order."""
    cfg = cf.standard_config(4, 2, [4800.37, 4840.0], 32.0, 512, subint_ns=B * 16000, xengine=xengine,
                             seed=seed, **kw)
    rng = np.random.default_rng(seed)
    for d in cfg.datastreams:
        d.lo_offset_hz = [float(rng.choice([0.0, 10.0, -2.5, 37.25])) for _ in d.recfreqs]
        d.clock_offset_us = [float(rng.uniform(-0.02, 0.02)) for _ in d.recfreqs]
        d.clock_offset_delta_us = [float(rng.uniform(-0.005, 0.005)) for _ in d.recfreqs]
        d.phase_offset_deg = [float(rng.uniform(-90.0, 90.0)) for _ in d.recfreqs]
    d1 = cfg.datastreams[1]
    d1.recbands = [(lf, "L" if p == "R" else "R") for lf, p in d1.recbands]
    for f in cfg.freqs:
        f.chanavg = 4
    return cfg


def case_offsets_fourstep(B: int = 4, xengine: str = "fp32", seed: int = 19, **kw):
    """The per-freq terms of case_offsets on the four-step F-engine: 3 stations dual pol,
    32 MHz with nchan 8192 (L = 16384), a fractional-MHz edge."""
    cfg = cf.standard_config(3, 2, [6000.37], 32.0, 8192, subint_ns=B * 256000, xengine=xengine,
                             seed=seed, max_delay_us=200.0, **kw)
    rng = np.random.default_rng(seed)
    for d in cfg.datastreams:
        d.lo_offset_hz = [37.25]
        d.clock_offset_us = [float(rng.uniform(-0.02, 0.02))]
        d.clock_offset_delta_us = [float(rng.uniform(-0.005, 0.005))]
        d.phase_offset_deg = [float(rng.uniform(-90.0, 90.0))]
    d1 = cfg.datastreams[1]
    d1.recbands = [(lf, "L" if p == "R" else "R") for lf, p in d1.recbands]
    return cfg


def case_dsb_fraclo(B: int = 16, xengine: str = "fp32", seed: int = 10, **kw):
    """Complex 8-bit double-sideband data, 4 stations, a 5 MHz band with a whole-MHz edge and a
    half-MHz centre (fractional LO, mode.cpp:83-101), USB and LSB freqs, nchan 256."""
    cfg = cf.standard_config(4, 1, [4800.0, 4900.0], 5.0, 256, subint_ns=B * 51200, nbits=8,
                             complex_sampling=2, xengine=xengine, seed=seed, max_delay_us=300.0, **kw)
    cfg.freqs[1].lsb = True
    return cfg


def case_usb_lsb(B: int = 32, xengine: str = "fp32", seed: int = 21, **kw):
    """USB x LSB: 4 stations dual pol; stations 0 and 1 record 4800-4832 MHz as USB (freq 0,
    edge 4800), stations 2 and 3 record the same sky band as LSB (freq 1, edge 4832),
    nchan 256.  Baselines within a sideband correlate their own freq; baselines across
    correlate freq 0 (the first datastream's) with the LSB bands on the second datastream
    (configuration.cpp:3125-3135)."""
    cfg = cf.standard_config(4, 2, [4800.0], 32.0, 256, subint_ns=B * 8000, xengine=xengine, seed=seed, **kw)
    cfg.freqs.append(cf.Freq(4832.0, 32.0, 256, lsb=True))
    for d in (2, 3):
        cfg.datastreams[d].recfreqs = [1]
    for b in cfg.baselines:
        if b.ds1 >= 2:
            b.freqs = [(1, prods) for _, prods in b.freqs]
    return cfg


def case_mark5b(B: int = 32, xengine: str = "fp32", **kw):
    """cfg1 shape (4 stations, 4 freqs dual pol) recorded as Mark5B frames."""
    cfg = case_cfg1(B=B, xengine=xengine, nstations=4, nfreq=4, **kw)
    for d in cfg.datastreams:
        d.format, d.headerbytes, d.payloadbytes = "MARK5B", 16, 10000
    return cfg


def case_real8(B: int = 32, xengine: str = "fp32", **kw):
    """8-bit real VDIF, 4 stations, 2 freqs dual pol, 16 MHz, nchan 256."""
    return cf.standard_config(4, 2, [1400.0, 1420.0], 16.0, 256, subint_ns=B * 16000, nbits=8,
                              xengine=xengine, seed=5, **kw)


def with_phase_centres(cfg, P: int = 3, seed: int = 40, interval_ffts: int = 8, threads: int = 3,
                       nbuf: int = 4):
    """P phase centres around the pointing centre: each station's phase-centre delays are its
    pointing delays plus a seeded offset of up to +-50 ns and a rate difference of up to
    +-2e-5 us/s (datastreams with identical pointing models belong to one station); uv-shift
    intervals of interval_ffts FFTs, rounded down to a multiple of nbuf buffered FFTs, over
    `threads` Core process threads."""
    rng = np.random.default_rng(seed)
    stations = []
    for ds in cfg.datastreams:
        k = next((i for i, m in enumerate(stations) if np.array_equal(m, ds.model)), None)
        if k is None:
            stations.append(ds.model)
    offs = {i: [(rng.uniform(-0.05, 0.05), rng.uniform(-2e-5, 2e-5)) for _ in range(P)]
            for i in range(len(stations))}
    for ds in cfg.datastreams:
        k = next(i for i, m in enumerate(stations) if np.array_equal(m, ds.model))
        ms = []
        for d0, r0 in offs[k]:
            m = ds.model.copy()
            m[:, 0] += d0
            m[:, 1] += r0
            ms.append(m)
        ds.phasecentre_models = ms
    cfg.nphasecentres = P
    cfg.process_threads = threads
    cfg.num_buffered_ffts = nbuf
    B = cfg.blockspersend()
    cfg.max_ns_between_xc_avg = int(interval_ffts * cfg.subint_ns / B) + 1
    return cfg


def slot(cfg, offsets=(0, 1, 0), seed: int = 11):
    return synth.noise_slot(cfg, list(offsets), seed=seed)


def damaged_slot(cfg, offsets=(0, 1, 0), seed: int = 12):
    """A noise slot with flagged FFTs (valid bits cleared) on station 0 and an invalid frame
    and a fill-pattern frame on station 1."""
    data, ctrl = synth.noise_slot(cfg, list(offsets), seed=seed)
    B = cfg.blockspersend()
    start = synth.buffer_start_ns(cfg, 0, list(offsets))
    ctrl[0] = synth.control_words(cfg, 0, list(offsets), start, invalid_ffts=[1, 2, B // 2, B - 1])
    ds = cfg.datastreams[1]
    buf = data[1].copy()
    fb, hb = ds.framebytes, ds.headerbytes
    nfr = len(buf) // fb
    # frames 2 and 5, or a third and two thirds into a shorter buffer
    f_inv, f_fill = (2, 5) if nfr >= 6 else (nfr // 3, (2 * nfr) // 3)
    if ds.format.upper() == "VDIF":
        buf[f_inv * fb + 3] |= 0x80                      # invalid bit
    else:
        buf[f_inv * fb + 5] |= 0x80
    fill = np.frombuffer(np.uint64(0x1122334411223344).tobytes(), np.uint8)
    buf[f_fill * fb + hb: f_fill * fb + hb + 8] = fill    # fill pattern opens the frame
    data[1] = buf
    return data, ctrl


FILL64 = np.frombuffer(np.uint64(0x1122334411223344).tobytes(), np.uint8)


def mark5_fill(payload: np.ndarray, words, half: int = -1):
    """Write the fill pattern over 64-bit words `words` of a payload (in place); `half` >= 0
    also fills the upper 32-bit half of that word."""
    for k in words:
        payload[8 * k: 8 * k + 8] = FILL64
    if half >= 0:
        payload[8 * half + 4: 8 * half + 8] = FILL64[:4]


def mark5b_partial_slot(cfg, offsets=(0, 1, 0), seed: int = 13):
    """A noise slot of Mark5B frames (10000-byte payloads, 1250 words) in which station 1's
    frames carry the fill pattern over parts of their payloads: the tail from word 625, the
    tail from word 313, all but word 0, the tail from 625 with the upper half of word 624
    (a straggler), the head to word 624, word 0 alone, both ends, and a middle run only."""
    data, ctrl = synth.noise_slot(cfg, list(offsets), seed=seed)
    ds = cfg.datastreams[1]
    buf = data[1].copy()
    fb, hb = ds.framebytes, ds.headerbytes
    nfr = len(buf) // fb
    plans = [(range(625, 1250), -1), (range(313, 1250), -1), (range(1, 1250), -1), (range(625, 1250), 624),
             (range(0, 625), -1), (range(0, 1), -1), (list(range(0, 3)) + list(range(1240, 1250)), -1),
             (range(100, 300), -1)]
    for j, (words, half) in enumerate(plans):
        f = 1 + (j * max(1, nfr - 2)) // len(plans)
        if f < nfr:
            mark5_fill(buf[f * fb + hb: (f + 1) * fb], words, half)
    data[1] = buf
    return data, ctrl


def record_errors(cfg, lay, got: np.ndarray, ref: np.ndarray):
    """Per-record errors in the diffDiFX convention (utilities/vis2screen/diffDiFX.py:137-151:
    mean|got-ref| / mean|ref| over the channels of one record) for every baseline product
    and every datastream's autocorrelation block."""
    out = {"cross": [], "auto": []}
    nb = len(cfg.baselines)
    P = getattr(lay, "nphasecentres", 1)
    for (fi, j), off in lay.baseline_off.items():
        fr = cfg.freqs[fi]
        prods = dict(cfg.baselines[j].freqs)[fi]
        nout = (fr.nchan + fr.chanavg - 1) // fr.chanavg
        for p in range(len(prods) * P):             # phase-centre blocks follow each other
            g = got[off + p * nout: off + (p + 1) * nout]
            r = ref[off + p * nout: off + (p + 1) * nout]
            den = np.mean(np.abs(r))
            out["cross"].append(float(np.mean(np.abs(g - r)) / den) if den > 0 else float(np.max(np.abs(g))))
    for d, ds in enumerate(cfg.datastreams):
        o = lay.autocorr_off[d]
        end = lay.autocorr_off[d + 1] if d + 1 < len(cfg.datastreams) else lay.acweight_off[0]
        n = end - o
        g, r = got[o: o + n], ref[o: o + n]
        den = np.mean(np.abs(r))
        if den > 0:
            out["auto"].append(float(np.mean(np.abs(g - r)) / den))
    return out


def channel_errors(cfg, lay, got: np.ndarray, ref: np.ndarray):
    """Per-channel errors of every cross record of a configuration whose datastreams hold one
    recorded band each (cfg2): |V_got - V_ref| / sqrt(A_1 A_2) with A the reference
    autocorrelations of the two inputs, and the phase difference arg(V_got conj(V_ref)) in
    degrees over the channels whose |V_ref| / sqrt(A_1 A_2) is at least the median (where the
    phase is defined; near |V| = 0 any error turns the phase freely).  Returns the two
    flattened arrays."""
    norm, dphi = [], []
    for (fi, j), off in lay.baseline_off.items():
        bl = cfg.baselines[j]
        fr = cfg.freqs[fi]
        assert fr.chanavg == 1
        n = fr.nchan
        a1 = ref[lay.autocorr_off[bl.ds1]: lay.autocorr_off[bl.ds1] + n].real
        a2 = ref[lay.autocorr_off[bl.ds2]: lay.autocorr_off[bl.ds2] + n].real
        assert len(cfg.datastreams[bl.ds1].recbands) == 1 and len(cfg.datastreams[bl.ds2].recbands) == 1
        g, r = got[off: off + n], ref[off: off + n]
        norm.append(np.abs(g - r) / np.sqrt(a1 * a2))
        coh = np.abs(r) / np.sqrt(a1 * a2)
        keep = coh >= np.median(coh)
        dphi.append(np.degrees(np.abs(np.angle(g[keep] * np.conj(r[keep])))))
    return np.concatenate(norm), np.concatenate(dphi)


def pcal_tone_slot(cfg, tones, offsets=(0, 1, 0), noise: float = 0.0, seed: int = 13):
    """8-bit real slot whose every band carries sum_j A_j cos(2 pi f_j t + phi_j) (+ Gaussian
    noise of rms `noise`), t the sample's time since the scan start, levels round(128 + 3.3 x)
    (the 8-bit decode is (level - 128) / 3.3).  tones: [(f_hz, A, phi)].  Returns (data,
    control, decoded) with decoded[d] the function m -> decoded value of datastream d's sample
    number m since the scan start (without the noise)."""
    rng = np.random.default_rng(seed)
    data, ctrl, funcs = [], [], []
    for d, ds in enumerate(cfg.datastreams):
        assert ds.nbits == 8 and not ds.complex_sampling
        start_ns = synth.buffer_start_ns(cfg, d, list(offsets))
        nbytes = synth.slot_bytes(cfg, d, list(offsets), start_ns)
        nframes = nbytes // ds.framebytes
        nb = len(ds.recbands)
        spf = (ds.framebytes - ds.headerbytes) // nb
        dt_ns = cfg.sampletime_us(d) * 1e3
        s_start = start_ns / dt_ns
        assert s_start == int(s_start), "buffer start on a sample boundary"
        fs = 1e9 / dt_ns

        def levels(m, fs=fs):
            m = np.asarray(m, dtype=np.int64)
            x = np.zeros(m.shape)
            for f, a, ph in tones:
                x += a * np.cos(2 * np.pi * ((int(f) * (m % int(fs))) % int(fs)) / fs + ph)
            return x

        m = int(s_start) + np.arange(nframes * spf, dtype=np.int64)
        x = levels(m)
        if noise:
            x = x + noise * rng.standard_normal(x.shape)
        lev = np.clip(np.round(128.0 + 3.3 * x), 0, 255).astype(np.uint8)
        payload = np.repeat(lev, nb)                          # the same signal in every band
        frame_ns = int(round(spf * dt_ns))
        data.append(synth.frame_payloads(payload, ds, int(start_ns // frame_ns)))
        ctrl.append(synth.control_words(cfg, d, list(offsets), start_ns))
        funcs.append(lambda mm, lv=levels: (np.clip(np.round(128.0 + 3.3 * lv(mm)), 0, 255) - 128.0) / 3.3)
    return data, ctrl, funcs


def pcal_complex_tone_slot(cfg, tones, offsets=(0, 1, 0)):
    """Complex 8-bit slot whose every band carries sum_j A_j exp(i (2 pi f_j m / fs + phi_j)),
    m the sample number since the scan start and fs = bw, levels round(128 + 3.3 x) for the
    real and imaginary parts (the 8-bit decode is (level - 128) / 3.3).  tones: [(f_hz, A, phi)]
    with f_hz whole and possibly negative.  Returns (data, control, decoded) with decoded[d]
    the function m -> decoded complex value of datastream d's sample number m."""
    data, ctrl, funcs = [], [], []
    for d, ds in enumerate(cfg.datastreams):
        assert ds.nbits == 8 and ds.complex_sampling
        start_ns = synth.buffer_start_ns(cfg, d, list(offsets))
        nbytes = synth.slot_bytes(cfg, d, list(offsets), start_ns)
        nframes = nbytes // ds.framebytes
        nb = len(ds.recbands)
        spf = (ds.framebytes - ds.headerbytes) // (2 * nb)
        dt_ns = cfg.sampletime_us(d) * 1e3
        s_start = start_ns / dt_ns
        assert s_start == int(s_start), "buffer start on a sample boundary"
        fs = int(round(1e9 / dt_ns))

        def signal(m, fs=fs):
            m = np.asarray(m, dtype=np.int64)
            x = np.zeros(m.shape, complex)
            for f, a, ph in tones:
                x += a * np.exp(1j * (2 * np.pi * ((int(f) * (m % fs)) % fs) / fs + ph))
            return x

        def lev(x):
            return np.clip(np.round(128.0 + 3.3 * x), 0, 255)

        m = int(s_start) + np.arange(nframes * spf, dtype=np.int64)
        x = signal(m)
        pair = np.stack([lev(x.real), lev(x.imag)], axis=1).astype(np.uint8)      # [samples, (re, im)]
        payload = np.repeat(pair[:, None, :], nb, axis=1).reshape(-1)           # every band the same
        frame_ns = int(round(spf * dt_ns))
        data.append(synth.frame_payloads(payload, ds, int(start_ns // frame_ns)))
        ctrl.append(synth.control_words(cfg, d, list(offsets), start_ns))
        funcs.append(lambda mm, sg=signal: ((lev(sg(mm).real) - 128.0) / 3.3) + 1j * ((lev(sg(mm).imag) - 128.0) / 3.3))
    return data, ctrl, funcs
