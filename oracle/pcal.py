"""CPU restatement of phase-cal extraction (test infrastructure only).

This module is a CHECKER, not the product: only tests/ may import it.

Tone table (Configuration, mpifxcorr/src/configuration.cpp:1765-1835), per recorded freq with
band edge lofreq (MHz), bandwidth bw and a whole-Hz comb interval:
  harmonic = (long)((lofreq - base 1e-6) / step), step = interval 1e-6 MHz
  USB: the first tone strictly above lofreq, counted while tone + n step <= lofreq + bw;
       offset = harmonic interval - long(1e6 lofreq)
  LSB: the first tone strictly below lofreq, counted while tone - n step >= lofreq - bw;
       offset = long(1e6 lofreq) - harmonic interval
  no tone in the band: one placeholder tone and offset -1.
  The offset leaves base_hz out, as the reference does, so a base can make it negative; a
  negative offset selects the dummy extractor (pcal.cpp:202-230, 1761-1775), whose tones stay 0.

Extraction (Mode::process, mode.cpp:688-772, and the PCal extractors of pcal.cpp): every FFT
that passes the range check and has a nonzero data weight hands its fftchannels unpacked samples
(real, or complex for complex-sampled data), starting at nearestsample, to the extractor,
numbered from the scan start by
m = nearestsample + (size_t)((datasec 1e9 + datans) / (sampletime 1e3) + 0.5) + n.  Every
extractor returns, per tone, tone_k = sum_m x[m] exp(-2 pi i (num_k / den_k) m) for a rational
frequency num_k / den_k in cycles per sample (extractor_tones):
  real data (fs = 2 bw; trivial, implicit-shift and shifting extractors, pcal.cpp:191-274):
    num/den = f_k / fs with f_k = offset + k interval, or (k + 1) interval when the offset is 0
    (the trivial extractor drops the DC bin, pcal.cpp:573-580);
  complex data (fs = bw): the implicit-shift complex extractor when No % Np == 0
    (No = fs / gcd(fs, offset), Np = fs / gcd(fs, interval)), else the shifting complex one;
    their DFT length N starts at No (implicit shift) or Np (shifting) and doubles while
    fs / (2 N) > 1e6 (PCal::setMinFrequencyResolution(1e6), mode.cpp:402; pcal.cpp:1419-1440,
    808-830).  Tone k is DFT bin b_k = floor(N offset / fs) + k floor(N interval / fs) of the
    folded samples (implicit shift, pcal.cpp:1571-1590), or bin k floor(N interval / fs) after
    the samples are mixed by exp(-2 pi i offset m / fs) (shifting, pcal.cpp:927-1020); a bin
    >= N leaves the tone at zero.  For double-sideband data the DFT output is reversed
    (bin b -> N - 1 - b) before the tones are read (pcal.cpp:1003-1008, 1560-1566), so the
    tone sits at -(b + 1) fs / N.
Here the sum is taken directly in FP64 with the phase reduced exactly in integers
((num_k (m mod den_k)) mod den_k).

A buffer that starts before the scan start (negative datasec) gives negative m; this
restatement (and the device) use m mod M in [0, M).  The reference converts the negative
double to size_t, which C leaves undefined; x86-64 wraps it to 2^64 + m, and that agrees with
m mod M whenever M divides 2^64 (M a power of two).

Pins (tests/test_pcal_cpu.py): tone tables worked by hand for USB and LSB bands, a band with no
tone and a zero offset; a planted cosine at a tone frequency returns its amplitude N A / 2 and
its phase; the FFTs that feed the sums are exactly those to which the FP64 F-engine restatement
gives a nonzero weight; folding by sample number modulo the tone period and an M-point DFT give
the same tones as the direct sum.
"""
from __future__ import annotations

import math
from typing import List, Optional, Tuple

import numpy as np

import fxoracle as fo


def tone_table(lofreq_mhz: float, bw_mhz: float, lsb: bool, interval_hz: int,
               base_hz: int = 0) -> Tuple[int, int, List[float]]:
    """(number of tones, baseband offset of the first tone in Hz, sky frequencies in Hz)."""
    step = interval_hz * 1e-6
    harmonic = int((lofreq_mhz - base_hz * 1e-6) / step)           # C truncation
    tonefreq = harmonic * step + base_hz * 1e-6
    n = 0
    if lsb:
        if tonefreq >= lofreq_mhz:
            tonefreq -= step
            harmonic -= 1
        if tonefreq >= lofreq_mhz - bw_mhz:
            while tonefreq - n * step >= lofreq_mhz - bw_mhz:
                n += 1
        if n == 0:
            return 1, -1, [0.0]
        sky = [float((harmonic - k) * interval_hz + base_hz) for k in range(n)]
        return n, int(1e6 * lofreq_mhz) - harmonic * interval_hz, sky
    if tonefreq <= lofreq_mhz:
        tonefreq += step
        harmonic += 1
    if tonefreq <= lofreq_mhz + bw_mhz:
        while tonefreq + n * step <= lofreq_mhz + bw_mhz:
            n += 1
    if n == 0:
        return 1, -1, [0.0]
    sky = [float((harmonic + k) * interval_hz + base_hz) for k in range(n)]
    return n, harmonic * interval_hz - int(1e6 * lofreq_mhz), sky


def baseband_freqs(offset_hz: int, interval_hz: int, ntones: int) -> List[int]:
    """Baseband frequency of each extracted tone (Hz)."""
    first = interval_hz if offset_hz == 0 else offset_hz
    return [first + k * interval_hz for k in range(ntones)]


def tone_period(fs_hz: int, offset_hz: int, interval_hz: int) -> int:
    """Samples after which every tone's phase repeats: fs / gcd(fs, offset, interval)."""
    return fs_hz // math.gcd(fs_hz, math.gcd(offset_hz, interval_hz))


MIN_FREQ_RESOLUTION_HZ = 1_000_000          # PCal::setMinFrequencyResolution(1e6), mode.cpp:402


def _dft_length(fs_hz: int, n0: int) -> int:
    """The extractors' N_bins: n0 doubled while fs / (2 N) > the minimum resolution (integer
    division, as the reference's long / int)."""
    n = n0
    while fs_hz // (2 * n) > MIN_FREQ_RESOLUTION_HZ:
        n *= 2
    return n


def extractor_tones(ntones: int, offset_hz: int, interval_hz: int, fs_hz: int, cplx: bool,
                    dsb: bool) -> List[Optional[Tuple[int, int]]]:
    """Per tone the reduced rational frequency (num, den) in cycles per sample that the
    reference's extractor reads, or None for a tone it leaves at zero (a flagged offset: the
    dummy extractor; a bin past the DFT length)."""
    if offset_hz < 0:
        return [None] * ntones
    out: List[Optional[Tuple[int, int]]] = []

    def red(num, den):
        num %= den
        g = math.gcd(num, den)
        return num // g, den // g

    if not cplx:
        first = interval_hz if offset_hz == 0 else offset_hz
        return [red(first + k * interval_hz, fs_hz) for k in range(ntones)]
    no = fs_hz // math.gcd(fs_hz, offset_hz)
    np_ = fs_hz // math.gcd(fs_hz, interval_hz)
    if no % np_ == 0:                          # PCalExtractorImplicitShiftComplex
        n = _dft_length(fs_hz, no)
        step = (n * interval_hz) // fs_hz
        obin = (n * offset_hz) // fs_hz
        for k in range(ntones):
            b = obin + k * step
            if b >= n:
                out.append(None)
                continue
            out.append(red(n - 1 - b if dsb else b, n))
        return out
    n = _dft_length(fs_hz, np_)                # PCalExtractorShiftingComplex
    step = (n * interval_hz) // fs_hz
    for k in range(ntones):
        b = k * step
        if b >= n:
            out.append(None)
            continue
        bb = n - 1 - b if dsb else b
        out.append(red(offset_hz * n + bb * fs_hz, fs_hz * n))
    return out


def fold_period(tones: List[Optional[Tuple[int, int]]]) -> int:
    """Samples after which every extracted tone's phase repeats: the lcm of the denominators
    (0 when no tone is read)."""
    m = 0
    for t in tones:
        if t is not None:
            m = t[1] if m == 0 else m * t[1] // math.gcd(m, t[1])
    return m


def valid_ffts(cfg, d: int, offsets, data: np.ndarray, control: np.ndarray):
    """(FFT index, nearestsample) of every FFT that feeds the extractor, and the sample number
    of the buffer start since the scan start, following Mode::setOffsets and Mode::process's
    exits (mode.cpp:1502-1532, 688-745)."""
    ds = cfg.datastreams[d]
    B = cfg.blockspersend()
    L = cfg.fftchannels(d)
    dt = cfg.sampletime_us(d)
    scan, datasec, datans = int(control[0]), int(control[1]), int(control[2])
    if datasec <= fo.INVALID_SUBINT or len(data) <= 1 or offsets[0] != scan:
        return [], 0
    f0 = cfg.freqs[ds.recfreqs[0]]
    timespan = B * f0.nchan / f0.bw_mhz / 1e6
    interp = fo.delay_interpolator(ds.model, ds.model_order, cfg.model_inc_secs,
                                   offsets[1] + offsets[2] / 1e9, timespan, B)
    if interp is None:
        return [], 0
    interp = interp.copy()
    interp[2] -= 1000000 * fo.intclockseconds(ds)
    _, valid = fo.decode_stream(data, ds)
    ns = fo.fft_params(cfg, d, offsets, control, interp, B)[0]
    nrec = len(ds.recbands)
    bpb_num, bpb_den = (nrec * ds.nbits) // 8, 1
    if bpb_num == 0:
        bpb_num, bpb_den = 1, 8 // (nrec * ds.nbits)
    out = []
    for i in range(B):
        if not ((int(control[3 + i // fo.FLAGS_PER_INT]) >> (i % fo.FLAGS_PER_INT)) & 1):
            continue
        n0 = int(ns[i])
        if n0 < -1 or ((n0 + L) * bpb_num) // bpb_den > len(data):
            continue
        n0 = max(n0, 0)
        if int(valid[n0:n0 + L].sum()) == 0:                 # zero data weight
            continue
        out.append((i, n0))
    s0 = int((datasec * 1e9 + datans) / (dt * 1e3) + 0.5)
    return out, s0


def band_tones(cfg, d: int, b: int, interval_hz: int, base_hz: int = 0):
    """(ntones, offset_hz, extractor tones) of recorded band b of datastream d."""
    ds = cfg.datastreams[d]
    f = cfg.freqs[ds.recfreqs[ds.recbands[b][0]]]
    cplx = bool(ds.complex_sampling)
    fs = int(round((1e6 if cplx else 2e6) * f.bw_mhz))
    ntones, offset, _ = tone_table(f.freq_mhz, f.bw_mhz, f.lsb, interval_hz, base_hz)
    return ntones, offset, extractor_tones(ntones, offset, interval_hz, fs, cplx, ds.complex_sampling == 2)


def extract(cfg, d: int, offsets, data: np.ndarray, control: np.ndarray, interval_hz: int,
            base_hz: int = 0) -> List[np.ndarray]:
    """Tones (complex128) of every recorded band of datastream d for one subint."""
    ds = cfg.datastreams[d]
    L = cfg.fftchannels(d)
    vals, _ = fo.decode_stream(data, ds)
    nsamp = vals.shape[1]
    ffts, s0 = valid_ffts(cfg, d, offsets, data, control)
    out = []
    for b in range(len(ds.recbands)):
        ntones, _, tones = band_tones(cfg, d, b, interval_hz, base_hz)
        acc = np.zeros(ntones, np.complex128)
        live = [k for k, t in enumerate(tones) if t is not None]
        if live:
            num = np.array([tones[k][0] for k in live], dtype=np.int64)
            den = np.array([tones[k][1] for k in live], dtype=np.int64)
            for _, n0 in ffts:
                seg = np.zeros(L, dtype=np.complex128)
                hi = min(n0 + L, nsamp)
                if hi > n0:
                    seg[: hi - n0] = vals[b, n0:hi]
                m = s0 + n0 + np.arange(L, dtype=np.int64)
                r = (num[:, None] * (m[None, :] % den[:, None])) % den[:, None]
                acc[live] += (np.exp(-2j * np.pi * r / den[:, None]) * seg[None, :]).sum(axis=1)
        out.append(acc)
    return out
