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
that passes the range check and has a nonzero data weight hands its fftchannels unpacked samples,
starting at nearestsample, to the extractor, numbered from the scan start by
m = nearestsample + (size_t)((datasec 1e9 + datans) / (sampletime 1e3) + 0.5) + n.  Every
extractor (trivial, implicit-shift, shifting) returns for tone k at baseband frequency f_k
  tone_k = sum_m x[m] exp(-2 pi i f_k m / fs),   fs = 2 bw (real data),
with f_k = offset + k interval, or (k + 1) interval when the offset is 0 (the trivial extractor
drops the DC bin, pcal.cpp:573-580).  Here the sum is taken directly in FP64 with the phase
reduced exactly in integers ((f_k (m mod fs)) mod fs).

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
from typing import List, Tuple

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


def extract(cfg, d: int, offsets, data: np.ndarray, control: np.ndarray, interval_hz: int,
            base_hz: int = 0) -> List[np.ndarray]:
    """Tones (complex128) of every recorded band of datastream d for one subint."""
    ds = cfg.datastreams[d]
    if ds.complex_sampling:
        raise NotImplementedError("phase-cal of complex-sampled data")
    L = cfg.fftchannels(d)
    vals, _ = fo.decode_stream(data, ds)
    nsamp = vals.shape[1]
    ffts, s0 = valid_ffts(cfg, d, offsets, data, control)
    out = []
    for b, (fi, _) in enumerate(ds.recbands):
        f = cfg.freqs[ds.recfreqs[fi]]
        fs = int(round(2e6 * f.bw_mhz))
        ntones, offset, _ = tone_table(f.freq_mhz, f.bw_mhz, f.lsb, interval_hz, base_hz)
        if offset < 0:                                         # flagged: the dummy extractor
            out.append(np.zeros(ntones, np.complex128))
            continue
        fk = np.array(baseband_freqs(offset, interval_hz, ntones), dtype=np.int64)
        acc = np.zeros(ntones, np.complex128)
        for _, n0 in ffts:
            seg = np.zeros(L)
            hi = min(n0 + L, nsamp)
            if hi > n0:
                seg[: hi - n0] = vals[b, n0:hi]
            m = (s0 + n0 + np.arange(L, dtype=np.int64)) % fs
            r = (fk[:, None] * m[None, :]) % fs
            acc += (np.exp(-2j * np.pi * r / fs) * seg[None, :]).sum(axis=1)
        out.append(acc)
    return out
