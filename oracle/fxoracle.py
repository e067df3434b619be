"""FP64 oracle of mpifxcorr's per-subintegration FX path (test infrastructure only).

This module is the CHECKER, not the product: only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py may import it.  It restates, in numpy with float64/complex128
arithmetic, what the reference computes for one subint and one Core slot:

  unpack          Mk5Mode::unpack (mk5mode.cpp:94-174) -> mark5_unpack_with_offset
                  (mark5_stream_unpacker.c:119-138), VDIF LUTs (format_vdif.c:66-145,
                  834-880, 1621-1680, 5215-5250), Mark5B LUT (mark5_format_mark5b.c:49-86,
                  1708-1760), frame validation / blanking (format_vdif.c:7619-7660,
                  blanker_mark5.c:43-172, 272-310, mark5_stream.c:144-193)
  delay model     Model::calculateDelayInterpolator (model.cpp:219-289), Mode::setOffsets
                  (mode.cpp:1502-1532)
  F-engine        Mode::process (mode.cpp:634-1374), fringe rotation order 1, fractional
                  sample correction, sideband channel selection, autocorrelations
  X-engine        Core::processdata XMAC (core.cpp:884-979), baseline weights
                  (core.cpp:1005-1052)
  accumulation    Core::uvshiftAndAverageBaselineFreq (core.cpp:1819-1853),
                  Core::averageAndSendAutocorrs (core.cpp:1258-1376), Mode::averageFrequency
                  (mode.cpp:1376-1401)
  layout          Configuration::populateResultLengths (configuration.cpp:2348-2630)
  normalisation   Visibility::writedata scale (visibility.cpp:526-527)

Where the reference rounds an intermediate to float32 (sin/cos arguments, the fractional
sample error, channel-frequency slopes) the oracle keeps float64: it is the exact
mathematics of the reference algorithm, against which both the FP32 and the FP16/BF16
device paths are measured.  The FFT is numpy's complex128 FFT, itself pinned by brute-force
DFT checks in tests/test_oracle.py.

Pins (tests/test_oracle.py): LUT closed forms over all 256 byte values; frame skipping and
blanking on hand-built frames; interpolator exactness for polynomials of degree <= 2;
per-FFT delay chain special cases (zero and integer-sample delays); rotator identity and
unit-phase cases; DFT against an O(L^2) brute force; XMAC/auto invariants (Hermitian,
autos real and non-negative, a station against itself equals its autocorrelation); the
result layout against hand-counted offsets; end to end against the statistical closed form
rho*0.88^2 of a 2-bit common signal (mode.cpp:1571).
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

TWO_PI = 6.283185307179586476925286766559   # architecture.h.in:430
HIMAG = 3.3359                                # OPTIMAL_2BIT_HIGH, mark5_stream.h:65
FLAGS_PER_INT = 30                            # mpifxcorr.h:29
INVALID_SUBINT = -99999999                    # mode.h:222
TINY = 1e-9                                   # mode.cpp:27
FILL_WORD64 = 0x1122334411223344              # blanker_mark5.c:36 (little endian)
DECORR = [0.63662, 0.88, 0.94, 0.96, 0.98, 0.99, 0.996, 0.998]  # mode.cpp:1571


# --------------------------------------------------------------------------------------
# Look-up tables
# --------------------------------------------------------------------------------------
def vdif_lut2bit() -> np.ndarray:
    """format_vdif.c:66-110: lut2bit[b][i] = {-HiMag,-1,+1,+HiMag}[(b >> 2i) & 3]."""
    lev = np.array([-HIMAG, -1.0, 1.0, HIMAG], dtype=np.float32)
    b = np.arange(256)
    return np.stack([lev[(b >> (2 * i)) & 3] for i in range(4)], axis=1)


def mark5b_lut2bit() -> np.ndarray:
    """mark5_format_mark5b.c:49-86: sign bit at 2i, magnitude bit at 2i+1,
    lut4level = {-HiMag, 1, -1, HiMag}[sign + 2*mag]."""
    lev = np.array([-HIMAG, 1.0, -1.0, HIMAG], dtype=np.float32)
    b = np.arange(256)
    return np.stack([lev[((b >> (2 * i)) & 1) + (((b >> (2 * i + 1)) & 1) << 1)]
                     for i in range(4)], axis=1)


def lut8bit(offset: float = 128.0) -> np.ndarray:
    """format_vdif.c:120: (b - 128)/3.3 (127.5 with SYMMETRIC_8_BIT_VDIF), rounded to float."""
    return ((np.arange(256) - offset) / 3.3).astype(np.float32)


# --------------------------------------------------------------------------------------
# Frames and unpacking
# --------------------------------------------------------------------------------------
FILL_WORD32 = 0x11223344                      # blanker_mark5.c:35 (little endian)
ZONE_ALL = 1 << 30                            # blanker_mark5.c:149, 153: "no bound"
ZONE_WORDS = 1 << (15 - 3)                    # blanker_mark5.c:121, 135: 32768-byte zones


def _first_invalid(w64: np.ndarray, w32: np.ndarray, start: int, end: int) -> int:
    """findfirstinvalid (blanker_mark5.c:43-76): a halving search from the middle of words
    [start, end) for the fill pattern, keeping the lowest probe that held it (end-1 if none),
    as a byte offset; 4 bytes less when the 32-bit word just before it is fill (a straggler)."""
    n = end - start
    i, r, s = n // 2, n - 1, n // 4
    while s > 0:
        if int(w64[start + i]) == FILL_WORD64:
            r = i
            i -= s
        else:
            i += s
        s >>= 1
    r = 8 * (start + r)
    if int(w32[r // 4 - 1]) == FILL_WORD32:
        r -= 4
    return r


def _first_valid(w64: np.ndarray, w32: np.ndarray, start: int, end: int) -> int:
    """findfirstvalid (blanker_mark5.c:78-111): the same search for a word that is not fill;
    4 bytes more when the 32-bit word two before the result is fill (the reference tests
    w32[r/4 - 2], the low half of the previous 64-bit word, and the restatement keeps that)."""
    n = end - start
    i, r, s = n // 2, n - 1, n // 4
    while s > 0:
        if int(w64[start + i]) != FILL_WORD64:
            r = i
            i -= s
        else:
            i += s
        s >>= 1
    r = 8 * (start + r)
    if int(w32[r // 4 - 2]) == FILL_WORD32:
        r += 4
    return r


def mark5_zone(payload: np.ndarray) -> Tuple[int, int]:
    """blanker_mark5 (blanker_mark5.c:112-172) on one payload: the byte range
    [startvalid, endvalid) of zone 0, the only zone the Mark5B decoders read
    (mark5_format_mark5b.c:241-243: sample bytes i with startvalid <= i < endvalid decode).
    Both end words valid: the whole payload, whatever lies between; both fill: nothing; one
    end fill: the valid side up to the boundary the search finds."""
    nword = len(payload) // 8
    if nword == 0:
        return 0, 0
    w64 = payload[: nword * 8].copy().view("<u8")
    w32 = payload[: nword * 8].copy().view("<u4")
    e = min(ZONE_WORDS, nword)
    start_ok = int(w64[0]) != FILL_WORD64
    end_ok = int(w64[e - 1]) != FILL_WORD64
    if start_ok and end_ok:
        return 0, ZONE_ALL
    if not start_ok and not end_ok:
        return ZONE_ALL, 0
    if start_ok:
        return 0, _first_invalid(w64, w32, 0, e)
    return _first_valid(w64, w32, 0, e), ZONE_ALL


def frame_valid(buf: np.ndarray, fmt: str, framebytes: int, headerbytes: int) -> np.ndarray:
    """Per-frame validity of a slot buffer (whole-frame part).  VDIF: header word 2 == 0 or
    invalid bit (word 0 bit 31) fails (format_vdif.c:7619-7660); blanker_vdif blanks a payload
    whose first or last 64-bit word is the fill pattern (blanker_mark5.c:272-310).  Mark5B:
    the invalid flag is header byte 5 bit 7 (mark5_format_mark5b.c:1723, buf[-11] & 0x80);
    its fill pattern blanks zones of the payload (mark5_zone, applied per sample by
    sample_valid).  Frames extending past the buffer are invalid."""
    nframes = len(buf) // framebytes
    out = np.zeros(nframes, dtype=bool)
    if nframes == 0:
        return out
    fr = buf[: nframes * framebytes].reshape(nframes, framebytes)
    if fmt == "VDIF":
        payload = fr[:, headerbytes:]
        first = payload[:, :8].copy().view("<u8")[:, 0]
        last = payload[:, -8:].copy().view("<u8")[:, 0]
        fill = (first == FILL_WORD64) | (last == FILL_WORD64)
        w0 = fr[:, 0:4].copy().view("<u4")[:, 0]
        w2 = fr[:, 8:12].copy().view("<u4")[:, 0]
        return (w2 != 0) & (((w0 >> 31) & 1) == 0) & ~fill
    return (fr[:, 5] & 0x80) == 0


def sample_valid(buf: np.ndarray, ds) -> np.ndarray:
    """Per-sample validity of a slot buffer: the frame's validity and, for Mark5B, the zone
    test on the byte where the time sample's fields begin, i = (n * bits per time sample) / 8
    within the payload (mark5_format_mark5b.c:1295-1343, 1868-1932: one test per sample, at
    the readposition before its bytes are read)."""
    fmt = ds.format.upper()
    fb, hb = ds.framebytes, ds.headerbytes
    nframes = len(buf) // fb
    spf = samples_per_frame(ds)
    valid_f = frame_valid(buf, fmt, fb, hb)
    valid = np.repeat(valid_f, spf)
    if fmt == "MARK5B" and nframes > 0:
        bps = ds.nbits * (2 if ds.complex_sampling else 1) * len(ds.recbands)
        byte = (np.arange(spf, dtype=np.int64) * bps) // 8
        fr = buf[: nframes * fb].reshape(nframes, fb)
        for f in range(nframes):
            sv, ev = mark5_zone(fr[f, hb:])
            if (sv, ev) != (0, ZONE_ALL):
                valid[f * spf: (f + 1) * spf] &= (byte >= sv) & (byte < ev)
    return valid


def decode_stream(buf: np.ndarray, ds) -> Tuple[np.ndarray, np.ndarray]:
    """Decode every sample of a slot buffer: returns (values[nbands, nsamples] float32 or
    complex64, valid[nsamples] bool).  Sample n of band b sits at sample-field q = n*nbands+b
    of its frame's payload; 2-bit fields are LSB-first in bytes (format_vdif.c:834-880,
    1621-1680), 8-bit complex samples are (re, im) byte pairs (format_vdif.c:5215-5250)."""
    fmt = ds.format.upper()
    nb = len(ds.recbands)
    fb, hb = ds.framebytes, ds.headerbytes
    pay = fb - hb
    nframes = len(buf) // fb
    fr = buf[: nframes * fb].reshape(nframes, fb)[:, hb:]
    if ds.nbits == 2 and not ds.complex_sampling:
        lut = vdif_lut2bit() if fmt == "VDIF" else mark5b_lut2bit()
        vals = lut[fr.reshape(-1)].reshape(-1)                 # 4 fields per byte, LSB first
        spf = pay * 4 // nb                                    # samples per frame
        vals = vals.reshape(nframes * spf, nb).T
    elif ds.nbits == 8 and ds.complex_sampling:
        lut = lut8bit()
        v = lut[fr.reshape(-1)].reshape(-1, 2)
        c = (v[:, 0].astype(np.float32) + 1j * v[:, 1].astype(np.float32)).astype(np.complex64)
        spf = pay // (2 * nb)
        vals = c.reshape(nframes * spf, nb).T
    elif ds.nbits == 8 and not ds.complex_sampling:
        lut = lut8bit()
        spf = pay // nb
        vals = lut[fr.reshape(-1)].reshape(nframes * spf, nb).T
    else:
        raise NotImplementedError(f"{ds.nbits}-bit complex={ds.complex_sampling}")
    valid = sample_valid(buf, ds)
    vals = np.where(valid[None, :], vals, np.zeros((), vals.dtype))   # blanked samples are +0.0
    return vals, valid


def samples_per_frame(ds) -> int:
    pay = ds.framebytes - ds.headerbytes
    bits = ds.nbits * (2 if ds.complex_sampling else 1) * len(ds.recbands)
    return pay * 8 // bits


# --------------------------------------------------------------------------------------
# Delay model
# --------------------------------------------------------------------------------------
def _delay_samples(model: np.ndarray, order_poly: int, model_inc_secs: float, offsettime: float,
                   timespan: float):
    """The three polynomial evaluations of Model::calculateDelayInterpolator (model.cpp:
    228-273): at the start, middle and end of the span, or None outside the model."""
    scansample = int(offsettime / float(model_inc_secs))
    nsamp = model.shape[0]
    if scansample == nsamp and (offsettime - scansample * model_inc_secs) < 1e-6:
        scansample -= 1
    if scansample < 0 or scansample >= nsamp:
        return None
    c = model[scansample, : order_poly + 1]

    def ev(dt):
        tp = dt ** np.arange(order_poly + 1)
        return float(np.dot(tp, c))

    base = scansample * model_inc_secs
    d1 = ev(offsettime + timespan / 2.0 - base)
    d0 = ev(offsettime - base)
    d2 = ev(offsettime + timespan - base)
    return d0, d1, d2


def delay_interpolator(model: np.ndarray, order_poly: int, model_inc_secs: float,
                       offsettime: float, timespan: float, numincrements: int) -> Optional[np.ndarray]:
    """Model::calculateDelayInterpolator(order 2) (model.cpp:219-289) for one scan whose
    polynomials start at the scan start (polyoffset = 0): three evaluations of the station
    delay polynomial at the start, middle and end of the span, fitted with a quadratic in
    increment units."""
    ds = _delay_samples(model, order_poly, model_inc_secs, offsettime, timespan)
    if ds is None:
        return None
    d0, d1, d2 = ds
    n = numincrements
    return np.array([(2.0 * d0 - 4.0 * d1 + 2.0 * d2) / (n * n),
                     (-3.0 * d0 + 4.0 * d1 - d2) / n,
                     d0])


def delay_interpolator_linear(model: np.ndarray, order_poly: int, model_inc_secs: float,
                              offsettime: float, timespan: float, numincrements: int):
    """Model::calculateDelayInterpolator(order 1) (model.cpp:279-283): [slope per increment,
    offset] of the straight line through the three evaluations."""
    ds = _delay_samples(model, order_poly, model_inc_secs, offsettime, timespan)
    if ds is None:
        return None
    d0, d1, d2 = ds
    c0 = (d2 - d0) / numincrements
    return np.array([c0, d0 + (d1 - (c0 * numincrements / 2.0 + d0)) / 3.0])


def intclockseconds(ds) -> int:
    """mode.cpp:69."""
    return int(math.floor(ds.clock_coeff0_us / 1000000.0 + 0.5))


# --------------------------------------------------------------------------------------
# Result layout (Configuration::populateResultLengths, single phase centre, no binning)
# --------------------------------------------------------------------------------------
@dataclass
class Layout:
    coreresultlength: int
    baseline_off: Dict[Tuple[int, int], int]   # (freq, baseline) -> cf32 offset
    bweight_off: Dict[Tuple[int, int], int]
    autocorr_off: List[int]
    acweight_off: List[int]
    bandsperautocorr: int
    shiftdecorr_off: Dict[Tuple[int, int], int] = None   # (freq, baseline) -> cf32 offset (P > 1)
    nphasecentres: int = 1
    pcal_off: List[int] = None               # per datastream: cf32 offset of its phase-cal tones


def band_freq_index(ds, k: int) -> int:
    """Freq table index of band k of a datastream (recorded bands, then zoom bands)."""
    nrec = len(ds.recbands)
    return ds.recfreqs[ds.recbands[k][0]] if k < nrec else ds.zoomfreqs[ds.zoombands[k - nrec][0]][0]


def opposite_sideband_freq(cfg, fi: int) -> int:
    """Configuration::getOppositeSidebandFreqIndex (configuration.cpp:774-791): the last freq
    of the other sideband with the same bandwidth, channels and averaging whose band covers
    the same sky range (exact comparisons, as there), or -1."""
    f1 = cfg.freqs[fi]
    r = -1
    for i, f in enumerate(cfg.freqs):
        if f.bw_mhz == f1.bw_mhz and f.lsb != f1.lsb and f.nchan == f1.nchan and f.chanavg == f1.chanavg:
            if f.lsb and f1.freq_mhz + f1.bw_mhz == f.freq_mhz:
                r = i
            if not f.lsb and f.freq_mhz + f.bw_mhz == f1.freq_mhz:
                r = i
    return r


def autocorr_freq_labels(cfg) -> List[int]:
    """The freq index each freq's autocorrelation records carry (visibility.cpp:912-925): with
    a USB x LSB product on any baseline (anyusbxlsb, configuration.cpp:3192-3203), an LSB freq
    correlated against a USB freq (correlatedagainstupper, set from the first product of each
    baseline freq, configuration.cpp:1131-1158) is written as its opposite sideband freq."""
    labels = list(range(len(cfg.freqs)))
    against_upper = [False] * len(cfg.freqs)
    anymixed = False
    for b in cfg.baselines:
        for _, prods in b.freqs:
            b1, b2 = prods[0]
            f1 = band_freq_index(cfg.datastreams[b.ds1], b1)
            f2 = band_freq_index(cfg.datastreams[b.ds2], b2)
            if cfg.freqs[f1].lsb != cfg.freqs[f2].lsb:
                anymixed = True
                against_upper[f1 if cfg.freqs[f1].lsb else f2] = True
    if anymixed:
        for i, f in enumerate(cfg.freqs):
            if f.lsb and against_upper[i]:
                o = opposite_sideband_freq(cfg, i)
                if o >= 0:
                    labels[i] = o
    return labels


def freq_used(cfg) -> Tuple[List[bool], List[bool]]:
    """configuration.cpp:2270-2320: a freq is used if some baseline correlates it; an
    'equivalent' freq (same edge, bandwidth, channels, averaging) of a used freq is also
    kept for its autocorrelations."""
    nf = len(cfg.freqs)
    used = [False] * nf
    for b in cfg.baselines:
        for fi, _ in b.freqs:
            used[fi] = True
    equiv = [False] * nf
    for j in range(nf):
        if not used[j]:
            for k in range(nf):
                fj, fk = cfg.freqs[j], cfg.freqs[k]
                bwdiff = fj.bw_mhz - fk.bw_mhz
                fdiff = fj.freq_mhz - fk.freq_mhz
                if fj.lsb:
                    fdiff -= fj.bw_mhz
                if fk.lsb:
                    fdiff += fk.bw_mhz
                if (abs(bwdiff) < TINY and abs(fdiff) < TINY and fj.nchan == fk.nchan
                        and fj.chanavg == fk.chanavg and used[k]):
                    equiv[j] = True
    return used, equiv


def nphasecentres(cfg) -> int:
    return max(1, int(getattr(cfg, "nphasecentres", 1) or 1))


def result_layout(cfg) -> Layout:
    """populateResultLengths (configuration.cpp:2348-2630) for one config, no pulsar bins:
    per used freq and baseline a cross block of P phase centres x products x output channels,
    then the baseline weights, then (P > 1) P shift-decorrelation floats per (freq, baseline),
    then the autocorrelations and their weights, then the phase-cal tones."""
    nf = len(cfg.freqs)
    P = nphasecentres(cfg)
    used, equiv = freq_used(cfg)
    maxprod = max((len(p) for b in cfg.baselines for (_, p) in b.freqs), default=1)
    bpa = 2 if maxprod > 2 else 1
    idx = 0
    boff, bw, sdo = {}, {}, {}
    for f in range(nf):                      # cross-correlations (configuration.cpp:2464-2521)
        if not used[f]:
            continue
        fr = cfg.freqs[f]
        for j, b in enumerate(cfg.baselines):
            for fi, prods in b.freqs:
                if fi == f:
                    boff[(f, j)] = idx
                    idx += P * len(prods) * ((fr.nchan + fr.chanavg - 1) // fr.chanavg)
    for f in range(nf):                      # baseline weights (:2523-2548)
        if not used[f]:
            continue
        for j, b in enumerate(cfg.baselines):
            for fi, prods in b.freqs:
                if fi == f:
                    bw[(f, j)] = idx
                    idx += len(prods) // 2 + len(prods) % 2
    for f in range(nf if P > 1 else 0):      # shift decorrelation factors (:2549-2566)
        if not used[f]:
            continue
        for j, b in enumerate(cfg.baselines):
            for fi, prods in b.freqs:
                if fi == f:
                    sdo[(f, j)] = idx
                    idx += P // 2 + P % 2
    ac = []
    for d in cfg.datastreams:                # autocorrelations (:2567-2586)
        ac.append(idx)
        for lf, _ in d.recbands:
            fi = d.recfreqs[lf]
            if used[fi] or equiv[fi]:
                fr = cfg.freqs[fi]
                idx += bpa * fr.nchan // fr.chanavg
        for lz, _ in d.zoombands:
            fi = d.zoomfreqs[lz][0]
            if used[fi] or equiv[fi]:
                fr = cfg.freqs[fi]
                idx += bpa * fr.nchan // fr.chanavg
    acw = []
    for d in cfg.datastreams:                # autocorrelation weights (:2587-2607)
        acw.append(idx)
        toadd = 0
        for lf, _ in d.recbands:
            fi = d.recfreqs[lf]
            if used[fi] or equiv[fi]:
                toadd += bpa
        for lz, _ in d.zoombands:
            fi = d.zoomfreqs[lz][0]
            if used[fi] or equiv[fi]:
                toadd += bpa
        toadd = (toadd + 1) // 2
        idx += max(toadd, 1)
    pco = []
    for d in cfg.datastreams:                # phase-cal tones (:2611-2619): per recorded freq, its
        pco.append(idx)                      # pols times its tones
        if d.phasecal_interval_hz > 0:
            import pcal as pc
            for lf, fi in enumerate(d.recfreqs):
                fr = cfg.freqs[fi]
                npol = sum(1 for lb, _ in d.recbands if lb == lf)
                nt = pc.tone_table(fr.freq_mhz, fr.bw_mhz, fr.lsb, d.phasecal_interval_hz, d.phasecal_base_hz)[0]
                idx += npol * nt
    return Layout(idx, boff, bw, ac, acw, bpa, sdo, P, pco)


# --------------------------------------------------------------------------------------
# The F-engine (Mode::process) for all FFTs of a subint
# --------------------------------------------------------------------------------------
@dataclass
class StationResult:
    spectra: np.ndarray        # [nbands_total, B, C] complex128 (zoom bands are parent windows)
    weight: np.ndarray         # [B] data weights
    autos: np.ndarray          # [2, nrecbands, C] complex128 (parallel, cross-pol)
    acweight: np.ndarray       # [2, nrecbands]


def fft_params(cfg, d: int, offsets, control, interp, B: int):
    """Per-FFT delay chain of Mode::process (mode.cpp:671-675, 743-757, 776-825) for all
    FFT indices: nearestsample, starttime, the interpolation coefficients a, b (and c for
    order 2), the integer delay and the FFT-centre delay.

    order 0: integerdelay = (int)averagedelay (mode.cpp:778-780); b = averagedelay - integerdelay
    order 1: a = d2 - d0, b = d0 + (d1 - (a/2 + d0))/3 - integerdelay (mode.cpp:781-788)
    order 2: a = q0, b = q1 + 2 i q0, c = q2 + i q1 + i^2 q0 - integerdelay (mode.cpp:800-805)"""
    ds = cfg.datastreams[d]
    L = cfg.fftchannels(d)
    dt = cfg.sampletime_us(d)
    i = np.arange(B, dtype=np.float64)
    q = lambda x: interp[0] * x * x + interp[1] * x + interp[2]
    avgdelay = q(i + 0.5)
    fftstart = i * L * dt
    datasec, datans = control[1], control[2]
    starttime = ((offsets[1] - datasec) * 1000000.0
                 + (float(offsets[2]) - float(datans)) / 1000.0 + fftstart - avgdelay)
    ns = np.trunc(starttime / dt + 0.5).astype(np.int64)
    order = cfg.fringe_rot_order
    c = np.zeros(B)
    if order == 0:
        a = np.zeros(B)
        intdelay = np.trunc(avgdelay).astype(np.int64)
        b = avgdelay - intdelay
    elif order == 1:
        d0, d1, d2 = q(i), q(i + 0.5), q(i + 1.0)
        a = d2 - d0
        b = d0 + (d1 - (a * 0.5 + d0)) / 3.0
        intdelay = np.trunc(b).astype(np.int64)
        b = b - intdelay
    elif order == 2:
        a = np.full(B, float(interp[0]))
        b = interp[1] + i * interp[0] * 2.0
        c = interp[2] + i * interp[1] + i * i * interp[0]
        intdelay = np.trunc(c).astype(np.int64)
        c = c - intdelay
    else:
        raise ValueError("fringe rotation order must be 0, 1 or 2")
    return ns, starttime, a, b, intdelay, avgdelay, c


def rotator_phase(order: int, L: int, S: int, a: float, b: float, c: float, lofreq: float,
                  intdelay: int, fractional_lo: bool, loff: float, sampletime_us: float,
                  intwall: int, fracwall: float) -> np.ndarray:
    """Fringe-rotator phase (turns) of every sample of one FFT, FP64
    (order 1: mode.cpp:883-955; order 2: mode.cpp:956-1026, with the step/sub
    decomposition of mode.cpp:254-261, 790-825).  The rotator is exp(-2 pi i phase).
    Order 0 has no pre-F rotator (zeros)."""
    n = np.arange(L)
    if order == 0:
        return np.zeros(L)
    nfr = L // S
    subxoff = np.arange(S) / float(L)
    stepxoff = (np.arange(nfr) * S) / float(L)
    st = sampletime_us / 1e6
    subtoff = np.arange(S) * st
    steptoff = np.arange(nfr) * S * st
    if order == 1:
        subphase = (subxoff * a + b) * lofreq
        stepphase = (stepxoff * a) * lofreq
        quadphase = np.zeros(S)
    else:
        subphase = (subxoff * (b + a * stepxoff[1]) + c) * lofreq
        quadphase = (subxoff * (2.0 * a * stepxoff[1])) * lofreq
        stepphase = (stepxoff * b + stepxoff * stepxoff * a) * lofreq
    if fractional_lo:
        subphase = subphase + (lofreq - int(lofreq)) * float(intdelay)
    if loff != 0.0:
        fraclo = abs(loff) - int(abs(loff))
        if loff < 0:
            fraclo = -fraclo
        subphase = subphase + subtoff * (-loff) - loff * fracwall
        stepphase = stepphase + steptoff * (-loff)
        if fraclo > TINY or fraclo < -TINY:
            subphase = subphase - fraclo * intwall
    k, j = n % S, n // S
    # order 2 builds sub x quad^j x step (mode.cpp:1017-1024); the quad term is zero for order 1
    return subphase[k] + j * quadphase[k] + stepphase[j]


def postf_phase(avgdelay: float, intdelay: int, lofreq: float, fractional_lo: bool, loff: float,
                walltime: float) -> float:
    """Order 0: the fringe phase applied after the FFT, in turns (mode.cpp:1040-1049); the
    channel phasors become exp(i (chanfreq e - 2 pi frac(phase)))."""
    ph = (avgdelay - intdelay) * lofreq
    if fractional_lo:
        ph += intdelay * (lofreq - int(lofreq))
    ph -= walltime * loff
    return ph


def is_fractional_lo(cfg, ds) -> bool:
    """Mode's fractionalLoFreq (mode.cpp:83-101): some recorded freq's LO is not a whole MHz.
    For double-sideband data the LO is the band centre: the edge moved by BW/2 towards the
    band (down for LSB, up for USB) before the test."""
    for fi in ds.recfreqs:
        fr = cfg.freqs[fi]
        lo = fr.freq_mhz
        if ds.complex_sampling == 2:
            lo += -fr.bw_mhz / 2.0 if fr.lsb else fr.bw_mhz / 2.0
        if abs(lo - int(lo + 0.5)) > TINY:
            return True
    return False


def pol_halfdelta(pol: str, halfdelta: float) -> float:
    """The clock-delta term of a band's fractional-sample phasor (mode.cpp:1030-1122,
    1257-1261): +delta/2 for R (fracsamprotatorA), -delta/2 for every other pol
    (fracsamprotatorB).  With delta 0 both phasors are the same."""
    return halfdelta if pol == "R" else -halfdelta


def phase_offset_phasors(ds) -> np.ndarray:
    """phasecorrA per recorded freq (mode.cpp:421-447): exp(+i pi phi / 360) for a pol phase
    offset of phi degrees, i.e. half the offset, applied to the second pol (mode.cpp:1318-1330)."""
    deg = np.asarray(ds.per_freq("phase_offset_deg"), dtype=np.float64)
    return np.exp(1j * np.radians(deg / 2.0))


def process_station(cfg, d: int, offsets, data: np.ndarray, control: np.ndarray,
                    B: Optional[int] = None) -> StationResult:
    """Mode::setValidFlags/setData/setOffsets + Mode::process for every FFT of the subint
    (core.cpp:695-801), with Mk5Mode::unpack, in float64."""
    ds = cfg.datastreams[d]
    if B is None:
        B = cfg.blockspersend()
    order = cfg.fringe_rot_order
    if order == 0 and ds.complex_sampling:
        raise ValueError("post-F fringe rotation of complex data is rejected (mode.cpp:1130-1133)")
    L = cfg.fftchannels(d)
    dt = cfg.sampletime_us(d)
    nrec = len(ds.recbands)
    nzoom = len(ds.zoombands)
    f0 = cfg.freqs[ds.recfreqs[0]]
    C = f0.nchan
    bw = f0.bw_mhz
    usecomplex = bool(ds.complex_sampling)
    usedouble = ds.complex_sampling == 2                                 # mode.cpp:41-49
    spectra = np.zeros((nrec + nzoom, B, C), dtype=np.complex128)
    weight = np.zeros(B)
    autos = np.zeros((2, nrec, C), dtype=np.complex128)
    acw = np.zeros((2, nrec))
    datalen = len(data)
    scan, datasec, datans = int(control[0]), int(control[1]), int(control[2])
    valid_bits = control[3:]

    # Mode::setOffsets (mode.cpp:1502-1532)
    if datasec <= INVALID_SUBINT or datalen <= 1:
        return StationResult(spectra[:nrec + nzoom], weight, autos, acw)
    if offsets[0] != scan:
        return StationResult(spectra, weight, autos, acw)
    # blockspersend*2*nchan*sampletime/1e6, halved for complex data (mode.cpp:1523-1524)
    timespan = B * C / bw / 1e6
    interp = delay_interpolator(ds.model, ds.model_order, cfg.model_inc_secs,
                                offsets[1] + offsets[2] / 1e9, timespan, B)
    if interp is None:
        return StationResult(spectra, weight, autos, acw)
    interp = interp.copy()
    interp[2] -= 1000000 * intclockseconds(ds)

    vals, valid = decode_stream(data, ds)
    nsamples_buf = vals.shape[1]
    ns, starttime, a, b, intdelay, avgdelay, cq = fft_params(cfg, d, offsets, control, interp, B)

    # bytes-per-block ratio of the reference's bounds check (mode.cpp:124-139, 688)
    nsb = ds.nbits * (2 if usecomplex else 1)
    bpb_num = (nrec * nsb) // 8
    bpb_den = 1
    if bpb_num == 0:
        bpb_num, bpb_den = 1, 8 // (nrec * nsb)
    S = ds.arraystridelen or calc_stride(C)
    fractional_lo = is_fractional_lo(cfg, ds)
    pcorr = phase_offset_phasors(ds)

    for i in range(B):
        if not ((int(valid_bits[i // FLAGS_PER_INT]) >> (i % FLAGS_PER_INT)) & 1):
            continue
        n0 = int(ns[i])
        if n0 < -1 or ((n0 + L) * bpb_num) // bpb_den > datalen:
            continue
        if n0 == -1:
            n0 = 0
        win = slice(n0, n0 + L)
        if n0 + L > nsamples_buf:
            # samples past the last whole frame are treated as blanked
            good = int(valid[n0:].sum()) if n0 < nsamples_buf else 0
            seg = np.zeros((nrec, L), dtype=vals.dtype)
            if n0 < nsamples_buf:
                seg[:, : nsamples_buf - n0] = vals[:, n0:]
        else:
            good = int(valid[win].sum())
            seg = vals[:, win]
        w = good / float(L)
        if not w > 0.0:
            continue
        weight[i] = w
        fse = starttime[i] - n0 * dt                        # mode.cpp:756-757 (kept in f64)
        walltime = cfg.scan_start_sec + offsets[1] + offsets[2] / 1.0e9 + i * L * dt / 1.0e6
        intwall = int(walltime)
        fracwall = walltime - intwall
        for lf, fi in enumerate(ds.recfreqs):
            fr = cfg.freqs[fi]
            lofreq = fr.freq_mhz
            if usecomplex and fr.lsb:
                lofreq = -lofreq                                         # mode.cpp:875-877
            if usedouble:
                lofreq += bw / 2.0                                       # mode.cpp:878-879
            loff = ds.per_freq("lo_offset_hz")[lf]
            # fringe rotator (pre-F for orders 1 and 2), phase in turns, f64
            phase = rotator_phase(order, L, S, a[i], b[i], cq[i], lofreq, int(intdelay[i]), fractional_lo,
                                  loff, dt, intwall, fracwall)
            rot = np.exp(-1j * TWO_PI * phase)
            # fractional-sample correction (mode.cpp:1030-1122), f64: the R band takes
            # fracsamprotatorA (+delta/2), every other pol fracsamprotatorB (-delta/2)
            halfdelta = ds.per_freq("clock_offset_delta_us")[lf] / 2
            k = np.arange(C)
            if usedouble:                                                # dsbstepchannelfreqs
                chanfreq = TWO_PI * (k - C / 2) * bw / C
            elif fr.lsb:                                                 # lsbstepchannelfreqs
                chanfreq = TWO_PI * (k - C) * bw / C
            else:
                chanfreq = TWO_PI * k * bw / C
            postf = 1.0
            if order == 0:                                               # mode.cpp:1040-1049
                pf = postf_phase(avgdelay[i], int(intdelay[i]), lofreq, fractional_lo, loff, walltime)
                postf = np.exp(-1j * TWO_PI * pf)
            match = [j for j, (blf, _) in enumerate(ds.recbands) if blf == lf]
            for j in match:
                e = fse - ds.per_freq("clock_offset_us")[lf] + pol_halfdelta(ds.recbands[j][1], halfdelta)
                frac = np.exp(1j * chanfreq * e) * postf
                x = seg[j].astype(np.complex128)
                X = np.fft.fft(x * rot)                                  # mode.cpp:1174-1183
                if order == 0:
                    # real FFT; LSB conj-flips bins 1..C (mode.cpp:1139-1156)
                    F = np.conj(X[C:0:-1]) if fr.lsb else X[:C]
                elif usecomplex:
                    # channel 0 = lowest sky frequency (mode.cpp:1186-1222)
                    if usedouble and fr.lsb:                             # two conj-flips
                        F = np.empty(C, dtype=np.complex128)
                        F[: C // 2 + 1] = np.conj(X[C // 2:: -1])
                        F[C // 2 + 1:] = np.conj(X[C - 1: C // 2: -1])
                    elif usedouble:                                      # swapped halves
                        F = np.concatenate([X[C // 2:], X[: C // 2]])
                    elif fr.lsb:                                         # DC kept, rest flipped
                        F = np.concatenate([X[:1], np.conj(X[C - 1: 0: -1])])
                    else:
                        F = X[:C]
                elif fr.lsb:
                    F = X[C: 2 * C]                                      # mode.cpp:1205
                else:
                    F = X[:C]                                            # mode.cpp:1215
                F = F * frac                                             # mode.cpp:1257-1261
                spectra[j, i] = F
                autos[0, j] += F * np.conj(F)                           # mode.cpp:1272
                acw[0, j] += w
            if len(match) > 1:
                # the second pol of a dual-pol freq turns by the pol phase offset, after the
                # parallel-hand autocorrelation and before the cross-pol ones
                # (mode.cpp:1318-1330)
                lcp = match[1] if ds.recbands[match[0]][1] == "R" else match[0]
                spectra[lcp, i] *= pcorr[lf]
            if len(match) > 1 and cfg.write_cross_autocorrs:            # mode.cpp:1333-1352
                j0, j1 = match[0], match[1]
                autos[1, j0] += spectra[j0, i] * np.conj(spectra[j1, i])
                autos[1, j1] += spectra[j1, i] * np.conj(spectra[j0, i])
                acw[1, j0] += w
                acw[1, j1] += w
    # zoom bands are windows of their parent band's spectrum (mode.cpp:184-196)
    for z, (lz, pol) in enumerate(ds.zoombands):
        zfi, parent_lf, choff = ds.zoomfreqs[lz]
        zn = cfg.freqs[zfi].nchan
        parent = [j for j, (blf, bp) in enumerate(ds.recbands) if blf == parent_lf and bp == pol][0]
        spectra[nrec + z, :, :zn] = spectra[parent, :, choff: choff + zn]
    return StationResult(spectra, weight, autos, acw)


def calc_stride(n: int) -> int:
    """configuration.cpp:38-51."""
    if n <= 16:
        return n
    g = int(math.sqrt(n + 0.1))
    while g > 0:
        if n % g == 0:
            return n // g
        g -= 1
    return n


def band_record_index(ds, band: int) -> int:
    """Record band of a (possibly zoom) band: the parent recorded band for zooms
    (getBDataStreamXRecordBandIndex)."""
    nrec = len(ds.recbands)
    if band < nrec:
        return band
    lz, pol = ds.zoombands[band - nrec]
    parent_lf = ds.zoomfreqs[lz][1]
    return [j for j, (blf, bp) in enumerate(ds.recbands) if blf == parent_lf and bp == pol][0]


# --------------------------------------------------------------------------------------
# Multi-phase-centre uv shift (Core::uvshiftAndAverageBaselineFreq, core.cpp:1605-1938)
# --------------------------------------------------------------------------------------
def xc_intervals(cfg, B: int):
    """The uv-shift intervals of one subint: each Core process thread owns a contiguous range
    of FFT blocks (core.cpp:2158-2166) and shifts/averages every maxxcblocks FFTs of it, the
    last interval shorter (core.cpp:761-776, 984-991, 1055-1057).  maxxcblocks is the
    model's maximum ns between averages over the FFT duration, rounded down to a multiple of
    numBufferedFFTs (at least numBufferedFFTs).  Returns (first FFT, FFTs, nsoffset, nswidth)
    per interval, nsoffset the interval centre in ns from the subint start."""
    T = max(1, int(getattr(cfg, "process_threads", 1) or 1))
    nbuf = cfg.num_buffered_ffts
    blockns = float(cfg.subint_ns) / float(B)
    maxxc = int(cfg.max_ns_between_xc_avg / blockns)
    maxxc -= maxxc % nbuf
    if maxxc == 0:
        maxxc = nbuf
    out = []
    start, nblk = 0, 0
    for t in range(T):
        start += nblk
        nblk = B // T + (1 if t < B % T else 0)
        k, shift = 0, 0
        while k < nblk:
            n = min(maxxc, nblk - k)
            out.append((start + k, n, (start + shift * maxxc + n / 2.0) * blockns, n * blockns))
            k += n
            shift += 1
    return out


def differential_delays(cfg, bl, offsets, nsoffset: float) -> np.ndarray:
    """differentialdelay[s] = (rate, delay) of every phase centre s of a baseline at the
    interval centre (core.cpp:1636-1670): order-1 interpolators over 1 us at
    offsets[1] + (offsets[2] + nsoffset)/1e9 of the pointing centre (model_delay_us) and of
    each phase centre, per station; the applied delay is the phase-centre minus pointing delay,
    stretched by the pointing rate."""
    d1, d2 = cfg.datastreams[bl.ds1], cfg.datastreams[bl.ds2]
    t = offsets[1] + (offsets[2] + nsoffset) / 1e9
    lin = lambda m, ds: delay_interpolator_linear(m, ds.model_order, cfg.model_inc_secs, t, 0.000001, 1)
    pt1, pt2 = lin(d1.model, d1), lin(d2.model, d2)
    P = nphasecentres(cfg)
    out = np.zeros((P, 2))
    for sc in range(P):
        pc1, pc2 = lin(d1.phasecentre_models[sc], d1), lin(d2.phasecentre_models[sc], d2)
        a1 = pc1[1] - pt1[1]
        a2 = pc2[1] - pt2[1]
        a1 += a1 * pt1[0]
        a2 += a2 * pt2[0]
        out[sc, 1] = a2 - a1
        out[sc, 0] = pc2[0] + pt1[0] - (pc1[0] + pt2[0])
    return out


def uvshift_phasor(fr, delay_us: float) -> np.ndarray:
    """exp(+2 pi i delay f) per channel of a freq, f the channel's sky frequency as the
    rotator of core.cpp:1699-1766 lays it out: the band edge plus c x channel width (USB), or
    the edge minus (nchan - 1 - c) channel widths (LSB)."""
    c = np.arange(fr.nchan)
    chbw = fr.bw_mhz / fr.nchan
    f = fr.freq_mhz - (fr.nchan - 1 - c) * chbw if fr.lsb else fr.freq_mhz + c * chbw
    if abs(delay_us) <= 1.0e-20:
        return np.ones(fr.nchan, dtype=np.complex128)
    return np.exp(2j * np.pi * delay_us * f)


def shift_decorrelation(fr, dd: np.ndarray, nswidth: float) -> float:
    """One interval's contribution nswidth x time-smearing x delay decorrelation
    (core.cpp:1887-1922)."""
    ts, dl = 1.0, 1.0
    if abs(dd[0]) > 1e-18:
        m = 2 * np.pi * dd[0] * (nswidth / 1000.0) * fr.freq_mhz
        ts = math.sin(m / 2.0) / (m / 2.0)
        if ts < 0.0:
            ts = 0.0
    if abs(dd[1]) > 1e-18:
        dl = 1.0 - abs(dd[1] / (fr.nchan / fr.bw_mhz))
        if dl < 0.0:
            dl = 0.0
    return nswidth * ts * dl


def process_subint(cfg, offsets, data: Sequence[np.ndarray], control: Sequence[np.ndarray],
                   B: Optional[int] = None, stations: Optional[List[StationResult]] = None):
    """Core::processdata for one slot (core.cpp:657-1129) with the accumulation of
    uvshiftAndAverageBaselineFreq (core.cpp:1819-1853) and averageAndSendAutocorrs
    (core.cpp:1258-1376).  Returns (results complex128[coreresultlength], stations)."""
    if B is None:
        B = cfg.blockspersend()
    lay = result_layout(cfg)
    res = np.zeros(lay.coreresultlength, dtype=np.complex128)
    if stations is None:
        stations = [process_station(cfg, d, offsets, data[d], control[d], B)
                    for d in range(len(cfg.datastreams))]
    fres = res.view(np.float64)
    used, equiv = freq_used(cfg)
    P = lay.nphasecentres
    # one phase centre: the whole subint at once, unrotated; several: per uv-shift interval,
    # rotated to each phase centre (core.cpp:1730-1879)
    intervals = xc_intervals(cfg, B) if P > 1 else [(0, B, 0.0, 0.0)]
    for j, bl in enumerate(cfg.baselines):
        s1, s2 = stations[bl.ds1], stations[bl.ds2]
        for i0, n, nsoff, nsw in intervals:
            dd = differential_delays(cfg, bl, offsets, nsoff) if P > 1 else np.zeros((1, 2))
            for fi, prods in bl.freqs:
                fr = cfg.freqs[fi]
                nch = fr.nchan
                avg = fr.chanavg
                nout = (nch + avg - 1) // avg
                off = lay.baseline_off[(fi, j)]
                rots = [uvshift_phasor(fr, dd[sc, 1]) if P > 1 else None for sc in range(P)]
                for p, (b1, b2) in enumerate(prods):
                    v = np.einsum("tc,tc->c", s1.spectra[b1, i0:i0 + n, :nch],
                                  np.conj(s2.spectra[b2, i0:i0 + n, :nch]))
                    for sc in range(P):
                        vr = v * rots[sc] if P > 1 else v
                        o = off + (sc * len(prods) + p) * nout
                        if avg == 1:
                            res[o: o + nch] += vr
                        else:
                            pad = np.zeros(nout * avg, dtype=np.complex128)
                            pad[:nch] = vr
                            res[o: o + nout] += pad.reshape(nout, avg).sum(1) / avg
                if P > 1:
                    so = lay.shiftdecorr_off[(fi, j)] * 2
                    for sc in range(P):
                        fres[so + sc] += shift_decorrelation(fr, dd[sc], nsw)
    for j, bl in enumerate(cfg.baselines):
        s1, s2 = stations[bl.ds1], stations[bl.ds2]
        for fi, prods in bl.freqs:
            # baseline weights (core.cpp:1005-1052, 1066-1112): sum over FFTs of w1*w2
            bwo = lay.bweight_off[(fi, j)] * 2
            for p, (b1, b2) in enumerate(prods):
                fres[bwo + p] += float(np.sum(s1.weight * s2.weight))
    # autocorrelations and their weights (core.cpp:1258-1376, mode.cpp:1376-1401)
    for d, ds in enumerate(cfg.datastreams):
        st = stations[d]
        nrec = len(ds.recbands)
        idx = lay.autocorr_off[d]
        for w in range(lay.bandsperautocorr):
            if w == 1 and not cfg.write_cross_autocorrs:
                break
            for k in range(nrec + len(ds.zoombands)):
                if k < nrec:
                    fi = ds.recfreqs[ds.recbands[k][0]]
                else:
                    fi = ds.zoomfreqs[ds.zoombands[k - nrec][0]][0]
                if not (used[fi] or equiv[fi]):
                    continue
                fr = cfg.freqs[fi]
                pf = cfg.freqs[ds.recfreqs[0]]
                cavg = pf.chanavg                 # Mode's channelstoaverage (first recorded freq)
                nout = fr.nchan // fr.chanavg
                if k < nrec:
                    a = st.autos[w, k]
                    av = a.reshape(-1, cavg).mean(1) if cavg > 1 else a
                    res[idx: idx + nout] += av[:nout]
                else:
                    lz, pol = ds.zoombands[k - nrec]
                    zfi, parent_lf, choff = ds.zoomfreqs[lz]
                    parent = band_record_index(ds, k)
                    a = st.autos[w, parent]
                    av = a.reshape(-1, cavg).mean(1) if cavg > 1 else a
                    res[idx: idx + nout] += av[choff // cavg: choff // cavg + nout]
                idx += nout
        widx = lay.acweight_off[d] * 2
        for w in range(lay.bandsperautocorr):
            if w == 1 and not cfg.write_cross_autocorrs:
                break
            for k in range(nrec + len(ds.zoombands)):
                if k < nrec:
                    fi = ds.recfreqs[ds.recbands[k][0]]
                else:
                    fi = ds.zoomfreqs[ds.zoombands[k - nrec][0]][0]
                if not (used[fi] or equiv[fi]):
                    continue
                fres[widx] += st.acweight[w, band_record_index(ds, k)]
                widx += 1
    # phase-cal tones of every recorded band, in band order (Core::copyPCalTones, core.cpp:1131-1163)
    for d, ds in enumerate(cfg.datastreams):
        if ds.phasecal_interval_hz <= 0:
            continue
        import pcal as pc
        tones = pc.extract(cfg, d, offsets, data[d], control[d], ds.phasecal_interval_hz, ds.phasecal_base_hz)
        o = lay.pcal_off[d]
        for t in tones:
            res[o: o + len(t)] += t
            o += len(t)
    return res, stations


def normalise(cfg, res: np.ndarray, lay: Layout, meansubints: float = 1.0) -> np.ndarray:
    """Visibility::writedata's scale without Tsys (visibility.cpp:407-415, 526-527):
    vis / (bweight/fftsperintegration * meansubints * blockspersend * 2 * nchan * chanavg)
    with fftsperintegration = blockspersend*meansubints, for cross-correlations."""
    out = res.copy()
    B = cfg.blockspersend()
    fres = res.view(np.float64)
    for (fi, j), off in lay.baseline_off.items():
        fr = cfg.freqs[fi]
        prods = dict(cfg.baselines[j].freqs)[fi]
        nout = (fr.nchan + fr.chanavg - 1) // fr.chanavg
        for p in range(len(prods)):
            bw = fres[lay.bweight_off[(fi, j)] * 2 + p] / (B * meansubints)
            sc = 0.0 if bw <= 0 else 1.0 / (bw * meansubints * B * 2 * fr.nchan * fr.chanavg)
            out[off + p * nout: off + (p + 1) * nout] *= sc
    return out
