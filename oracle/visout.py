"""CPU restatement of Visibility::writedata's amplitude scaling and of the DIFX binary record
writer (test infrastructure only).

This module is a CHECKER, not the product: only tests/ may import it.

  scaling   Visibility::writedata (mpifxcorr/src/visibility.cpp:353-646) for one phase
            centre and no pulsar bins: baseline and autocorrelation weights = stored weight /
            fftsperintegration (:402-413, :451-462); without Tsys a cross-correlation is scaled
            by 1 / (weight * meansubints * (float)(blockspersend * 2 * freqchannels * chanavg))
            times sqrt(zoom bandwidth / parent bandwidth) for each zoom side (:520-541); with
            Tsys on both datastreams by sqrt(Tsys1 Tsys2 / divisor) / weight, divisor =
            decorr(nbits1) decorr(nbits2) acal1 acal2 / (acw1 acw2) from the band means of
            the parallel autocorrelations (:466-476, :510-515); autocorrelations, when written,
            by 1 / (weight * ...) times the zoom ratio, or Tsys / sqrt(acal_k acal_matching)
            (:588-646).  Float and double steps are kept (numpy float32 / float64).
  records   Visibility::writeSWIN (:781-931) with appendSWINHeaderBuffered (:1205-1234):
            per record a 74-byte header (SYNC_WORD 0xFF00FF00, version 1, baseline, MJD,
            seconds, config, source, freq, polarisation pair, pulsar bin, weight, u, v, w)
            and freqchannels cf32; cross-correlations by baseline, freq, product; then the
            autocorrelations with baseline number 257 (telescope + 1); weight-0 products
            skipped.
  parse     the header layout of the reference's reader (libraries/python/parseDiFX/
            Common.py:28-124), for the tests.
  STA       Core::averageAndSendAutocorrs (mpifxcorr/src/core.cpp:1166-1258): a recorded
            band's parallel autocorrelation summed to the dump channels and scaled by
            1/(2 nchan weight); no record below 0.333 of a full span's weight.

Pins (tests/test_difx_cpu.py): the scale's closed forms (scale x weight x meansubints x
2 x blockspersend x nchan = 1; the zoom factors sqrt(ratio) and ratio; with Tsys1 = Tsys2 and
acal = acw the Tsys scale reduces to Tsys / (decorr^2 ... ) by hand), records parsed back
field by field with the reference reader's layout, and the STA data of a flat
autocorrelation in closed form.
"""
from __future__ import annotations

import struct
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

import fxoracle as fo
import pcal as pcal_restatement

SYNC_WORD = 0xFF00FF00
HEADER_BYTES = 74
DECORR = [0.63662, 0.88, 0.94, 0.96, 0.98, 0.99, 0.996, 0.998]      # mode.cpp:1571
F32 = np.float32
F64 = np.float64


def band_freq(ds, k: int) -> int:
    nrec = len(ds.recbands)
    return ds.recfreqs[ds.recbands[k][0]] if k < nrec else ds.zoomfreqs[ds.zoombands[k - nrec][0]][0]


def band_pol(ds, k: int) -> str:
    nrec = len(ds.recbands)
    return ds.recbands[k][1] if k < nrec else ds.zoombands[k - nrec][1]


def zoom_ratio(cfg, ds, k: int) -> np.float32:
    """modifier = 1; *= zoom bandwidth; /= parent bandwidth, in float (visibility.cpp:525-530)."""
    zf, parent_local, _ = ds.zoomfreqs[ds.zoombands[k - len(ds.recbands)][0]]
    m = F32(1.0)
    m = F32(F64(m) * cfg.freqs[zf].bw_mhz)
    m = F32(F64(m) / cfg.freqs[ds.recfreqs[parent_local]].bw_mhz)
    return m


def matching_band(ds, k: int) -> int:
    """Configuration::getDMatchingBand (configuration.cpp:885-904), recorded bands."""
    if k >= len(ds.recbands):
        return -1
    for i, (lf, _) in enumerate(ds.recbands):
        if i != k and lf == ds.recbands[k][0]:
            return i
    return -1


def _autocorr_bands(cfg, d: int, used, equiv):
    ds = cfg.datastreams[d]
    return [k for k in range(len(ds.recbands) + len(ds.zoombands))
            if used[band_freq(ds, k)] or equiv[band_freq(ds, k)]]


def normalise(cfg, res: np.ndarray, lay, meansubints: float, tsys: Optional[Sequence[float]] = None) -> np.ndarray:
    """The scaled copy of a complex64 results array."""
    out = np.array(res, dtype=np.complex64, copy=True)
    fr32 = out.view(np.float32)
    src = np.asarray(res, dtype=np.complex64).view(np.float32)
    B = cfg.blockspersend()
    msi = F64(meansubints)
    ffts = msi * B
    nds = len(cfg.datastreams)
    tsys = [0.0] * nds if tsys is None else list(tsys)
    used, equiv = fo.freq_used(cfg)
    # parallel autocorrelation band means and weights (j = 0)
    acal: Dict[Tuple[int, int], np.float32] = {}
    acw0: Dict[Tuple[int, int], np.float32] = {}
    for d in range(nds):
        ds = cfg.datastreams[d]
        off, widx = lay.autocorr_off[d], lay.acweight_off[d] * 2
        for k in _autocorr_bands(cfg, d, used, equiv):
            f = cfg.freqs[band_freq(ds, k)]
            fc = f.nchan // f.chanavg
            if tsys[d] > 0:
                acal[(d, k)] = F32(np.mean(src[2 * off:2 * (off + fc):2].astype(np.float64)))
            acw0[(d, k)] = F32(F64(src[widx]) / ffts)
            off += fc
            widx += 1

    def apply(off, n, sc):
        if sc > 0:
            fr32[2 * off:2 * (off + n)] = (fr32[2 * off:2 * (off + n)] * F32(sc)).astype(np.float32)

    for f in range(len(cfg.freqs)):                          # cross-correlations
        if not used[f]:
            continue
        fr = cfg.freqs[f]
        fc = (fr.nchan + fr.chanavg - 1) // fr.chanavg
        chanfactor = F32(B * 2 * fc * fr.chanavg)
        for j, b in enumerate(cfg.baselines):
            for fi, prods in b.freqs:
                if fi != f:
                    continue
                for p, (b1, b2) in enumerate(prods):
                    w = F32(F64(src[lay.bweight_off[(f, j)] * 2 + p]) / ffts)
                    d1, d2 = b.ds1, b.ds2
                    ds1, ds2 = cfg.datastreams[d1], cfg.datastreams[d2]
                    if tsys[d1] > 0 and tsys[d2] > 0:
                        div = F32(DECORR[ds1.nbits - 1]) * F32(DECORR[ds2.nbits - 1]) * acal.get((d1, b1), F32(0)) \
                            * acal.get((d2, b2), F32(0)) / (acw0.get((d1, b1), F32(0)) * acw0.get((d2, b2), F32(0)))
                        div = F32(div)
                        sc = F32(np.sqrt(tsys[d1] * tsys[d2] / F64(div)) / F64(w)) if (div > 0 and w > 0) else F32(0)
                    elif w > 0:
                        sc = F32(1.0 / (F64(w) * msi * F64(chanfactor)))
                        if b1 >= len(ds1.recbands):
                            sc = F32(sc * np.sqrt(zoom_ratio(cfg, ds1, b1)))
                        if b2 >= len(ds2.recbands):
                            sc = F32(sc * np.sqrt(zoom_ratio(cfg, ds2, b2)))
                    else:
                        sc = F32(0)
                    P = lay.nphasecentres
                    for pc in range(P):
                        scp = sc
                        if P > 1:
                            # the shift decorrelation over the integration time (meansubints
                            # subints), visibility.cpp:427-440, 555-557
                            dec = F32(F64(src[lay.shiftdecorr_off[(f, j)] * 2 + pc])
                                      / (1000000000.0 * (msi * F64(cfg.subint_ns) / 1e9)))
                            scp = F32(sc / dec) if dec != 0 else F32(np.inf if sc > 0 else 0.0)
                        apply(lay.baseline_off[(f, j)] + (pc * len(prods) + p) * fc, fc, scp)
    if cfg.write_cross_autocorrs:                            # autocorrelations
        for d in range(nds):
            ds = cfg.datastreams[d]
            off, widx = lay.autocorr_off[d], lay.acweight_off[d] * 2
            for j in range(lay.bandsperautocorr):
                for k in _autocorr_bands(cfg, d, used, equiv):
                    f = cfg.freqs[band_freq(ds, k)]
                    fc = f.nchan // f.chanavg
                    w = F32(F64(src[widx]) / ffts)
                    if tsys[d] > 0:
                        mk = k if j == 0 else matching_band(ds, k)
                        div = F32(np.sqrt(F32(acal.get((d, k), F32(0)) * acal.get((d, mk), F32(0)))))
                        sc = F32(tsys[d] / F64(div)) if div > 0 else F32(0)
                    elif w > 0:
                        sc = F32(1.0 / (F64(w) * msi * F64(F32(B * 2 * fc * f.chanavg))))
                        if k >= len(ds.recbands):
                            sc = F32(sc * zoom_ratio(cfg, ds, k))
                    else:
                        sc = F32(0)
                    apply(off, fc, sc)
                    off += fc
                    widx += 1
    for d in range(nds):                                     # phase-cal tones (visibility.cpp:650-675)
        ds = cfg.datastreams[d]
        if ds.phasecal_interval_hz <= 0:
            continue
        off = lay.pcal_off[d]
        for k in range(len(ds.recbands)):
            f = cfg.freqs[band_freq(ds, k)]
            nt = pcal_restatement.tone_table(f.freq_mhz, f.bw_mhz, f.lsb, ds.phasecal_interval_hz,
                                             ds.phasecal_base_hz)[0]
            acw = acw0.get((d, k), F32(0))
            if acw == 0 and len(ds.zoombands) > 0:               # zoom mode without matching autos
                acw = F32(1.0)
            if acw > 0:
                fc = f.nchan // f.chanavg
                apply(off, nt, F32(1.0 / (F64(acw) * msi * F64(F32(B * 2 * fc * f.chanavg)))))
            off += nt
    return out


def opposite_pol(p: str) -> str:
    return {"R": "L", "L": "R", "X": "Y", "Y": "X"}.get(p, "X")


def header(bl, mjd, sec, cfgi, src, freq, pol, weight, uvw, pulsarbin=0) -> bytes:
    return struct.pack("<Iiiidiiiccidddd", SYNC_WORD, 1, bl, mjd, sec, cfgi, src, freq,
                       pol[0].encode(), pol[1].encode(), pulsarbin, weight, *uvw)


def records(cfg, res: np.ndarray, lay, mjd: int, seconds: float, configindex: int, sourceindex: int,
            fftsperintegration: float, baselinenumber: Sequence[int], uvw: np.ndarray,
            telescopeindex: Sequence[int], phasecentre: int = 0) -> bytes:
    """writeSWIN's records of a (scaled) results array; with several phase centres those of
    one phase centre's file (visibility.cpp:848-885), the autocorrelations in the first."""
    src = np.asarray(res, dtype=np.complex64)
    f32 = src.view(np.float32)
    used, equiv = fo.freq_used(cfg)
    P = lay.nphasecentres
    pc = phasecentre if P > 1 else 0
    inttime_s = fftsperintegration / cfg.blockspersend() * F64(cfg.subint_ns) / 1e9
    out = []
    for i, b in enumerate(cfg.baselines):
        for f in range(len(cfg.freqs)):
            prods = dict(b.freqs).get(f)
            if prods is None:
                continue
            fr = cfg.freqs[f]
            fc = fr.nchan // fr.chanavg
            core = lay.baseline_off[(f, i)]
            for p, (b1, b2) in enumerate(prods):
                w = F32(F64(f32[lay.bweight_off[(f, i)] * 2 + p]) / fftsperintegration)
                if not w > 0:
                    continue
                if P > 1:                                   # visibility.cpp:869-873
                    w = F32(w * F32(F64(f32[lay.shiftdecorr_off[(f, i)] * 2 + pc]) / (1000000000.0 * inttime_s)))
                pol = band_pol(cfg.datastreams[b.ds1], b1) + band_pol(cfg.datastreams[b.ds2], b2)
                out.append(header(baselinenumber[i], mjd, seconds, configindex, sourceindex, f, pol, float(w),
                                  [float(x) for x in uvw[i]]))
                o = core + (pc * len(prods) + p) * fc
                out.append(src[o: o + fc].tobytes())
    if cfg.write_cross_autocorrs and pc == 0:
        label = fo.autocorr_freq_labels(cfg)            # USB x LSB (visibility.cpp:912-925)
        for d, ds in enumerate(cfg.datastreams):
            off, widx = lay.autocorr_off[d], lay.acweight_off[d] * 2
            for j in range(lay.bandsperautocorr):
                for k in range(len(ds.recbands) + len(ds.zoombands)):
                    fi = label[band_freq(ds, k)]
                    if not (used[fi] or equiv[fi]):
                        continue
                    fc = cfg.freqs[fi].nchan // cfg.freqs[fi].chanavg
                    w = F32(F64(f32[widx]) / fftsperintegration)
                    if w > 0:
                        p0 = band_pol(ds, k)
                        out.append(header(257 * (telescopeindex[d] + 1), mjd, seconds, configindex, sourceindex, fi,
                                          p0 + (p0 if j == 0 else opposite_pol(p0)), float(w), [0.0, 0.0, 0.0]))
                        out.append(src[off:off + fc].tobytes())
                    off += fc
                    widx += 1
    return b"".join(out)


def parse(buf: bytes, nchan_of_freq: Dict[int, int]) -> List[dict]:
    """Records in the layout of the reference reader (parseDiFX/Common.py:79-124)."""
    recs, pos = [], 0
    while pos < len(buf):
        (sync, ver, bl, mjd, sec, cfgi, src, freq, p0, p1, pbin, w, u, v, ww) = \
            struct.unpack_from("<Iiiidiiiccidddd", buf, pos)
        assert sync == SYNC_WORD and ver == 1
        pos += HEADER_BYTES
        n = nchan_of_freq[freq]
        vis = np.frombuffer(buf, dtype=np.complex64, count=n, offset=pos)
        pos += 8 * n
        recs.append(dict(baseline=bl, mjd=mjd, seconds=sec, config=cfgi, source=src, freq=freq,
                         pol=(p0 + p1).decode(), bin=pbin, weight=w, uvw=(u, v, ww), vis=vis))
    return recs


def sta_autocorr(cfg, res: np.ndarray, lay, d: int, band: int, stadumpchannels: int, nswidth: float):
    """Core::averageAndSendAutocorrs' STA record data (core.cpp:1166-1258) of recorded band
    `band` of datastream d from a results array: None when its weight is below 0.333 of a
    full span's (no record)."""
    src = np.asarray(res, dtype=np.complex64)
    f32 = src.view(np.float32)
    used, equiv = fo.freq_used(cfg)
    minpostav = min(cfg.freqs[f].nchan // cfg.freqs[f].chanavg for f in range(len(cfg.freqs)) if used[f])
    averaged = minpostav >= stadumpchannels
    ds = cfg.datastreams[d]
    bands = _autocorr_bands(cfg, d, used, equiv)
    off, widx = lay.autocorr_off[d], lay.acweight_off[d] * 2
    for k in bands:
        if k == band:
            break
        f = cfg.freqs[band_freq(ds, k)]
        off += f.nchan // f.chanavg
        widx += 1
    fr = cfg.freqs[band_freq(ds, band)]
    fc = fr.nchan
    stasamples = 0.001 * nswidth * 2 * fr.bw_mhz
    if F64(f32[widx]) < 0.333 * stasamples / (2 * fc):
        return None
    renorm = F32(1.0 / F64(F32(2 * fc) * f32[widx]))
    if averaged:
        renorm = F32(renorm / F32(fr.chanavg))
        fc //= fr.chanavg
    n = min(stadumpchannels, fc)
    cta = fc // n
    out = np.zeros(n, dtype=np.float32)
    for k in range(n):
        a = f32[2 * (off + k * cta)]
        for l in range(1, cta):
            a = F32(a + f32[2 * (off + k * cta + l)])
        out[k] = F32(a * renorm)
    return out
