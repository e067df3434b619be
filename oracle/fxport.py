"""ctypes front end of the C restatement oracle/fxport.c (test infrastructure; timed CPU
baseline).  The per-subint bookkeeping -- delay interpolators (fxoracle.delay_interpolator),
the product list and the placement of products into the populateResultLengths layout
(fxoracle.result_layout) -- is done here; the per-sample work runs in C."""
from __future__ import annotations

import ctypes as C
import math
import os
from pathlib import Path
from typing import List, Optional, Sequence

import numpy as np

import fxoracle as fo

HERE = Path(__file__).resolve().parent
LIB = HERE / "libfxport.so"
_lib = None


class port_freq(C.Structure):
    _fields_ = [("lofreq", C.c_double), ("lo_offset_hz", C.c_double), ("fraclooffset", C.c_double),
                ("clockoff", C.c_double), ("halfdelta", C.c_double),
                ("lsb", C.c_int), ("looff", C.c_int), ("isfraclooffset", C.c_int), ("nbands", C.c_int),
                ("band", C.c_int * 4), ("bhd", C.c_double * 4), ("pcre", C.c_float * 4),
                ("pcim", C.c_float * 4)]


class port_ds(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("fmt", "nbits", "cplx", "framebytes", "headerbytes", "spf", "nrec",
                                       "nrecfreqs", "L", "C", "S", "NS", "NFS", "bpb_num", "bpb_den",
                                       "fractionallo", "valid", "dsb")] + [
        ("sampletime_us", C.c_double), ("interp", C.c_double * 3),
        ("datasec", C.c_int), ("datans", C.c_int), ("datalen", C.c_int), ("nframes", C.c_int),
        ("data", C.c_void_p), ("vflags", C.c_void_p), ("chanfreq", C.c_void_p),
        ("rf", port_freq * 32)]


class port_prod(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("ds1", "band1", "choff1", "ds2", "band2", "choff2", "nchan")] + [
        ("acc_off", C.c_longlong)]


def available() -> bool:
    try:
        _load()
        return True
    except OSError:
        return False


def _load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise OSError(f"{LIB} not built (python oracle/build_port.py)")
        lib = C.CDLL(str(LIB))
        lib.fxport_run.argtypes = [C.c_int, C.POINTER(port_ds), C.c_int, C.c_int, C.c_int, C.c_double,
                                   C.c_int, C.POINTER(port_prod), C.c_longlong, C.c_void_p, C.c_void_p,
                                   C.c_int, C.c_int, C.c_int]
        lib.fxport_run.restype = C.c_int
        lib.fxport_fft_backend.restype = C.c_int
        _lib = lib
    return _lib


def default_threads() -> int:
    return max(1, min(16, os.cpu_count() or 1))


def fft_backend() -> str:
    return "FFTW3 API (MKL libmkl_rt)" if _load().fxport_fft_backend() == 1 else "own radix-2"


def _chanfreq(C_, S, NFS, bw):
    """Channel slopes as the reference computes them (mode.cpp:303-324): sub, USB step,
    LSB step and double-sideband step tables."""
    two_pi = fo.TWO_PI
    sub = [np.float32((two_pi * i * bw) / C_) for i in range(S)]
    usb = [np.float32((two_pi * i * S * bw) / C_) for i in range(NFS)]
    lsb = [np.float32((-two_pi * ((NFS - i) * S) * bw) / C_) for i in range(NFS)]
    dsb = [np.float32((two_pi * i * S * bw) / C_ - two_pi * bw / 2.0) for i in range(NFS)]
    return np.array(sub + usb + lsb + dsb, dtype=np.float32)


def process_subint(cfg, offsets, data, control, nthreads: Optional[int] = None, nbuffered: Optional[int] = None):
    """Results (complex64[coreresultlength]) of one slot, computed by the C restatement."""
    lib = _load()
    if nthreads is None:
        nthreads = default_threads()
    if nbuffered is None:
        nbuffered = cfg.num_buffered_ffts
    B = cfg.blockspersend()
    nds = len(cfg.datastreams)
    arr = (port_ds * nds)()
    keep: List = []
    for d, ds in enumerate(cfg.datastreams):
        x = arr[d]
        f0 = cfg.freqs[ds.recfreqs[0]]
        Cn = f0.nchan
        L = cfg.fftchannels(d)
        S = ds.arraystridelen or fo.calc_stride(Cn)
        nrec = len(ds.recbands)
        nsb = ds.nbits * (2 if ds.complex_sampling else 1)
        x.fmt = 1 if ds.format.upper() == "MARK5B" else 0
        x.nbits, x.cplx = ds.nbits, int(bool(ds.complex_sampling))
        x.dsb = int(ds.complex_sampling == 2)
        x.framebytes, x.headerbytes = ds.framebytes, ds.headerbytes
        x.spf = fo.samples_per_frame(ds)
        x.nrec, x.nrecfreqs = nrec, len(ds.recfreqs)
        x.L, x.C, x.S, x.NS, x.NFS = L, Cn, S, L // S, Cn // S
        num = (nrec * nsb) // 8
        x.bpb_num, x.bpb_den = (num, 1) if num else (1, 8 // (nrec * nsb))
        x.fractionallo = int(fo.is_fractional_lo(cfg, ds))
        x.sampletime_us = cfg.sampletime_us(d)
        ctrl = np.ascontiguousarray(control[d], dtype=np.int32)
        buf = np.ascontiguousarray(data[d], dtype=np.uint8)
        keep += [ctrl, buf]
        x.datasec, x.datans = int(ctrl[1]), int(ctrl[2])
        x.datalen = len(buf)
        x.nframes = len(buf) // ds.framebytes
        x.data = buf.ctypes.data
        x.vflags = ctrl[3:].ctypes.data
        valid = x.datasec > fo.INVALID_SUBINT and x.datalen > 1 and int(ctrl[0]) == offsets[0]
        interp = fo.delay_interpolator(ds.model, ds.model_order, cfg.model_inc_secs,
                                       offsets[1] + offsets[2] / 1e9, B * Cn / f0.bw_mhz / 1e6, B)
        if interp is None:
            valid = False
            interp = np.zeros(3)
        interp = interp.copy()
        interp[2] -= 1e6 * fo.intclockseconds(ds)
        x.valid = int(valid)
        for k in range(3):
            x.interp[k] = float(interp[k])
        cf = _chanfreq(Cn, S, Cn // S, f0.bw_mhz)
        keep.append(cf)
        x.chanfreq = cf.ctypes.data
        for lf, fi in enumerate(ds.recfreqs):
            fr = cfg.freqs[fi]
            r = x.rf[lf]
            r.lofreq = -fr.freq_mhz if (ds.complex_sampling and fr.lsb) else fr.freq_mhz
            if ds.complex_sampling == 2:
                r.lofreq += fr.bw_mhz / 2.0                       # mode.cpp:878-879
            lo = ds.per_freq("lo_offset_hz")[lf]
            r.lo_offset_hz = lo
            r.looff = int(lo != 0.0)
            fl = abs(lo) - int(abs(lo))
            r.isfraclooffset = int(r.looff and fl > fo.TINY)
            r.fraclooffset = (-fl if lo < 0 else fl) if r.looff else 0.0
            r.clockoff = ds.per_freq("clock_offset_us")[lf]
            r.halfdelta = ds.per_freq("clock_offset_delta_us")[lf] / 2
            r.lsb = int(fr.lsb)
            bands = [j for j, (blf, _) in enumerate(ds.recbands) if blf == lf]
            r.nbands = len(bands)
            lcp = -1
            if len(bands) > 1:                  # mode.cpp:1318-1330: the second pol turns
                lcp = bands[1] if ds.recbands[bands[0]][1] == "R" else bands[0]
            deg = ds.per_freq("phase_offset_deg")[lf]
            # phasecorrA (mode.cpp:421-447), double rounded to float
            pcre = float(np.float32(math.cos(-deg / 2 * math.pi / 180.0)))
            pcim = float(np.float32(-math.sin(-deg / 2 * math.pi / 180.0)))
            for k, b in enumerate(bands[:4]):
                r.band[k] = b
                r.bhd[k] = fo.pol_halfdelta(ds.recbands[b][1], r.halfdelta)
                r.pcre[k], r.pcim[k] = (pcre, pcim) if b == lcp else (1.0, 0.0)
    # products: baseline products then autocorrelation products (parallel, cross-pol)
    lay = fo.result_layout(cfg)
    used, equiv = fo.freq_used(cfg)
    prods, place = [], []
    acc = 0

    def add(d1, b1, c1, d2, b2, c2, n):
        nonlocal acc
        p = port_prod()
        p.ds1, p.band1, p.choff1, p.ds2, p.band2, p.choff2, p.nchan = d1, b1, c1, d2, b2, c2, n
        p.acc_off = acc
        acc += n
        prods.append(p)
        return p.acc_off

    def src(ds, band):
        nrec = len(ds.recbands)
        if band < nrec:
            return band, 0
        lz, pol = ds.zoombands[band - nrec]
        return fo.band_record_index(ds, band), ds.zoomfreqs[lz][2]

    for j, bl in enumerate(cfg.baselines):
        d1, d2 = cfg.datastreams[bl.ds1], cfg.datastreams[bl.ds2]
        for fi, pp in bl.freqs:
            fr = cfg.freqs[fi]
            nout = (fr.nchan + fr.chanavg - 1) // fr.chanavg
            for p, (b1, b2) in enumerate(pp):
                r1, c1 = src(d1, b1)
                r2, c2 = src(d2, b2)
                o = add(bl.ds1, r1, c1, bl.ds2, r2, c2, fr.nchan)
                place.append((o, lay.baseline_off[(fi, j)] + p * nout, nout, fr.chanavg))
    for d, ds in enumerate(cfg.datastreams):
        nrec = len(ds.recbands)
        Cn = cfg.freqs[ds.recfreqs[0]].nchan
        cavg = cfg.freqs[ds.recfreqs[0]].chanavg
        aa = [add(d, b, 0, d, b, 0, Cn) for b in range(nrec)]
        xa = [None] * nrec
        if cfg.write_cross_autocorrs:
            for b in range(nrec):
                lf = ds.recbands[b][0]
                m = [k for k, (blf, _) in enumerate(ds.recbands) if blf == lf]
                if len(m) > 1 and b in m[:2]:
                    other = m[1] if b == m[0] else m[0]
                    xa[b] = add(d, b, 0, d, other, 0, Cn)
        idx = lay.autocorr_off[d]
        for w in range(lay.bandsperautocorr):
            if w == 1 and not cfg.write_cross_autocorrs:
                break
            for k in range(nrec + len(ds.zoombands)):
                fi = ds.recfreqs[ds.recbands[k][0]] if k < nrec else ds.zoomfreqs[ds.zoombands[k - nrec][0]][0]
                if not (used[fi] or equiv[fi]):
                    continue
                fr = cfg.freqs[fi]
                nout = fr.nchan // fr.chanavg
                rec, choff = src(ds, k)
                o = aa[rec] if w == 0 else xa[rec]
                if o is not None:
                    place.append((o + choff, idx, nout, cavg))
                idx += nout
    parr = (port_prod * max(len(prods), 1))(*prods)
    accbuf = np.zeros(max(acc, 1), np.complex64)
    weights = np.zeros((nds, B), np.float32)
    lib.fxport_run(nds, arr, B, int(offsets[1]), int(offsets[2]), float(cfg.scan_start_sec),
                   len(prods), parr, acc, accbuf.ctypes.data, weights.ctypes.data, int(nthreads), int(nbuffered),
                   int(cfg.fringe_rot_order))
    res = np.zeros(lay.coreresultlength, np.complex64)
    for o, dst, nout, avg in place:
        v = accbuf[o: o + nout * avg]
        res[dst: dst + nout] += v.reshape(nout, avg).sum(1) / avg if avg > 1 else v
    fres = res.view(np.float32)
    for (fi, j), off in lay.bweight_off.items():
        bl = cfg.baselines[j]
        npp = len(dict(bl.freqs)[fi])
        s = np.float32(np.sum(weights[bl.ds1].astype(np.float64) * weights[bl.ds2]))
        for p in range(npp):
            fres[off * 2 + p] += s
    for d, ds in enumerate(cfg.datastreams):
        nrec = len(ds.recbands)
        widx = lay.acweight_off[d] * 2
        s = np.float32(np.sum(weights[d].astype(np.float64)))
        for w in range(lay.bandsperautocorr):
            if w == 1 and not cfg.write_cross_autocorrs:
                break
            for k in range(nrec + len(ds.zoombands)):
                fi = ds.recfreqs[ds.recbands[k][0]] if k < nrec else ds.zoomfreqs[ds.zoombands[k - nrec][0]][0]
                if not (used[fi] or equiv[fi]):
                    continue
                rec, _ = src(ds, k)
                if w == 0:
                    fres[widx] += s
                else:
                    lf = ds.recbands[rec][0]
                    m = [kk for kk, (blf, _) in enumerate(ds.recbands) if blf == lf]
                    if len(m) > 1 and rec in m[:2]:
                        fres[widx] += s
                widx += 1
    return res, weights
