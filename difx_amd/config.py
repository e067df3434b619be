"""Job description for the FX hot path and its conversion to the C ABI structures.

The dataclasses carry the parts of a DiFX Configuration that the per-subint FX path reads:
the frequency table, the datastream table (with the per-station delay polynomials of the
Model), the baseline table with its polarisation products, and the per-config parameters
(mpifxcorr/src/configuration.h:781-950, model.h).  `standard_config` builds the array
layouts used by the benchmarks and tests in the same way vex2difx lays them out: one
datastream per station, every recorded band in one sample word, baselines in station order
with parallel-hand products before cross-hand products.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import fxg_abi as abi

TINY = 1e-9  # Mode::TINY (mode.cpp:27)


@dataclass
class Freq:
    freq_mhz: float          # band-edge sky frequency, MHz
    bw_mhz: float
    nchan: int
    lsb: bool = False
    chanavg: int = 1


@dataclass
class Datastream:
    recfreqs: List[int]                      # freq table indices of the recorded freqs
    recbands: List[Tuple[int, str]]          # (local recorded freq, pol) per band, sample-word order
    model: np.ndarray                        # [nsamples, 6] delay coefficients, microseconds
    format: str = "VDIF"                     # VDIF or MARK5B
    nbits: int = 2
    complex_sampling: int = 0                # 0 real, 1 complex, 2 complex double sideband
    payloadbytes: int = 8000
    headerbytes: int = 32
    arraystridelen: int = 0
    clock_offset_us: Optional[List[float]] = None
    clock_offset_delta_us: Optional[List[float]] = None
    phase_offset_deg: Optional[List[float]] = None
    lo_offset_hz: Optional[List[float]] = None
    zoomfreqs: List[Tuple[int, int, int]] = field(default_factory=list)  # (freq idx, parent local freq, chan offset)
    zoombands: List[Tuple[int, str]] = field(default_factory=list)       # (local zoom freq, pol)
    clock_coeff0_us: float = 0.0
    model_order: int = 5
    # delay polynomials of phase centres 1..P, each [nsamples, 6] like `model` (which is then
    # the pointing centre's); only read when the config has more than one phase centre
    phasecentre_models: Optional[List[np.ndarray]] = None
    # phase-cal tone interval and base (Configuration's phasecalintervalhz, phasecalbasehz):
    # a nonzero interval extracts the tones every subint into the results' phase-cal block
    phasecal_interval_hz: int = 0
    phasecal_base_hz: int = 0

    @property
    def framebytes(self) -> int:
        return self.payloadbytes + self.headerbytes

    def per_freq(self, name: str) -> List[float]:
        v = getattr(self, name)
        return list(v) if v is not None else [0.0] * len(self.recfreqs)


@dataclass
class Baseline:
    ds1: int
    ds2: int
    freqs: List[Tuple[int, List[Tuple[int, int]]]]  # (freq table index, [(band1, band2) per product])


@dataclass
class Config:
    freqs: List[Freq]
    datastreams: List[Datastream]
    baselines: List[Baseline]
    subint_ns: int
    guard_ns: int = 1000
    fringe_rot_order: int = 1
    xmac_stride_len: int = 0
    num_buffered_ffts: int = 10
    write_cross_autocorrs: bool = True
    model_inc_secs: float = 120.0
    scan_start_sec: float = 0.0
    xengine: str = "fp32"
    chunk_ffts: int = 0
    fengine_mode: int = 0       # 0 automatic, 1 one launch per FFT length, 2 one persistent launch
    fengine_threads: int = 0    # F-engine workgroup size 256 or 512 (0: by FFT length)
    autotune: int = 0           # 1: time the launch candidates at create and keep the fastest
    nphasecentres: int = 1      # phase centres of the scan (Model::getNumPhaseCentres)
    process_threads: int = 1    # Core process threads, whose block ranges set the uv-shift intervals
    max_ns_between_xc_avg: int = 2_000_000_000   # Model::getMaxNSBetweenXCAvg

    # ---- derived quantities (configuration.cpp) ----
    def blockspersend(self) -> int:
        """configuration.cpp:1843-1852: FFTs per subint from the first datastream's first
        recorded freq; a non-integer count rejects the job."""
        f = self.freqs[self.datastreams[0].recfreqs[0]]
        ffttime = 1000.0 * f.nchan / f.bw_mhz
        b = self.subint_ns / ffttime
        bi = int(b + 0.5)
        if abs(b - bi) > TINY:
            raise ValueError(f"subint {self.subint_ns} ns is not an integer number of FFTs ({ffttime} ns)")
        return bi

    def fftchannels(self, ds: int) -> int:
        d = self.datastreams[ds]
        n = self.freqs[d.recfreqs[0]].nchan
        return n if d.complex_sampling else 2 * n

    def sampletime_us(self, ds: int) -> float:
        d = self.datastreams[ds]
        bw = self.freqs[d.recfreqs[0]].bw_mhz
        return (1.0 / bw) if d.complex_sampling else 1.0 / (2.0 * bw)

    def maxproducts(self) -> int:
        return max(len(p) for b in self.baselines for (_, p) in b.freqs)

    # ---- conversion to the C ABI ----
    def to_abi(self):
        """Return (fxg_config, keepalive) with all arrays owned by keepalive."""
        keep = []
        nf = len(self.freqs)
        freqs = (abi.fxg_freq * nf)()
        for i, f in enumerate(self.freqs):
            freqs[i].freq_mhz = f.freq_mhz
            freqs[i].bw_mhz = f.bw_mhz
            freqs[i].lsb = int(f.lsb)
            freqs[i].nchan = f.nchan
            freqs[i].chanavg = f.chanavg
        keep.append(freqs)
        nd = len(self.datastreams)
        dss = (abi.fxg_datastream * nd)()
        for i, d in enumerate(self.datastreams):
            x = dss[i]
            x.format = abi.FXG_FMT_MARK5B if d.format.upper() == "MARK5B" else abi.FXG_FMT_VDIF
            x.nbits = d.nbits
            x.complex_sampling = int(d.complex_sampling)
            x.framebytes = d.framebytes
            x.headerbytes = d.headerbytes
            x.arraystridelen = d.arraystridelen
            x.nrecfreqs = len(d.recfreqs)
            for j, fi in enumerate(d.recfreqs):
                x.recfreq_index[j] = fi
            for j, v in enumerate(d.per_freq("clock_offset_us")):
                x.clock_offset_us[j] = v
            for j, v in enumerate(d.per_freq("clock_offset_delta_us")):
                x.clock_offset_delta_us[j] = v
            for j, v in enumerate(d.per_freq("phase_offset_deg")):
                x.phase_offset_deg[j] = v
            for j, v in enumerate(d.per_freq("lo_offset_hz")):
                x.lo_offset_hz[j] = v
            x.nrecbands = len(d.recbands)
            for j, (lf, pol) in enumerate(d.recbands):
                x.recband_localfreq[j] = lf
                x.recband_pol[j] = ord(pol)
            x.nzoomfreqs = len(d.zoomfreqs)
            for j, (fi, parent, off) in enumerate(d.zoomfreqs):
                x.zoomfreq_index[j] = fi
                x.zoomfreq_parent_localfreq[j] = parent
                x.zoomfreq_chanoffset[j] = off
            x.nzoombands = len(d.zoombands)
            for j, (lz, pol) in enumerate(d.zoombands):
                x.zoomband_localfreq[j] = lz
                x.zoomband_pol[j] = ord(pol)
            x.clock_coeff0_us = d.clock_coeff0_us
            x.model_order = d.model_order
            m = np.ascontiguousarray(d.model, dtype=np.float64)
            assert m.ndim == 2 and m.shape[1] == abi.FXG_MAX_POLY
            keep.append(m)
            x.model_nsamples = m.shape[0]
            x.model_delay_us = m.ctypes.data_as(C.POINTER(C.c_double))
            if self.nphasecentres > 1:
                pcm = np.ascontiguousarray(np.stack(d.phasecentre_models[: self.nphasecentres]), dtype=np.float64)
                assert pcm.shape == (self.nphasecentres,) + m.shape
                keep.append(pcm)
                x.phasecentre_delay_us = pcm.ctypes.data_as(C.POINTER(C.c_double))
            x.phasecal_interval_hz = int(d.phasecal_interval_hz)
            x.phasecal_base_hz = int(d.phasecal_base_hz)
        keep.append(dss)
        nb = len(self.baselines)
        bls = (abi.fxg_baseline * max(nb, 1))()
        for i, b in enumerate(self.baselines):
            x = bls[i]
            x.ds1, x.ds2 = b.ds1, b.ds2
            x.nfreqs = len(b.freqs)
            for j, (fi, prods) in enumerate(b.freqs):
                x.freq_index[j] = fi
                x.npolprod[j] = len(prods)
                for p, (b1, b2) in enumerate(prods):
                    x.ds1band[j][p] = b1
                    x.ds2band[j][p] = b2
        keep.append(bls)
        cfg = abi.fxg_config()
        cfg.nfreqs = nf
        cfg.freqs = C.cast(freqs, C.POINTER(abi.fxg_freq))
        cfg.ndatastreams = nd
        cfg.datastreams = C.cast(dss, C.POINTER(abi.fxg_datastream))
        cfg.nbaselines = nb
        cfg.baselines = C.cast(bls, C.POINTER(abi.fxg_baseline))
        cfg.subint_ns = int(self.subint_ns)
        cfg.guard_ns = int(self.guard_ns)
        cfg.fringe_rot_order = self.fringe_rot_order
        cfg.xmac_stride_len = self.xmac_stride_len
        cfg.num_buffered_ffts = self.num_buffered_ffts
        cfg.write_cross_autocorrs = int(self.write_cross_autocorrs)
        cfg.model_inc_secs = self.model_inc_secs
        cfg.scan_start_sec = self.scan_start_sec
        cfg.xengine = {"fp32": abi.FXG_X_FP32, "fp16": abi.FXG_X_FP16,
                       "bf16": abi.FXG_X_BF16}[self.xengine]
        cfg.chunk_ffts = self.chunk_ffts
        cfg.fengine_mode = self.fengine_mode
        cfg.fengine_threads = self.fengine_threads
        cfg.autotune = self.autotune
        cfg.nphasecentres = self.nphasecentres
        cfg.process_threads = self.process_threads
        cfg.max_ns_between_xc_avg = int(self.max_ns_between_xc_avg)
        return cfg, keep


def _with(cfg: Config, extra: dict) -> Config:
    """Set further Config fields (fringe_rot_order, fengine_mode, fengine_threads, autotune, ...)."""
    for k, v in extra.items():
        if not hasattr(cfg, k):
            raise TypeError(f"Config has no field {k!r}")
        setattr(cfg, k, v)
    return cfg


def random_delay_model(rng: np.random.Generator, nsamples: int, max_delay_us: float = 1e4,
                       max_rate_us_per_s: float = 1.0, model_inc_secs: float = 120.0) -> np.ndarray:
    """Seeded quintic delay polynomials (microseconds, t in seconds from interval start):
    |delay| <= max_delay_us, |rate| <= max_rate_us_per_s; higher terms small so the rate
    bound holds over each interval."""
    m = np.zeros((nsamples, abi.FXG_MAX_POLY))
    d0 = rng.uniform(-max_delay_us * 0.9, max_delay_us * 0.9)
    r0 = rng.uniform(-max_rate_us_per_s * 0.8, max_rate_us_per_s * 0.8)
    T = model_inc_secs
    hi = rng.uniform(-1, 1, size=4)
    for s in range(nsamples):
        t0 = s * T
        # a smooth global quintic, re-expanded about each interval start
        c = np.array([d0, r0, 1e-4 * hi[0], 1e-7 * hi[1], 1e-10 * hi[2], 1e-13 * hi[3]])
        # Taylor shift of the global polynomial to t0
        shifted = np.zeros(6)
        for k in range(6):
            for j in range(k, 6):
                shifted[k] += c[j] * math.comb(j, k) * t0 ** (j - k)
        m[s] = shifted
    return m


def calcstridelength(n: int) -> int:
    """configuration.cpp:38-51: the divisor of n closest to but not below sqrt(n)."""
    if n <= 16:
        return n
    g = int(math.sqrt(n + 0.1))
    while g > 0:
        if n % g == 0:
            return n // g
        g -= 1
    return n


def standard_config(nstations: int, npol: int, freqs_mhz: Sequence[float], bw_mhz: float,
                    nchan: int, subint_ns: int, nbits: int = 2, complex_sampling: int = 0,
                    payloadbytes: int = 8000, seed: int = 1, xengine: str = "fp32",
                    lsb: bool = False, fringe_rot_order: int = 1, chunk_ffts: int = 0,
                    write_cross_autocorrs: bool = True, max_delay_us: float = 1e4,
                    max_rate_us_per_s: float = 1.0, model_nsamples: int = 1,
                    start_offset_s: int = 0, **extra) -> Config:
    """One datastream per station, all npol x len(freqs) bands in one sample word
    (band order: freq-major, pol-minor, as vex2difx records them), and every baseline
    correlating every freq with all npol**2 products (parallel hands first)."""
    rng = np.random.default_rng(seed)
    pols = "RL"[:npol]
    freqs = [Freq(f, bw_mhz, nchan, lsb=lsb) for f in freqs_mhz]
    nf = len(freqs)
    dss = []
    for s in range(nstations):
        model = random_delay_model(rng, model_nsamples, max_delay_us, max_rate_us_per_s)
        recbands = [(fi, p) for fi in range(nf) for p in pols]
        dss.append(Datastream(recfreqs=list(range(nf)), recbands=recbands, model=model,
                              nbits=nbits, complex_sampling=complex_sampling,
                              payloadbytes=payloadbytes))
    bls = []
    for a in range(nstations):
        for b in range(a + 1, nstations):
            fl = []
            for fi in range(nf):
                idx = {p: fi * npol + k for k, p in enumerate(pols)}
                prods = [(idx[p], idx[p]) for p in pols]
                if npol == 2:
                    prods += [(idx["R"], idx["L"]), (idx["L"], idx["R"])]
                fl.append((fi, prods))
            bls.append(Baseline(a, b, fl))
    return _with(Config(freqs=freqs, datastreams=dss, baselines=bls, subint_ns=subint_ns,
                        fringe_rot_order=fringe_rot_order, xengine=xengine, chunk_ffts=chunk_ffts,
                        write_cross_autocorrs=write_cross_autocorrs), extra)


def batched_config(nstations: int, groups: Sequence[Tuple[float, int, Sequence[float]]],
                   zooms: Sequence[Tuple[int, int, float, float]], subint_ns: int, npol: int = 2,
                   nbits: int = 2, payloadbytes: int = 8000, seed: int = 1, xengine: str = "fp32",
                   chunk_ffts: int = 0, write_cross_autocorrs: bool = True, max_delay_us: float = 1e4,
                   max_rate_us_per_s: float = 1.0, model_nsamples: int = 1,
                   correlate_parents: bool = True, **extra) -> Config:
    """Uneven bands: one datastream per (station, bandwidth group), since a DiFX datastream
    has one recorded bandwidth and channel count (mode.h:71, configuration.cpp:937-941).

    groups: (bandwidth MHz, nchan, band-edge sky frequencies) per group; every group must
            give the same FFT duration nchan / bandwidth so that all datastreams run the same
            blockspersend (configuration.cpp:1843-1852, 3008-3028).
    zooms:  (group, local parent freq, offset of the zoom's lower edge in MHz, zoom bandwidth
            MHz); a zoom band is a window of its parent's channels at the parent resolution
            (configuration.cpp:1733, mode.cpp:184-196), recorded in every polarisation.

    Datastreams are ordered station-major.  Baselines pair the datastreams of one group on two
    stations and correlate its parent freqs (unless correlate_parents is False) and its zoom
    freqs, each with npol**2 products (parallel hands first)."""
    rng = np.random.default_rng(seed)
    pols = "RL"[:npol]
    freqs: List[Freq] = []
    gfreqs: List[List[int]] = []
    for bw, nchan, edges in groups:
        gfreqs.append([])
        for e in edges:
            gfreqs[-1].append(len(freqs))
            freqs.append(Freq(float(e), float(bw), int(nchan)))
    gz: List[List[Tuple[int, int, int]]] = [[] for _ in groups]   # (zoom freq idx, parent lf, choff)
    for g, lf, off, zbw in zooms:
        bw, nchan, edges = groups[g]
        chw = bw / nchan
        zn = int(round(zbw / chw))
        choff = int(off / chw + 0.5)
        if abs(zn * chw - zbw) > TINY or choff + zn > nchan:
            raise ValueError("zoom band is not a whole channel window of its parent")
        gz[g].append((len(freqs), lf, choff))
        freqs.append(Freq(float(edges[lf]) + choff * chw, float(zbw), zn))
    dss: List[Datastream] = []
    for s in range(nstations):
        model = random_delay_model(rng, model_nsamples, max_delay_us, max_rate_us_per_s)
        for g in range(len(groups)):
            nf = len(gfreqs[g])
            recbands = [(lf, p) for lf in range(nf) for p in pols]
            zb = [(lz, p) for lz in range(len(gz[g])) for p in pols]
            dss.append(Datastream(recfreqs=list(gfreqs[g]), recbands=recbands, model=model.copy(),
                                  nbits=nbits, payloadbytes=payloadbytes, zoomfreqs=list(gz[g]),
                                  zoombands=zb))
    ng = len(groups)
    bls: List[Baseline] = []
    for a in range(nstations):
        for b in range(a + 1, nstations):
            for g in range(ng):
                nrec = len(gfreqs[g]) * npol
                fl = []

                def prods(first):
                    idx = {p: first + k for k, p in enumerate(pols)}
                    pr = [(idx[p], idx[p]) for p in pols]
                    if npol == 2:
                        pr += [(idx["R"], idx["L"]), (idx["L"], idx["R"])]
                    return pr

                if correlate_parents:
                    for lf, fi in enumerate(gfreqs[g]):
                        fl.append((fi, prods(lf * npol)))
                for lz, (zfi, _, _) in enumerate(gz[g]):
                    fl.append((zfi, prods(nrec + lz * npol)))
                bls.append(Baseline(a * ng + g, b * ng + g, fl))
    return _with(Config(freqs=freqs, datastreams=dss, baselines=bls, subint_ns=subint_ns, xengine=xengine,
                        chunk_ffts=chunk_ffts, write_cross_autocorrs=write_cross_autocorrs), extra)
