"""Seeded synthetic station data and slot inputs for the FX path.

The recorded-data side that the reference gets from DataStream (datastream.cpp:246-627) and
the vdifio/mark5access writers: 2-bit quantisation with 4-level thresholds (vdifsim's
+-0.96 sigma is not used; this module uses generateVDIF's +-0.95 sigma,
libraries/vdifio/utils/generateVDIF.c:789-790), VDIF frames with 32-byte headers
(vdifio.h) or Mark5B frames with 16-byte headers, and the control words
{scan, sec, ns, validflags...} that DataStream::calculateControlParams produces
(datastream.cpp:380, 581-605).  The data start chosen here is frame aligned and one frame
ahead of the earliest sample the subint needs, so the reference and this build read the same
buffer with identical control words.
"""
from __future__ import annotations

import math
from typing import List, Optional, Sequence, Tuple

import numpy as np

THRESH = 0.95            # 2-bit threshold in units of sigma (generateVDIF.c:789-790)
FILL_WORD64 = 0x1122334411223344


def samples_per_frame(ds) -> int:
    pay = ds.framebytes - ds.headerbytes
    bits = ds.nbits * (2 if ds.complex_sampling else 1) * len(ds.recbands)
    return pay * 8 // bits


def quantize_2bit(x: np.ndarray, fmt: str = "VDIF", thresh: float = THRESH) -> np.ndarray:
    """4-level codes of x (unit sigma): VDIF code = {0:-H, 1:-1, 2:+1, 3:+H}
    (format_vdif.c:70), Mark5B field = sign + 2*mag with {0:-H, 1:+1, 2:-1, 3:+H}
    (mark5_format_mark5b.c:52)."""
    lvl = np.digitize(x, [-thresh, 0.0, thresh]).astype(np.uint8)   # 0..3 from low to high
    if fmt.upper() == "MARK5B":
        return np.array([0, 2, 1, 3], dtype=np.uint8)[lvl]
    return lvl


def random_2bit_codes(rng: np.random.Generator, n: int, fmt: str = "VDIF") -> np.ndarray:
    """Codes with the probabilities of a unit Gaussian at +-THRESH, drawn without
    generating Gaussians (fast path for large buffers)."""
    p_hi = 0.5 * math.erfc(THRESH / math.sqrt(2.0))
    u = rng.random(n, dtype=np.float32)
    lvl = np.digitize(u, [p_hi, 0.5, 1.0 - p_hi]).astype(np.uint8)
    if fmt.upper() == "MARK5B":
        return np.array([0, 2, 1, 3], dtype=np.uint8)[lvl]
    return lvl


def pack_2bit(codes: np.ndarray) -> np.ndarray:
    """codes[nsamples, nbands] -> payload bytes, field q = n*nbands + b at byte q//4,
    bit 2*(q%4) (LSB first)."""
    q = codes.reshape(-1)
    pad = (-len(q)) % 4
    if pad:
        q = np.concatenate([q, np.zeros(pad, np.uint8)])
    q = q.reshape(-1, 4).astype(np.uint8)
    return (q[:, 0] | (q[:, 1] << 2) | (q[:, 2] << 4) | (q[:, 3] << 6)).astype(np.uint8)


def quantize_8bit(x: np.ndarray, sigma_counts: float = 20.0) -> np.ndarray:
    """8-bit offset-binary samples with the reference's 128 offset (format_vdif.c:120)."""
    return np.clip(np.round(x * sigma_counts + 128.0), 0, 255).astype(np.uint8)


def vdif_header(sec: int, frame_in_sec: int, framebytes: int, nchan: int, nbits: int,
                is_complex: bool, invalid: bool = False, station: int = 0x4142) -> np.ndarray:
    w = np.zeros(8, dtype=np.uint32)
    w[0] = (sec & 0x3FFFFFFF) | ((1 << 31) if invalid else 0)
    w[1] = frame_in_sec & 0xFFFFFF
    w[2] = ((framebytes // 8) & 0xFFFFFF) | ((int(round(math.log2(nchan))) & 0x1F) << 24)
    w[3] = (station & 0xFFFF) | (((nbits - 1) & 0x1F) << 26) | ((1 << 31) if is_complex else 0)
    return w.view(np.uint8)


def mark5b_header(frame_in_sec: int, invalid: bool = False) -> np.ndarray:
    w = np.zeros(4, dtype=np.uint32)
    w[0] = 0xABADDEED
    w[1] = (frame_in_sec & 0x7FFF) | ((1 << 15) if invalid else 0)
    return w.view(np.uint8)


def frame_payloads(payload: np.ndarray, ds, start_frame: int, sec0: int = 0) -> np.ndarray:
    """Cut a payload byte stream into frames and prepend headers."""
    pay = ds.framebytes - ds.headerbytes
    nframes = len(payload) // pay
    out = np.zeros((nframes, ds.framebytes), dtype=np.uint8)
    out[:, ds.headerbytes:] = payload[: nframes * pay].reshape(nframes, pay)
    for f in range(nframes):
        if ds.format.upper() == "MARK5B":
            out[f, :16] = mark5b_header(start_frame + f)
        else:
            out[f, :32] = vdif_header(sec0, start_frame + f, ds.framebytes,
                                      max(1, len(ds.recbands)), ds.nbits, ds.complex_sampling)
    return out.reshape(-1)


def data_bytes(cfg, d: int) -> int:
    """Configuration::getDataBytes (configuration.cpp:794-820)."""
    ds = cfg.datastreams[d]
    f = cfg.freqs[ds.recfreqs[0]]
    B = cfg.blockspersend()
    nrec = len(ds.recbands)
    valid = (B * nrec * 2 * ds.nbits * f.nchan) // 8
    valid += (int(cfg.guard_ns / (1000.0 / (f.bw_mhz * 2.0)) + 1.0) * nrec * ds.nbits) // 8
    pay = ds.framebytes - ds.headerbytes
    nframes = valid // pay + 2
    if ds.format.upper() == "MARK5B":
        nframes += nframes % 2
    return nframes * ds.framebytes


def eval_delay(model: np.ndarray, order: int, t: float, inc: float) -> float:
    s = min(int(t / inc), model.shape[0] - 1)
    dt = t - s * inc
    return float(sum(model[s, k] * dt ** k for k in range(order + 1)))


def buffer_start_ns(cfg, d: int, offsets: Sequence[int]) -> int:
    """Frame-aligned station time (ns from scan start) of the slot buffer: one frame before
    the first sample of the subint's first FFT (delay at the subint start)."""
    ds = cfg.datastreams[d]
    spf = samples_per_frame(ds)
    dt_ns = cfg.sampletime_us(d) * 1000.0
    frame_ns = spf * dt_ns
    if abs(frame_ns - round(frame_ns)) > 1e-6:
        raise ValueError("frame duration is not an integer number of ns")
    frame_ns = int(round(frame_ns))
    t0 = offsets[1] * 1e9 + offsets[2]
    tau_ns = eval_delay(ds.model, ds.model_order, t0 / 1e9, cfg.model_inc_secs) * 1000.0
    tau_ns -= 1e9 * math.floor(ds.clock_coeff0_us / 1e6 + 0.5)
    start = t0 - tau_ns
    return int(math.floor(start / frame_ns) - 1) * frame_ns


def slot_bytes(cfg, d: int, offsets: Sequence[int], start_ns: int) -> int:
    """Bytes of a slot buffer starting at start_ns: Configuration::getDataBytes, extended
    when a short subint needs more frames to cover its last FFT after the start offset."""
    ds = cfg.datastreams[d]
    spf = samples_per_frame(ds)
    dt_ns = cfg.sampletime_us(d) * 1000.0
    frame_ns = spf * dt_ns
    t_end = offsets[1] * 1e9 + offsets[2] + cfg.subint_ns
    tau_end = eval_delay(ds.model, ds.model_order, t_end / 1e9, cfg.model_inc_secs) * 1000.0
    tau_end -= 1e9 * math.floor(ds.clock_coeff0_us / 1e6 + 0.5)
    # the last FFT ends at the subint end minus the delay there; one frame of margin
    need_ns = t_end - tau_end - start_ns
    need_frames = int(math.ceil(need_ns / frame_ns)) + 1
    return max(data_bytes(cfg, d), need_frames * ds.framebytes)


def control_words(cfg, d: int, offsets, start_ns: int, B: Optional[int] = None,
                  invalid_ffts: Sequence[int] = ()) -> np.ndarray:
    """{scan, sec, ns, validflags...} with 30 valid bits per int (mpifxcorr.h:29)."""
    if B is None:
        B = cfg.blockspersend()
    nflag = (B + 29) // 30
    c = np.zeros(3 + nflag, dtype=np.int32)
    sec = int(start_ns) // 1000000000
    c[0] = offsets[0]
    c[1] = sec
    c[2] = int(start_ns) - sec * 1000000000
    flags = np.zeros(nflag * 30, dtype=np.int64)
    flags[:B] = 1
    for i in invalid_ffts:
        flags[i] = 0
    c[3:] = (flags.reshape(nflag, 30) << np.arange(30, dtype=np.int64)).sum(axis=1).astype(np.int32)
    return c


def noise_slot(cfg, offsets, seed: int = 0, nbytes_extra_frames: int = 0):
    """Independent 2-bit (or 8-bit complex) noise per station.  Returns (data, control)
    lists of numpy arrays."""
    rng = np.random.default_rng(seed)
    data, control = [], []
    for d, ds in enumerate(cfg.datastreams):
        start_ns = buffer_start_ns(cfg, d, offsets)
        nbytes = slot_bytes(cfg, d, offsets, start_ns) + nbytes_extra_frames * ds.framebytes
        nframes = nbytes // ds.framebytes
        pay = ds.framebytes - ds.headerbytes
        nb = len(ds.recbands)
        if ds.nbits == 2 and not ds.complex_sampling:
            codes = random_2bit_codes(rng, nframes * pay * 4, ds.format)
            payload = pack_2bit(codes.reshape(-1, 1))
        elif ds.nbits == 8:
            payload = quantize_8bit(rng.standard_normal(nframes * pay).astype(np.float32))
        else:
            raise NotImplementedError
        spf = samples_per_frame(ds)
        frame_ns = int(round(spf * cfg.sampletime_us(d) * 1000.0))
        first = start_ns // frame_ns
        data.append(frame_payloads(payload, ds, int(first)))
        control.append(control_words(cfg, d, offsets, start_ns))
    return data, control


def correlated_slot(cfg, offsets, seed: int = 0, amp: float = 0.05, lo_phase: bool = True):
    """Gaussian noise plus a common band-limited signal of relative amplitude `amp`, delayed
    for each station by its (constant) model delay tau_k: the analytic common signal s_a is
    read at T + tau_k and carries the LO phase exp(2 pi i f_LO tau_k), so that after the
    correlator removes the model delay and fringe phase the cross spectra are flat in phase.
    Real USB sampling only; one recorded freq per station band set; model delays must be
    constant over the subint (only the c0 term)."""
    rng = np.random.default_rng(seed)
    ds0 = cfg.datastreams[0]
    dt_us = cfg.sampletime_us(0)
    starts = [buffer_start_ns(cfg, d, offsets) for d in range(len(cfg.datastreams))]
    lens = [slot_bytes(cfg, d, offsets, starts[d]) for d in range(len(cfg.datastreams))]
    nsamps = [(lens[d] // ds.framebytes) * samples_per_frame(ds)
              for d, ds in enumerate(cfg.datastreams)]
    taus = [float(ds.model[0, 0]) for ds in cfg.datastreams]
    t_lo = min(s / 1000.0 + t for s, t in zip(starts, taus)) - 10 * dt_us   # us
    span = max(s / 1000.0 + t + n * dt_us for s, t, n in zip(starts, taus, nsamps)) - t_lo
    N = 1 << int(math.ceil(math.log2(span / dt_us + 64)))
    nbands_freq = {}
    out_data, out_ctrl = [], []
    # one common analytic signal per recorded freq index (positive frequencies only)
    fidx = sorted({fi for ds in cfg.datastreams for fi in ds.recfreqs})
    common = {}
    for fi in fidx:
        S = np.zeros(N, dtype=np.complex128)
        S[1: N // 2] = (rng.standard_normal(N // 2 - 1) + 1j * rng.standard_normal(N // 2 - 1))
        common[fi] = S * math.sqrt(2.0 * N)  # Re(ifft) then has unit variance
    fgrid = np.fft.fftfreq(N, d=dt_us)        # MHz
    for d, ds in enumerate(cfg.datastreams):
        n = nsamps[d]
        nb = len(ds.recbands)
        vals = np.zeros((n, nb))
        shift_us = starts[d] / 1000.0 + taus[d] - t_lo
        for b, (lf, pol) in enumerate(ds.recbands):
            fi = ds.recfreqs[lf]
            f_lo = cfg.freqs[fi].freq_mhz
            ph = np.exp(2j * np.pi * fgrid * shift_us)
            if lo_phase:
                ph = ph * np.exp(2j * np.pi * f_lo * taus[d])
            s = np.fft.ifft(common[fi] * ph).real[:n]
            vals[:, b] = rng.standard_normal(n) + amp * s
        sd = np.sqrt(1.0 + amp * amp)
        codes = quantize_2bit(vals / sd, ds.format)
        payload = pack_2bit(codes)
        spf = samples_per_frame(ds)
        frame_ns = int(round(spf * dt_us * 1000.0))
        out_data.append(frame_payloads(payload, ds, starts[d] // frame_ns))
        out_ctrl.append(control_words(cfg, d, offsets, starts[d]))
    return out_data, out_ctrl
