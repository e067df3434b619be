"""CPU restatement of the VDIF thread corner turn (test infrastructure only).

This module is a CHECKER, not the product: only tests/ and the cpu_baseline leg of bench.py
may import it.  It restates what VDIFMuxer (mpifxcorr/src/datamuxer.cpp) does to one segment
of multi-thread VDIF: single-channel frames of N threads, in any order, become N-channel
frames whose 32-bit words interleave the threads sample by sample, the format Mk5Mode
unpacks.

  deinterlace   VDIFMuxer::deinterlace (datamuxer.cpp:764-833): a frame's slot is its frame
                count since the reference frame, (mjd - refmjd)*86400 + sec - refsec)*fps +
                frame - refframe, minus the segment start; its thread index comes from the
                thread map; frames before the segment are dropped; a later copy of a slot
                replaces an earlier one
  multiplex     VDIFMuxer::multiplex (datamuxer.cpp:670-757): per slot, when every thread is
                present, thread 0's header with invalid 0, log2(N) channels, the output frame
                length and thread id 0, then the corner-turned payload; when some are present,
                the first present thread's header with invalid 1, the output frame length and
                thread id 0 (payload not written); when none are, the previous output
                header with invalid 1 and the frame number advanced by one (wrapping into the
                next second at fps); leading slots with no thread present are not output
  corner turn   VDIFMuxer::cornerturn_generic (datamuxer.cpp:243-266): output word i*N + j
                holds samples j*spw + k (k < spw = 32/(bits*N)) of every thread l at sample
                position k*N + l
  header        vdif_header bit fields (vdifio.h:68-92), getVDIFEpochMJD / ymd2mjd
                (vdifio.c:127-131, dateutils.c:35-63), setVDIFNumChannels (vdifio.c:134-144)

Pins (tests/test_vdifmux_cpu.py): the corner turn against the reference's own mask-and-shift
turners for 2, 4, 8 and 16 threads of 2-bit data (datamuxer.cpp:279-532); a mux of N
single-band streams decodes (fxoracle.decode_stream, the unpack pins) to the same samples as
the N-band stream they were split from; hand-built segments for every header rule.
"""
from __future__ import annotations

from typing import Dict, List, Sequence, Tuple

import numpy as np

HEADER_BYTES = 32                                   # VDIF_HEADER_BYTES


# ----------------------------------------------------------------------------- header fields
def header_words(frame: np.ndarray) -> np.ndarray:
    return np.frombuffer(frame[:HEADER_BYTES].tobytes(), dtype="<u4").copy()


def fields(w: np.ndarray) -> Dict[str, int]:
    """vdif_header (vdifio.h:68-92), little-endian bit fields."""
    w = [int(x) for x in w]
    return {"seconds": w[0] & 0x3FFFFFFF, "legacy": (w[0] >> 30) & 1, "invalid": (w[0] >> 31) & 1,
            "frame": w[1] & 0xFFFFFF, "epoch": (w[1] >> 24) & 0x3F,
            "framebytes": (w[2] & 0xFFFFFF) * 8, "lognchan": (w[2] >> 24) & 0x1F,
            "threadid": (w[3] >> 16) & 0x3FF}


def ymd2doy(yr: int, mo: int, day: int) -> int:
    """dateutils.c:35-53 (integer arithmetic of C: truncating division)."""
    monstart1 = [0, 31, 59, 90, 120, 151, 181, 212, 243, 273, 304, 334]
    monstart2 = [0, 31, 60, 91, 121, 152, 182, 213, 244, 274, 305, 335]
    L2 = yr // 4 - (yr + 7) // 4 - yr // 100 + (yr + 99) // 100 + yr // 400 - (yr + 399) // 400
    return day + (monstart2[mo - 1] if L2 == -1 else monstart1[mo - 1])


def ymd2mjd(yr: int, mo: int, day: int) -> int:
    """dateutils.c:55-63."""
    yr1 = yr - 1
    return ymd2doy(yr, mo, day) - 678576 + 365 * yr1 + yr1 // 4 - yr1 // 100 + yr1 // 400


def epoch_mjd(epoch: int) -> int:
    """getVDIFEpochMJD (vdifio.c:127-131): epochs are half years from 2000."""
    return ymd2mjd(2000 + epoch // 2, (epoch % 2) * 6 + 1, 1)


def full_seconds(w: np.ndarray) -> int:
    """getVDIFFrameMJD * 86400 + getVDIFFrameSecond (vdifio.c:158-163, vdifio.h:327): the
    seconds since MJD 0 of the frame."""
    f = fields(w)
    return epoch_mjd(f["epoch"]) * 86400 + f["seconds"]


def set_invalid(w: np.ndarray, v: int) -> None:
    w[0] = (int(w[0]) & 0x7FFFFFFF) | ((v & 1) << 31)


def set_framebytes(w: np.ndarray, nbytes: int) -> None:
    w[2] = (int(w[2]) & 0xFF000000) | ((nbytes // 8) & 0xFFFFFF)


def set_nchan(w: np.ndarray, n: int) -> None:
    """setVDIFNumChannels (vdifio.c:134-144): floor(log2(n))."""
    lg = 0
    while n > 1:
        n //= 2
        lg += 1
    w[2] = (int(w[2]) & ~(0x1F << 24) & 0xFFFFFFFF) | ((lg & 0x1F) << 24)


def set_threadid(w: np.ndarray, t: int) -> None:
    w[3] = (int(w[3]) & ~(0x3FF << 16) & 0xFFFFFFFF) | ((t & 0x3FF) << 16)


# ----------------------------------------------------------------------------- corner turns
def cornerturn_generic(tw: np.ndarray, bits: int) -> np.ndarray:
    """datamuxer.cpp:243-266.  tw: [N, W] uint32 payload words of the N threads; returns the
    N*W output words."""
    N, W = tw.shape
    spw = (32 // bits) // N                          # samplesperoutputword (datamuxer.cpp:154-155)
    mask = (1 << bits) - 1
    out = np.zeros(W * N, dtype=np.uint64)
    t64 = tw.astype(np.uint64)
    for j in range(N):
        acc = np.zeros(W, dtype=np.uint64)
        for k in range(spw):
            for l in range(N):
                acc |= ((t64[l] >> np.uint64((j * spw + k) * bits)) & np.uint64(mask)) << np.uint64((k * N + l) * bits)
        out[np.arange(W) * N + j] = acc
    return out.astype(np.uint32)


def _mask_shift(x: np.ndarray, masks, shifts) -> np.ndarray:
    y = np.zeros_like(x)
    for m, s in zip(masks, shifts):
        y |= ((x & np.uint32(m)) >> np.uint32(-s)) if s < 0 else ((x & np.uint32(m)) << np.uint32(s))
    return y


def cornerturn_2bit_special(tb: np.ndarray) -> np.ndarray:
    """The reference's mask-and-shift turners for 2, 4, 8 and 16 threads of 2-bit data
    (datamuxer.cpp:279-532), restated from their masks and byte shuffles.  tb: [N, B] payload
    bytes; returns the output payload bytes."""
    N, B = tb.shape
    t = tb.astype(np.uint32)
    if N == 2:                                       # :279-324
        x = (t[1, 1::2] << 24) | (t[0, 1::2] << 16) | (t[1, 0::2] << 8) | t[0, 0::2]
        y = _mask_shift(x, [0xC003C003, 0x30003000, 0x000C000C, 0x0C000C00, 0x00300030, 0x03000300, 0x00C000C0],
                        [0, -2, 2, -4, 4, -6, 6])
        return y.astype("<u4").view(np.uint8)
    m4 = [0xC0300C03, 0x300C0300, 0x00C0300C, 0x0C030000, 0x0000C030, 0x03000000, 0x000000C0]
    s4 = [0, -6, 6, -12, 12, -18, 18]
    if N == 4:                                       # :326-384
        x = (t[3] << 24) | (t[2] << 16) | (t[1] << 8) | t[0]
        return _mask_shift(x, m4, s4).astype("<u4").view(np.uint8)
    if N == 8:                                       # :386-448
        y1 = _mask_shift((t[3] << 24) | (t[2] << 16) | (t[1] << 8) | t[0], m4, s4)
        y2 = _mask_shift((t[7] << 24) | (t[6] << 16) | (t[5] << 8) | t[4], m4, s4)
        b1 = [(y1 >> np.uint32(8 * i)) & np.uint32(0xFF) for i in range(4)]
        b2 = [(y2 >> np.uint32(8 * i)) & np.uint32(0xFF) for i in range(4)]
        o = np.empty(2 * B, dtype=np.uint32)
        o[0::2] = (b2[1] << 24) | (b1[1] << 16) | (b2[0] << 8) | b1[0]
        o[1::2] = (b2[3] << 24) | (b1[3] << 16) | (b2[2] << 8) | b1[2]
        return o.astype("<u4").view(np.uint8)
    if N == 16:                                      # :450-532
        ys = [_mask_shift((t[4 * g + 3] << 24) | (t[4 * g + 2] << 16) | (t[4 * g + 1] << 8) | t[4 * g], m4, s4)
              for g in range(4)]
        bs = [[(y >> np.uint32(8 * i)) & np.uint32(0xFF) for i in range(4)] for y in ys]
        o = np.empty(4 * B, dtype=np.uint32)
        for i in range(4):
            o[i::4] = (bs[3][i] << 24) | (bs[2][i] << 16) | (bs[1][i] << 8) | bs[0][i]
        return o.astype("<u4").view(np.uint8)
    raise ValueError("special turners exist for 2, 4, 8 and 16 threads")


# ----------------------------------------------------------------------------- the segment
def mux_segment(buf: np.ndarray, framebytes: int, threadmap: Sequence[int], fps: int, bits: int,
                ref_seconds: int, ref_frame: int, start: int, nslots: int, out: np.ndarray | None = None
                ) -> Tuple[np.ndarray, int, Dict[str, int]]:
    """Corner-turn one segment: `buf` holds input frames of `framebytes` (any order, any of the
    mapped threads); slot s is the frame at count start + s after the reference frame
    (ref_seconds = its MJD seconds, ref_frame its frame number).  Returns the output frames
    (nslots of (framebytes - 32) * N + 32 bytes, written from the start; the unwritten rest of
    `out`, or zeros, is left as it was), the number of output frames, and counters (dropped
    frames: before the segment, after it, of an unmapped thread)."""
    N = len(threadmap)
    pay = framebytes - HEADER_BYTES
    ofb = pay * N + HEADER_BYTES
    if out is None:
        out = np.zeros(nslots * ofb, dtype=np.uint8)
    nfr = len(buf) // framebytes
    table = -np.ones((nslots, N), dtype=np.int64)
    stats = {"early": 0, "late": 0, "unmapped": 0}
    for i in range(nfr):                             # deinterlace (datamuxer.cpp:773-829)
        fr = buf[i * framebytes:(i + 1) * framebytes]
        w = header_words(fr)
        f = fields(w)
        if f["framebytes"] != framebytes:
            raise ValueError("frame length changed")  # the reference aborts (:784-787)
        if f["threadid"] not in threadmap:
            stats["unmapped"] += 1
            continue
        ti = list(threadmap).index(f["threadid"])
        cur = (full_seconds(w) - ref_seconds) * fps + f["frame"] - ref_frame - start
        if cur < 0:
            stats["early"] += 1
            continue
        if cur >= nslots:
            stats["late"] += 1
            continue
        table[cur, ti] = i                           # a later copy replaces an earlier one (:827)
    spw = (32 // bits) // N
    assert spw >= 1, "too many threads for the sample width"
    nout = 0
    prev = None
    for s in range(nslots):                          # multiplex (datamuxer.cpp:676-752)
        present = table[s] >= 0
        o = out[nout * ofb:(nout + 1) * ofb]
        if present.all():
            w = header_words(buf[table[s, 0] * framebytes:])
            set_invalid(w, 0)
            set_nchan(w, N)
            set_framebytes(w, ofb)
            set_threadid(w, 0)
            tw = np.stack([np.frombuffer(buf[table[s, l] * framebytes + HEADER_BYTES:
                                             (table[s, l] + 1) * framebytes].tobytes(), dtype="<u4")
                           for l in range(N)])
            o[HEADER_BYTES:] = cornerturn_generic(tw, bits).astype("<u4").view(np.uint8)
        elif present.any():
            l = int(np.flatnonzero(present)[0])
            w = header_words(buf[table[s, l] * framebytes:])
            set_invalid(w, 1)
            set_framebytes(w, ofb)
            set_threadid(w, 0)
        elif nout > 0:
            w = prev.copy()
            set_invalid(w, 1)
            fn = (int(w[1]) & 0xFFFFFF) + 1
            w[1] = (int(w[1]) & 0xFF000000) | (fn & 0xFFFFFF)
            if (fn & 0xFFFFFF) == fps:
                w[1] = int(w[1]) & 0xFF000000
                secs = ((int(w[0]) & 0x3FFFFFFF) + 1) & 0x3FFFFFFF   # setVDIFFrameSecond(+1)
                w[0] = (int(w[0]) & 0xC0000000) | secs
        else:
            continue                                 # nothing to copy: the slot is not output
        o[:HEADER_BYTES] = w.astype("<u4").view(np.uint8)
        prev = w
        nout += 1
    return out, nout, stats


# ----------------------------------------------------------------------------- test helpers
def make_header(seconds: int, frame: int, epoch: int, framebytes: int, threadid: int, bits: int,
                lognchan: int = 0, stationid: int = 0x4142, invalid: int = 0) -> np.ndarray:
    """A VDIF header (vdifio.h:68-92) as 8 little-endian words."""
    w = np.zeros(8, dtype=np.uint32)
    w[0] = (seconds & 0x3FFFFFFF) | ((invalid & 1) << 31)
    w[1] = (frame & 0xFFFFFF) | ((epoch & 0x3F) << 24)
    w[2] = ((framebytes // 8) & 0xFFFFFF) | ((lognchan & 0x1F) << 24)
    w[3] = (stationid & 0xFFFF) | ((threadid & 0x3FF) << 16) | (((bits - 1) & 0x1F) << 26)
    return w


def split_threads(buf: np.ndarray, framebytes: int, nthreads: int, bits: int, threadids: Sequence[int]
                  ) -> List[np.ndarray]:
    """The inverse of the corner turn: the N single-channel thread streams a stream of
    N-channel frames was made from (each input frame gives one frame per thread, same header
    with 1 channel, the thread's id and the thread frame length).  Used to build test inputs."""
    pay = framebytes - HEADER_BYTES
    tpay = pay // nthreads
    tfb = tpay + HEADER_BYTES
    spw = (32 // bits) // nthreads
    mask = (1 << bits) - 1
    nfr = len(buf) // framebytes
    outs = [np.zeros(nfr * tfb, dtype=np.uint8) for _ in range(nthreads)]
    for f in range(nfr):
        fr = buf[f * framebytes:(f + 1) * framebytes]
        w = header_words(fr)
        ow = np.frombuffer(fr[HEADER_BYTES:].tobytes(), dtype="<u4").astype(np.uint64)
        W = len(ow) // nthreads                      # input words per thread frame
        for l in range(nthreads):
            th = np.zeros(W, dtype=np.uint64)
            for j in range(nthreads):
                src = ow[np.arange(W) * nthreads + j]
                for k in range(spw):
                    th |= ((src >> np.uint64((k * nthreads + l) * bits)) & np.uint64(mask)) << np.uint64((j * spw + k) * bits)
            hw = w.copy()
            set_nchan(hw, 1)
            set_framebytes(hw, tfb)
            set_threadid(hw, threadids[l])
            o = outs[l][f * tfb:(f + 1) * tfb]
            o[:HEADER_BYTES] = hw.astype("<u4").view(np.uint8)
            o[HEADER_BYTES:] = th.astype(np.uint32).astype("<u4").view(np.uint8)
    return outs
