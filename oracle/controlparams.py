"""CPU restatement of DataStream::calculateControlParams (test infrastructure only).

This module is a CHECKER, not the product: only tests/ may import it.  It restates, in
Python with C's integer semantics (truncating division and casts, remainders that take the
sign of the dividend), how mpifxcorr's DataStream chooses where a subint's data starts in
its segmented read buffer and which FFT blocks have data
(mpifxcorr/src/datastream.cpp:368-627), and the per-segment constants of
DataStream::updateConfig (datastream.cpp:745-763).

  delay span      datastream.cpp:380-409: first sample ns = offsetns - int(delay1 us * 1000),
                  last = offsetns + data span + int(-1000 delay2 + 0.5), one second borrowed
                  (at most twice) while last < 0
  segment move    :413-457: invalid when the scan is earlier, the segment unread or the start
                  more than 3 s before the segment; move on while the start is at or past the
                  segment's valid nanoseconds
  far checks      :460-508
  byte index      :515-524 (16-bit aligned), blocks before the segment skipped with the
                  delay change over them (:541-568), back to a whole-nanosecond byte (:570-582)
  valid flags     :584-620, 30 per control word (mpifxcorr.h:29)
  framed streams  VDIFDataStream / Mark5BDataStream::calculateControlParams
                  (vdiffile.cpp:348-472, mark5bfile.cpp:121-247): back to a frame boundary
                  that is a multiple of the frame granularity, the frame's time, the frame's
                  byte; their updateConfig's frame-based nsinc and sendbytes

Pins (tests/test_control_cpu.py): hand-derived cases whose byte index, start time and flags
follow from the buffer geometry alone (aligned start, a delay of whole samples, a partly
valid segment, a move to the next segment, each invalid exit, the partial-word quirk of the
good-data check), and the C function of the product equal to this one on random buffer
states.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Callable, List, Optional, Tuple

import numpy as np

INVALID_SUBINT = -99999999                  # mode.h:222
FLAGS_PER_INT = 30                          # mpifxcorr.h:29
NS = 1000000000


class StateError(RuntimeError):
    pass


def cdiv(a: int, b: int) -> int:
    """C integer division (truncates toward zero)."""
    q = abs(a) // abs(b)
    return q if (a >= 0) == (b >= 0) else -q


def cmod(a: int, b: int) -> int:
    """C remainder (sign of the dividend)."""
    return a - b * cdiv(a, b)


def cint(x: float) -> int:
    """C cast of a double to an integer (truncation)."""
    return int(x)


@dataclass
class Segment:
    """DataStream::readinfo (datastream.h:87-113), the fields the choice reads."""
    scan: int = 0
    scanseconds: int = 0
    scanns: int = 0
    validbytes: int = 0
    readto: bool = True
    sendbytes: int = 0
    blockspersend: int = 1
    bytespersamplenum: int = 1
    bytespersampledenom: int = 1
    bytesbetweenintegerns: int = 1
    nsinc: int = 0
    numchannels: int = 1
    complex_sampling: bool = False
    sampletimens: float = 1.0


@dataclass
class Buffer:
    segs: List[Segment]
    bufferbytes: int
    readbytes: int
    keepreading: bool = True
    lastvalidsegment: int = 0
    intclockseconds: int = 0
    framing: int = 0                        # 0 none, 1 VDIF, 2 Mark5B
    framebytes: int = 0
    payloadbytes: int = 0
    framespersecond: int = 0
    framegranularity: int = 1


def segment_constants(nbits: int, nrecbands: int, complex_sampling: bool, bw_mhz: float, nchan: int,
                      blockspersend: int, databytes: int, subint_ns: int, guard_ns: int,
                      bufferbytes: int, numdatasegments: int) -> Segment:
    """DataStream::updateConfig (datastream.cpp:745-763) with Configuration's bytes per
    sample fraction (configuration.cpp:1656-1664, decimation 1)."""
    g = Segment()
    g.sendbytes = cint((databytes * (subint_ns + guard_ns)) // subint_ns)
    g.blockspersend = blockspersend
    sb = nbits * (2 if complex_sampling else 1)
    g.bytespersamplenum = (nrecbands * sb) // 8
    if g.bytespersamplenum == 0:
        g.bytespersamplenum = 1
        g.bytespersampledenom = 8 // (nrecbands * sb)
    else:
        g.bytespersampledenom = 1
    g.numchannels = nchan
    g.complex_sampling = complex_sampling
    g.sampletimens = 500.0 / bw_mhz * (2 if complex_sampling else 1)
    tiny50 = float(np.float32(50) * np.float32(1e-9))      # 50 * Mode::TINY in float (mode.cpp:27)
    acc, g.bytesbetweenintegerns = 0.0, 0
    while True:
        acc += g.bytespersampledenom * g.sampletimens
        g.bytesbetweenintegerns += g.bytespersamplenum
        if abs(acc - cint(acc + 0.5)) < tiny50:
            break
    g.nsinc = cint((g.sampletimens * (bufferbytes // numdatasegments) * g.bytespersampledenom)
                   / g.bytespersamplenum + 0.5)
    return g


def framed_constants(g: Segment, framebytes: int, framespersecond: int, subint_ns: int, guard_ns: int,
                     bufferbytes: int, numdatasegments: int) -> Segment:
    """VDIFDataStream / Mark5BDataStream::updateConfig (vdiffile.cpp:484-491,
    mark5bfile.cpp:258-266): nsinc from whole frames, sendbytes without the guard."""
    g.nsinc = cint(((bufferbytes // numdatasegments) // framebytes) * (1e9 / float(framespersecond)) + 0.5)
    g.sendbytes = cint((float(g.sendbytes) * float(subint_ns)) / (subint_ns + guard_ns) + 0.5)
    return g


def frames_per_second(bw_mhz: float, nrecbands: int, nbits: int, payloadbytes: int) -> int:
    """Configuration::getFramesPerSecond (configuration.cpp:654-668), held in an int."""
    return cint(2.0e6 * bw_mhz * nrecbands * nbits / (8.0 * payloadbytes))


def frame_granularity(fps: int) -> int:
    """The fewest frames that span whole nanoseconds (vdifmux.c:140)."""
    import math
    return fps // math.gcd(fps, NS)


def control_params(buf: Buffer, at: int, scan: int, offsetsec: int, offsetns: int, delay_ok: bool,
                   delayus1: float, delayus2: float,
                   advance: Optional[Callable[[int], bool]] = None) -> Tuple[List[int], int, int]:
    """Returns (control words, buffer index, segment) for the subint at (scan, offsetsec,
    offsetns).  An invalid subint has control[1] = INVALID_SUBINT, zero flags, index 0.
    `advance(at)` is called before each move to the next segment and returns True to stop
    (StateError); without it the choice stops after one pass round the ring."""
    segs, rb = buf.segs, buf.readbytes
    n = len(segs)

    def words(g):
        return [0] * (3 + cdiv(g.blockspersend + FLAGS_PER_INT - 1, FLAGS_PER_INT))

    ctrl = words(segs[at])
    ctrl[0] = scan

    def invalid(c):
        c[1] = INVALID_SUBINT
        for k in range(3, len(c)):
            c[k] = 0
        return c, 0, at

    if not delay_ok:
        return invalid(ctrl)
    g0 = segs[at]
    d1 = delayus1 - buf.intclockseconds * 1000000
    first = offsetns - cint(d1 * 1000)
    span = cint(g0.numchannels * g0.blockspersend * 2 * g0.sampletimens + 0.5)
    if g0.complex_sampling:
        span = cdiv(span, 2)
    d2 = delayus2 - buf.intclockseconds * 1000000
    last = offsetns + span + cint(-1000 * d2 + 0.5)
    for _ in range(2):
        if last < 0:
            offsetsec -= 1
            first += NS
            last += NS
    if scan < g0.scan or not g0.readto:
        return invalid(ctrl)
    if scan == g0.scan and offsetsec < g0.scanseconds - 3:
        return invalid(ctrl)

    def nsdiff(g):
        return (offsetsec - g.scanseconds) * NS + first - g.scanns

    def valid_ns(g):
        return cdiv(g.validbytes * g.nsinc, rb)

    nsd, vns = nsdiff(segs[at]), valid_ns(segs[at])
    moves = 0
    while (scan > segs[at].scan or (scan == segs[at].scan and nsd >= vns)) and \
            (buf.keepreading or at != buf.lastvalidsegment):
        # the reference waits here for its reader; `advance` returns True to stop
        moves += 1
        if (advance(at) if advance is not None else moves > n):
            raise StateError("the ring holds no segment at or after the subint start")
        at = (at + 1) % n
        nsd, vns = nsdiff(segs[at]), valid_ns(segs[at])
    g, nx = segs[at], segs[(at + 1) % n]
    ctrl = words(g)
    ctrl[0] = scan
    if scan < g.scan or nsd < -g.nsinc or scan > g.scan:
        return invalid(ctrl)
    fftch = g.numchannels if g.complex_sampling else 2 * g.numchannels
    blockbytes = cdiv(fftch * g.bytespersamplenum, g.bytespersampledenom)

    def dindex():
        return at * rb + ((((offsetsec - g.scanseconds) * 1.0e9 + (first - g.scanns)) / g.sampletimens)
                          * g.bytespersamplenum) / g.bytespersampledenom

    cur_t = g.scanseconds + g.scanns / 1.0e9
    nxt_t = nx.scanseconds + nx.scanns / 1.0e9
    if g.validbytes < rb or nx.scanseconds - g.scanseconds > 2 or nxt_t < cur_t:
        dbi = dindex()
        if nxt_t < cur_t and cint(dbi - at * rb) > g.validbytes:
            return invalid(ctrl)
        if dbi < -blockbytes * g.blockspersend or dbi > buf.bufferbytes:
            return invalid(ctrl)
    ctrl[1] = g.scanseconds
    samples = cint(((offsetsec - g.scanseconds) * NS + (first - g.scanns)) / g.sampletimens + 0.5)
    bi = at * rb + cdiv(samples * g.bytespersamplenum, g.bytespersampledenom)
    if cmod(bi, 2) != 0:
        bi -= 1
    if bi == buf.bufferbytes:
        if g.scan != nx.scan or ((nx.scanseconds - g.scanseconds) * NS + nx.scanns - g.scanns - g.nsinc) != 0:
            return invalid(ctrl)
    count = 0
    while bi < at * rb and count < g.blockspersend:
        bi += blockbytes
        count += 1
    if bi < at * rb:
        return invalid(ctrl)
    if count > 0:
        def tosub(k):
            return cdiv(cint(k * 1000.0 * (d2 - d1) / (g.sampletimens * g.blockspersend) + 0.5)
                        * g.bytespersamplenum, g.bytespersampledenom)
        if bi - tosub(count) < at * rb:
            count += 1
            bi += blockbytes
            if count == g.blockspersend:
                return invalid(ctrl)
            bi -= tosub(count)
        bi -= cmod(bi, 2)
    segoff = bi - at * rb
    bi -= cmod(segoff, g.bytesbetweenintegerns)
    segoff -= cmod(segoff, g.bytesbetweenintegerns)
    segoffns = cint(float(cdiv(segoff * g.bytespersampledenom, g.bytespersamplenum)) * g.sampletimens + 0.5)
    a = ctrl[1] + cdiv(segoffns, NS)
    b = g.scanns + cmod(segoffns, NS)
    a += cdiv(b, NS)
    b = cmod(b, NS)
    ctrl[1], ctrl[2] = a, b
    contig = g.validbytes == rb and \
        ((nx.scanseconds - g.scanseconds) * NS + nx.scanns - g.scanns) == g.nsinc

    def setflag(i):
        ctrl[3 + i // FLAGS_PER_INT] |= 1 << (i % FLAGS_PER_INT)

    if (g.validbytes - segoff >= g.sendbytes) or (contig and g.validbytes - segoff + nx.validbytes > g.sendbytes):
        for i in range(count, g.blockspersend):
            setflag(i)
    else:
        dbi = dindex()
        if dbi < -blockbytes * g.blockspersend or dbi > buf.bufferbytes:
            return invalid(ctrl)
        for i in range(count, g.blockspersend):
            if bi + i * blockbytes - at * rb < g.validbytes:
                setflag(i)
            elif contig and (g.validbytes - segoff + nx.validbytes) > i * blockbytes:
                setflag(i)
        if not any(ctrl[3 + i] > 0 for i in range(g.blockspersend // FLAGS_PER_INT)):
            return invalid(ctrl)
    if buf.framing:                         # vdiffile.cpp:418-467, mark5bfile.cpp:174-240
        off = bi - at * rb
        if off < 0:
            return invalid(ctrl)
        fin = cdiv(off, buf.payloadbytes)
        if cmod(fin, buf.framegranularity) != 0:
            fin -= cmod(fin, buf.framegranularity)
            if buf.framing == 2 and fin < 0:
                fin += buf.framegranularity
        if buf.framing == 1:
            segoffns = g.scanns + cint((1e9 * fin) / buf.framespersecond)
        else:
            segoffns = g.scanns + cdiv(NS * fin, buf.framespersecond)
        ctrl[1] = g.scanseconds + cdiv(segoffns, NS)
        ctrl[2] = cmod(segoffns, NS)
        bi = at * rb + fin * buf.framebytes
        if bi == buf.bufferbytes and (g.scan != nx.scan or
                                      (nx.scanseconds - g.scanseconds) * NS + nx.scanns - g.scanns - g.nsinc != 0):
            return invalid(ctrl)
        if bi > buf.bufferbytes:
            return invalid(ctrl)
    return ctrl, bi, at
