"""DataStream::calculateControlParams (datastream.cpp:368-627): the product's host function
fxg_control_params against hand-derived cases and against the restatement in
oracle/controlparams.py on random buffer states; fxg_segment_params against
DataStream::updateConfig's constants (datastream.cpp:745-763)."""
import ctypes as C
from pathlib import Path

import numpy as np
import pytest

import controlparams as cp

ROOT = Path(__file__).resolve().parent.parent

# 16 bands x 2 bits = 4 bytes per sample, 32 MHz (15.625 ns samples, whole ns every 8 samples
# = 32 bytes), nchan 1024: 8192 bytes per FFT block
R = 8192 * 200                       # bytes per segment
NSINC = 6_400_000                    # 15.625 ns x R / 4


def _seg(scanseconds=10, scanns=0, validbytes=R, B=64, scan=0, readto=True):
    return cp.Segment(scan=scan, scanseconds=scanseconds, scanns=scanns, validbytes=validbytes, readto=readto,
                      sendbytes=64 * 8192 + 8192, blockspersend=B, bytespersamplenum=4, bytespersampledenom=1,
                      bytesbetweenintegerns=32, nsinc=NSINC, numchannels=1024, complex_sampling=False,
                      sampletimens=15.625)


def _ring(**kw0):
    segs = [_seg(scanns=k * NSINC, **kw0) for k in range(4)]
    return cp.Buffer(segs=segs, bufferbytes=4 * R, readbytes=R)


def _lib():
    if not (ROOT / "difx_amd" / "libfxg.so").exists():
        pytest.skip("libfxg.so not built")
    from difx_amd.core import load_library
    return load_library()


def _c_call(buf: cp.Buffer, at, scan, sec, ns, ok, d1, d2):
    """fxg_control_params through the C ABI; returns (rc, control, index, segment, advances)."""
    from difx_amd import fxg_abi as abi
    lib = _lib()
    n = len(buf.segs)
    arr = (abi.fxg_segment * n)()
    for i, g in enumerate(buf.segs):
        a = arr[i]
        a.scan, a.scanseconds, a.scanns, a.validbytes = g.scan, g.scanseconds, g.scanns, g.validbytes
        a.readto, a.sendbytes, a.blockspersend = int(g.readto), g.sendbytes, g.blockspersend
        a.bytespersamplenum, a.bytespersampledenom = g.bytespersamplenum, g.bytespersampledenom
        a.bytesbetweenintegerns, a.nsinc, a.numchannels = g.bytesbetweenintegerns, g.nsinc, g.numchannels
        a.complex_sampling, a.sampletimens = int(g.complex_sampling), g.sampletimens
    db = abi.fxg_databuffer(n, buf.bufferbytes, buf.readbytes, int(buf.keepreading), buf.lastvalidsegment,
                            buf.intclockseconds, buf.framing, buf.framebytes, buf.payloadbytes,
                            buf.framespersecond, buf.framegranularity, 0, arr)
    calls = []

    def adv(user, frm):                      # a reader that never refills: stop after one pass
        calls.append(frm)
        return 1 if len(calls) > n else 0
    cb = abi.FXG_ADVANCE_FN(adv)
    maxw = 3 + max((g.blockspersend + 29) // 30 for g in buf.segs)
    ctrl = (C.c_int32 * maxw)()
    bi = C.c_int32(-7)
    atc = C.c_int32(at)
    rc = lib.fxg_control_params(C.byref(db), C.byref(atc), scan, sec, ns, int(ok), d1, d2, cb, None, ctrl,
                                C.byref(bi))
    nw = 3 + (buf.segs[atc.value].blockspersend + 29) // 30
    return rc, list(ctrl)[:nw], bi.value, atc.value, calls


def _both(buf, at, scan, sec, ns, ok=True, d1=0.0, d2=0.0):
    calls = []
    n = len(buf.segs)

    def adv(frm):
        calls.append(frm)
        return len(calls) > n
    try:
        ref = cp.control_params(buf, at, scan, sec, ns, ok, d1, d2, advance=adv)
    except cp.StateError:
        ref = None
    rc, ctrl, bi, atc, ccalls = _c_call(buf, at, scan, sec, ns, ok, d1, d2)
    assert ccalls == calls
    if ref is None:
        from difx_amd import fxg_abi as abi
        assert rc == abi.FXG_E_STATE
        return (None, None, None), calls
    assert rc == 0
    assert (ctrl, bi, atc) == (ref[0], ref[1], ref[2]), (ctrl, bi, atc, ref)
    return ref, calls


FULL = (1 << 30) - 1


def test_aligned_start_all_blocks_valid():
    """Zero delay, subint 1 ms into a full segment: 64000 samples = 256000 bytes in, start
    time 10 s + 1 ms, all 64 blocks valid."""
    (ctrl, bi, at), _ = _both(_ring(), 0, 0, 10, 1_000_000)
    assert bi == 256000 and at == 0
    assert ctrl == [0, 10, 1_000_000, FULL, FULL, 0xF]


def test_delay_of_whole_samples_moves_the_start_back():
    """A 2.5 us delay starts the data 160 samples (640 bytes) earlier."""
    (ctrl, bi, _), _ = _both(_ring(), 0, 0, 10, 1_000_000, d1=2.5, d2=2.5)
    assert bi == 256000 - 640 and ctrl[1:3] == [10, 997_500]


def test_start_backs_off_to_a_whole_nanosecond_byte():
    """A 10 ns delay: 63999 samples (255996 bytes), backed off to the previous multiple of
    32 bytes (8 samples = 125 ns): 255968 bytes = 63992 samples = 999875 ns."""
    (ctrl, bi, _), _ = _both(_ring(), 0, 0, 10, 1_000_000, d1=0.01, d2=0.01)
    assert bi == 255968 and ctrl[1:3] == [10, 999_875]


def test_partly_valid_segment_flags_the_blocks_with_data():
    """The segment holds 20 blocks past the start and the next one is 10 s later: blocks
    0-19 valid."""
    buf = _ring()
    buf.segs[0].validbytes = 256000 + 20 * 8192
    buf.segs[1].scanseconds = 20
    (ctrl, bi, _), _ = _both(buf, 0, 0, 10, 1_000_000)
    assert bi == 256000 and ctrl[3:] == [(1 << 20) - 1, 0, 0]


def test_good_data_check_ignores_a_partial_flag_word():
    """The same with 16 blocks per subint: the reference's good-data check reads whole flag
    words only (datastream.cpp:607-614), finds none and drops the subint."""
    buf = _ring(B=16)
    buf.segs[0].validbytes = 256000 + 10 * 8192
    buf.segs[1].scanseconds = 20
    (ctrl, bi, _), _ = _both(buf, 0, 0, 10, 1_000_000)
    assert ctrl[1] == cp.INVALID_SUBINT and bi == 0


def test_start_in_the_next_segment_moves_on():
    """7 ms in: past segment 0's 6.4 ms, so the choice moves to segment 1 (advance called
    once, from 0) and the start is 0.6 ms = 153600 bytes into it."""
    (ctrl, bi, at), calls = _both(_ring(), 0, 0, 10, 7_000_000)
    assert calls == [0] and at == 1 and bi == R + 153600 and ctrl[1:3] == [10, 7_000_000]


def test_start_before_the_segment_skips_a_block():
    """A 0.1 us delay at the segment's first nanosecond: the start is 5 samples before the
    buffer, so block 0 is skipped (index 8172, backed off to 8160 = 2040 samples = 31875 ns)
    and flagged invalid."""
    (ctrl, bi, _), _ = _both(_ring(), 0, 0, 10, 0, d1=0.1, d2=0.1)
    assert bi == 8160 and ctrl[1:3] == [10, 31_875] and ctrl[3:] == [FULL - 1, FULL, 0xF]


@pytest.mark.parametrize("case", ["no_model", "earlier_scan", "unread", "too_early", "later_scan"])
def test_invalid_exits(case):
    buf, at, scan, sec, ok = _ring(), 0, 0, 10, True
    if case == "no_model":
        ok = False
    elif case == "earlier_scan":
        for g in buf.segs:
            g.scan = 1
    elif case == "unread":
        buf.segs[0].readto = False
    elif case == "too_early":
        sec = 6
    elif case == "later_scan":
        scan = 1
        buf.keepreading, buf.lastvalidsegment = False, 0
    (ctrl, bi, _), _ = _both(buf, at, scan, sec, 1_000_000, ok=ok)
    assert ctrl[1] == cp.INVALID_SUBINT and bi == 0 and not any(ctrl[3:])


def test_random_buffer_states_match_the_restatement():
    """3000 random rings (2-16 segments; 1-32 bands of 2 or 8 bits, real or complex; 1-64 MHz;
    partly read, gapped and next-scan segments) and subint times around them."""
    rng = np.random.default_rng(17)
    n_valid = n_state = 0
    for case in range(3000):
        nbits = int(rng.choice([2, 8]))
        cplx = bool(rng.random() < 0.3)
        nbands = int(rng.choice([1, 2, 4, 8, 16, 32]))
        bw = float(rng.choice([1.0, 2.0, 4.0, 8.0, 16.0, 32.0, 64.0]))
        nchan = int(rng.choice([64, 128, 256, 1024]))
        B = int(rng.integers(4, 100))
        nseg = int(rng.integers(2, 17))
        blocks_per_seg = int(rng.integers(max(2, B // 4), 3 * B))
        proto = cp.segment_constants(nbits, nbands, cplx, bw, nchan, B, 1, 1, 0, 1 << 20, 1)
        fftch = nchan if cplx else 2 * nchan
        blockbytes = fftch * proto.bytespersamplenum // proto.bytespersampledenom
        readbytes = blocks_per_seg * blockbytes
        readbytes -= readbytes % 8
        if readbytes * nseg >= 2**31 - 1 or readbytes < 64:
            continue
        g0 = cp.segment_constants(nbits, nbands, cplx, bw, nchan, B, B * blockbytes, 1_000_000, 10_000,
                                  readbytes * nseg, nseg)
        segs = []
        t = int(rng.integers(0, 5)) * cp.NS + int(rng.integers(0, cp.NS))
        scan = 0
        for k in range(nseg):
            g = cp.Segment(**g0.__dict__)
            if rng.random() < 0.1:
                scan += 1
            g.scan = scan
            g.scanseconds, g.scanns = t // cp.NS, t % cp.NS
            g.validbytes = readbytes if rng.random() < 0.7 else int(rng.integers(0, readbytes + 1))
            g.readto = bool(rng.random() < 0.95)
            segs.append(g)
            t += g.nsinc + (0 if rng.random() < 0.8 else int(rng.integers(1, 3 * cp.NS)))
        buf = cp.Buffer(segs=segs, bufferbytes=readbytes * nseg, readbytes=readbytes,
                        keepreading=bool(rng.random() < 0.8), lastvalidsegment=int(rng.integers(0, nseg)),
                        intclockseconds=int(rng.integers(-1, 2)))
        at = int(rng.integers(0, nseg))
        g = segs[at]
        tq = g.scanseconds * cp.NS + g.scanns + int(rng.integers(-2 * g.nsinc, 3 * g.nsinc))
        d1 = float(rng.normal(0, 3000)) + buf.intclockseconds * 1e6
        d2 = d1 + float(rng.normal(0, 2))
        ref, _ = _both(buf, at, g.scan + int(rng.integers(-1, 2) if rng.random() < 0.1 else 0),
                       int(tq // cp.NS), int(tq % cp.NS), ok=bool(rng.random() < 0.98), d1=d1, d2=d2)
        if ref[0] is None:
            n_state += 1
        else:
            n_valid += ref[0][1] != cp.INVALID_SUBINT
    assert n_valid > 500 and n_state < 1000


def test_segment_constants_from_a_configuration():
    """fxg_segment_params (no device) equals updateConfig's constants (the VDIF frame-based
    nsinc and sendbytes included) for the 8 x 32 MHz dual-pol 2-bit datastream: 4 bytes per
    sample, 15.625 ns samples, whole nanoseconds every 32 bytes; and for a 2-band complex
    8-bit datastream."""
    lib = _lib()
    from difx_amd import config as cf
    from difx_amd import fxg_abi as abi
    for kind in ("real2", "cplx8"):
        if kind == "real2":
            cfg = cf.standard_config(2, 2, [4800.0 + 40.0 * k for k in range(8)], 32.0, 1024, subint_ns=8_000_000)
            args = (2, 16, False, 32.0, 1024)
        else:
            cfg = cf.standard_config(2, 2, [4800.0], 16.0, 256, subint_ns=8_000_000, nbits=8, complex_sampling=1)
            args = (8, 2, True, 16.0, 256)
        fa, keep = cfg.to_abi()
        seg = abi.fxg_segment()
        assert lib.fxg_segment_params(C.byref(fa), 0, 4, 4 * R, C.byref(seg)) == 0
        from difx_amd import synth
        B = cfg.blockspersend()
        want = cp.segment_constants(*args, B, synth.data_bytes(cfg, 0), 8_000_000, cfg.guard_ns, 4 * R, 4)
        ds = cfg.datastreams[0]
        fps = cp.frames_per_second(args[3], len(ds.recbands), args[0], ds.framebytes - 32)
        want = cp.framed_constants(want, ds.framebytes, fps, 8_000_000, cfg.guard_ns, 4 * R, 4)
        got = (seg.bytespersamplenum, seg.bytespersampledenom, seg.bytesbetweenintegerns, seg.nsinc,
               seg.numchannels, seg.blockspersend, seg.complex_sampling, seg.sampletimens, seg.sendbytes)
        assert got == (want.bytespersamplenum, want.bytespersampledenom, want.bytesbetweenintegerns, want.nsinc,
                       want.numchannels, B, int(want.complex_sampling), want.sampletimens, want.sendbytes)
        if kind == "real2":       # 1638400 / 8032 = 203 whole frames of 31250 ns
            assert got[:4] == (4, 1, 32, 203 * 31250) and seg.sampletimens == 15.625
        else:                     # 2 bands x 16 bits: 4 bytes per 62.5 ns complex sample
            assert got[:3] == (4, 1, 8) and seg.sampletimens == 62.5


def _vdif_ring():
    """The same stream as 8032-byte VDIF frames (8000-byte payloads, 2000 samples = 31250 ns,
    32000 frames/s): 200 frames per segment, 6.25 ms each."""
    rf = 200 * 8032
    segs = [_seg(scanns=k * 6_250_000) for k in range(4)]
    for g in segs:
        g.validbytes, g.nsinc = rf, 6_250_000
    return cp.Buffer(segs=segs, bufferbytes=4 * rf, readbytes=rf, framing=1, framebytes=8032, payloadbytes=8000,
                     framespersecond=32000, framegranularity=1)


def test_vdif_framing_steps_back_to_a_frame():
    """1 ms in: 256000 payload bytes = 32 frames, whose start is 1 ms (32 x 31250 ns) and byte
    32 x 8032; with a 10 ns delay the payload index 255968 is in frame 31 (968750 ns)."""
    (ctrl, bi, _), _ = _both(_vdif_ring(), 0, 0, 10, 1_000_000)
    assert bi == 32 * 8032 and ctrl[1:3] == [10, 1_000_000] and ctrl[3:] == [FULL, FULL, 0xF]
    (ctrl, bi, _), _ = _both(_vdif_ring(), 0, 0, 10, 1_000_000, d1=0.01, d2=0.01)
    assert bi == 31 * 8032 and ctrl[1:3] == [10, 968_750]


def test_mark5b_framing_with_a_granularity_of_two():
    """Mark5B frames of 10000 payload bytes at 25600 frames/s (39062.5 ns, granularity 2):
    1 ms in is 25.6 frames, stepped back to frame 24 (937500 ns, byte 24 x 10016)."""
    rf = 100 * 10016
    segs = [_seg(scanns=k * 3_906_250) for k in range(4)]
    for g in segs:
        g.validbytes, g.nsinc = rf, 3_906_250
        g.sampletimens, g.bytesbetweenintegerns = 1e9 / (25600 * 2500), 32   # 2500 samples per frame
    buf = cp.Buffer(segs=segs, bufferbytes=4 * rf, readbytes=rf, framing=2, framebytes=10016, payloadbytes=10000,
                    framespersecond=25600, framegranularity=2)
    assert cp.frame_granularity(25600) == 2
    (ctrl, bi, _), _ = _both(buf, 0, 0, 10, 1_000_000)
    assert bi == 24 * 10016 and ctrl[1:3] == [10, 937_500]


def test_random_framed_states_match_the_restatement():
    """1500 random VDIF and Mark5B rings against the restatement, framing included."""
    rng = np.random.default_rng(29)
    n_valid = 0
    for case in range(1500):
        mk5 = bool(rng.random() < 0.5)
        nbands = int(rng.choice([1, 2, 4, 8, 16]))
        bw = float(rng.choice([4.0, 8.0, 16.0, 32.0, 64.0]))
        nchan = int(rng.choice([128, 256, 1024]))
        B = int(rng.integers(8, 90))
        payload = 10000 if mk5 else int(rng.choice([1000, 5000, 8000]))
        fb = payload + (16 if mk5 else 32)
        g0 = cp.segment_constants(2, nbands, False, bw, nchan, B, B * 2 * nchan * nbands * 2 // 8 + 4 * fb,
                                  1_000_000, 2000, 1, 1)
        fps = cp.frames_per_second(bw, nbands, 2, payload)
        if fps < 1 or (2e6 * bw * nbands * 2) % (8 * payload):
            continue
        nseg = int(rng.integers(2, 9))
        rf = fb * int(rng.integers(20, 400))
        g0 = cp.segment_constants(2, nbands, False, bw, nchan, B, g0.sendbytes, 1_000_000, 2000, rf * nseg, nseg)
        g0 = cp.framed_constants(g0, fb, fps, 1_000_000, 2000, rf * nseg, nseg)
        t = int(rng.integers(0, 3)) * cp.NS + int(rng.integers(0, cp.NS))
        segs = []
        for k in range(nseg):
            g = cp.Segment(**g0.__dict__)
            g.scanseconds, g.scanns = t // cp.NS, t % cp.NS
            g.validbytes = rf if rng.random() < 0.8 else int(rng.integers(0, rf + 1))
            segs.append(g)
            t += g.nsinc + (0 if rng.random() < 0.85 else int(rng.integers(1, 2 * cp.NS)))
        buf = cp.Buffer(segs=segs, bufferbytes=rf * nseg, readbytes=rf, keepreading=bool(rng.random() < 0.8),
                        lastvalidsegment=int(rng.integers(0, nseg)), framing=2 if mk5 else 1, framebytes=fb,
                        payloadbytes=payload, framespersecond=fps, framegranularity=cp.frame_granularity(fps))
        at = int(rng.integers(0, nseg))
        g = segs[at]
        tq = g.scanseconds * cp.NS + g.scanns + int(rng.integers(-g.nsinc, 3 * g.nsinc))
        d1 = float(rng.normal(0, 2000))
        ref, _ = _both(buf, at, 0, int(tq // cp.NS), int(tq % cp.NS), d1=d1, d2=d1 + float(rng.normal(0, 1)))
        if ref[0] is not None:
            n_valid += ref[0][1] != cp.INVALID_SUBINT
    assert n_valid > 300


def test_buffer_and_framed_segment_params_from_a_configuration():
    """fxg_buffer_params and fxg_segment_params of the 8 x 32 MHz VDIF datastream (8000-byte
    payloads): 32000 frames/s, granularity 1, nsinc = whole frames per segment x 31250 ns,
    sendbytes without the guard."""
    lib = _lib()
    from difx_amd import config as cf
    from difx_amd import fxg_abi as abi
    from difx_amd import synth
    cfg = cf.standard_config(2, 2, [4800.0 + 40.0 * k for k in range(8)], 32.0, 1024, subint_ns=8_000_000)
    fa, keep = cfg.to_abi()
    rf = 200 * 8032
    db = abi.fxg_databuffer()
    assert lib.fxg_buffer_params(C.byref(fa), 0, 4, 4 * rf, C.byref(db)) == 0
    assert (db.readbytes, db.framing, db.framebytes, db.payloadbytes, db.framespersecond, db.framegranularity) == \
        (rf, abi.FXG_FRAMING_VDIF, 8032, 8000, 32000, 1)
    seg = abi.fxg_segment()
    assert lib.fxg_segment_params(C.byref(fa), 0, 4, 4 * rf, C.byref(seg)) == 0
    assert seg.nsinc == 6_250_000
    want = cp.segment_constants(2, 16, False, 32.0, 1024, cfg.blockspersend(), synth.data_bytes(cfg, 0),
                                8_000_000, cfg.guard_ns, 4 * rf, 4)
    want = cp.framed_constants(want, 8032, 32000, 8_000_000, cfg.guard_ns, 4 * rf, 4)
    assert (seg.sendbytes, seg.nsinc) == (want.sendbytes, want.nsinc)
