"""Pins of the VDIF corner-turn restatement (oracle/vdifmux.py) on the CPU: the reference's own
mask-and-shift turners, decoding the muxed stream against its single-thread parts, a
split/shuffle/mux round trip, and hand-built segments for every header rule of
VDIFMuxer::multiplex (datamuxer.cpp:670-757) and deinterlace (:764-833)."""
from types import SimpleNamespace

import numpy as np
import pytest

import fxoracle as fo
import vdifmux as vm


@pytest.mark.parametrize("n", [2, 4, 8, 16])
def test_generic_turner_equals_the_2bit_special_turners(n):
    """cornerturn_generic (datamuxer.cpp:243-266) against the reference's 2-bit turners for
    2, 4, 8 and 16 threads (:279-532) on random payloads."""
    rng = np.random.default_rng(n)
    tb = rng.integers(0, 256, size=(n, 64), dtype=np.uint8)
    tw = np.stack([np.frombuffer(t.tobytes(), dtype="<u4") for t in tb])
    gen = vm.cornerturn_generic(tw, 2).astype("<u4").view(np.uint8)
    np.testing.assert_array_equal(gen, vm.cornerturn_2bit_special(tb))


def _multichannel_stream(nframes, nchan, bits, pay, fps, rng, sec=1000, epoch=40):
    fb = pay + vm.HEADER_BYTES
    buf = np.zeros(nframes * fb, dtype=np.uint8)
    lg = int(np.log2(nchan))
    for f in range(nframes):
        h = vm.make_header(sec + f // fps, f % fps, epoch, fb, 0, bits, lognchan=lg)
        buf[f * fb:f * fb + 32] = h.astype("<u4").view(np.uint8)
        buf[f * fb + 32:(f + 1) * fb] = rng.integers(0, 256, size=pay, dtype=np.uint8)
    return buf, fb


@pytest.mark.parametrize("nthreads,bits", [(16, 2), (4, 2), (2, 8), (8, 1)])
def test_split_shuffle_mux_round_trip(nthreads, bits):
    """An N-channel stream split into N single-channel threads (thread ids scrambled), frames
    shuffled across threads, then muxed: the original bytes come back."""
    rng = np.random.default_rng(7 + nthreads)
    fps = 5
    buf, fb = _multichannel_stream(12, nthreads, bits, 8 * 64, fps, rng)
    tids = [int(x) for x in rng.permutation(1000)[:nthreads]]
    parts = vm.split_threads(buf, fb, nthreads, bits, tids)
    tfb = (fb - 32) // nthreads + 32
    frames = [p[i * tfb:(i + 1) * tfb] for p in parts for i in range(len(p) // tfb)]
    order = rng.permutation(len(frames))
    inbuf = np.concatenate([frames[i] for i in order])
    w0 = vm.header_words(buf)
    out, nout, st = vm.mux_segment(inbuf, tfb, tids, fps, bits, vm.full_seconds(w0), 0, 0, 12)
    assert nout == 12 and st == {"early": 0, "late": 0, "unmapped": 0}
    np.testing.assert_array_equal(out, buf)


def test_muxed_stream_decodes_to_the_thread_samples():
    """Decoding (the unpack restatement, pinned in test_oracle.py) the muxed 16-channel stream
    gives, for channel l, the samples of thread l decoded as a one-channel stream."""
    rng = np.random.default_rng(3)
    n, bits, fps = 16, 2, 4
    tfb = 32 + 8 * 32
    tids = list(range(100, 100 + n))
    threads = []
    for l in range(n):
        t = np.zeros(8 * tfb, dtype=np.uint8)
        for f in range(8):
            t[f * tfb:f * tfb + 32] = vm.make_header(50 + f // fps, f % fps, 40, tfb, tids[l], bits).astype("<u4").view(np.uint8)
            t[f * tfb + 32:(f + 1) * tfb] = rng.integers(0, 256, size=tfb - 32, dtype=np.uint8)
        threads.append(t)
    inbuf = np.concatenate(threads)
    out, nout, _ = vm.mux_segment(inbuf, tfb, tids, fps, bits,
                                  vm.full_seconds(vm.header_words(threads[0])), 0, 0, 8)
    assert nout == 8
    ofb = (tfb - 32) * n + 32
    dsm = SimpleNamespace(format="VDIF", recbands=[None] * n, nbits=2, complex_sampling=0,
                          framebytes=ofb, headerbytes=32)
    vals, valid = fo.decode_stream(out, dsm)
    assert valid.all()
    ds1 = SimpleNamespace(format="VDIF", recbands=[None], nbits=2, complex_sampling=0,
                          framebytes=tfb, headerbytes=32)
    for l in range(n):
        v1, _ = fo.decode_stream(threads[l], ds1)
        np.testing.assert_array_equal(vals[l], v1[0])


def _frame(sec, frame, tid, pay_byte, tfb=64, bits=2, epoch=40):
    f = np.full(tfb, pay_byte, dtype=np.uint8)
    f[:32] = vm.make_header(sec, frame, epoch, tfb, tid, bits).astype("<u4").view(np.uint8)
    return f


def test_header_rules_on_a_hand_built_segment():
    """Two threads (ids 7, 3), fps 4, segment start at the reference frame (second 10, frame 2).
    slot 0: no thread (leading: not output); slot 1: both (muxed; thread 7's frame appears
    twice, the later copy wins); slot 2: only thread 3 (its header, invalid, output length,
    thread id 0, channel count untouched); slot 3: none (slot 2's header, frame + 1 -> wraps
    to frame 0 of second 11); slot 4: none (frame 1 of second 11).  Also one frame before the
    segment, one after it and one of an unmapped thread, all dropped."""
    tids, fps, tfb = [7, 3], 4, 64
    ofb = (tfb - 32) * 2 + 32
    frames = [
        _frame(10, 3, 7, 0x11),          # slot 1, thread 0 (first copy)
        _frame(10, 3, 3, 0x22),          # slot 1, thread 1
        _frame(10, 3, 7, 0x55),          # slot 1, thread 0 again: replaces the first copy
        _frame(11, 0, 3, 0x33),          # slot 2, thread 1 only
        _frame(10, 1, 7, 0x44),          # before the segment
        _frame(11, 3, 7, 0x44),          # after it (slot 5 of 5)
        _frame(11, 0, 9, 0x44),          # unmapped thread
    ]
    buf = np.concatenate(frames)
    ref_sec = vm.epoch_mjd(40) * 86400 + 10
    out, nout, st = vm.mux_segment(buf, tfb, tids, fps, 2, ref_sec, 2, 0, 5)
    assert nout == 4 and st == {"early": 1, "late": 1, "unmapped": 1}
    hdr = lambda o: np.frombuffer(out[o * ofb:o * ofb + 32].tobytes(), dtype="<u4")
    # slot 1 -> output 0: thread 7's (second copy) header, invalid 0, 2 channels, length, tid 0
    np.testing.assert_array_equal(hdr(0), vm.make_header(10, 3, 40, ofb, 0, 2, lognchan=1))
    tw = np.stack([np.full(8, 0x55555555, dtype=np.uint32), np.full(8, 0x22222222, dtype=np.uint32)])
    np.testing.assert_array_equal(out[32:ofb], vm.cornerturn_generic(tw, 2).astype("<u4").view(np.uint8))
    # slot 2 -> output 1: thread 3's header, invalid, output length, tid 0, 1 channel
    np.testing.assert_array_equal(hdr(1), vm.make_header(11, 0, 40, ofb, 0, 2, lognchan=0, invalid=1))
    assert not out[ofb + 32:2 * ofb].any()            # payload not written
    # slot 3 -> output 2: previous header, frame + 1
    np.testing.assert_array_equal(hdr(2), vm.make_header(11, 1, 40, ofb, 0, 2, lognchan=0, invalid=1))
    # slot 4 -> output 3
    np.testing.assert_array_equal(hdr(3), vm.make_header(11, 2, 40, ofb, 0, 2, lognchan=0, invalid=1))


def test_frame_number_wraps_into_the_next_second():
    """A none-slot after frame fps-1 takes frame 0 of the next second (datamuxer.cpp:722-725)."""
    tids, fps, tfb = [1], 4, 64
    ofb = tfb
    buf = np.concatenate([_frame(20, 3, 1, 0x66)])
    ref_sec = vm.epoch_mjd(40) * 86400 + 20
    out, nout, _ = vm.mux_segment(buf, tfb, tids, fps, 2, ref_sec, 3, 0, 3)
    assert nout == 3
    hdr = lambda o: np.frombuffer(out[o * ofb:o * ofb + 32].tobytes(), dtype="<u4")
    np.testing.assert_array_equal(hdr(0), vm.make_header(20, 3, 40, ofb, 0, 2))
    np.testing.assert_array_equal(hdr(1), vm.make_header(21, 0, 40, ofb, 0, 2, invalid=1))
    np.testing.assert_array_equal(hdr(2), vm.make_header(21, 1, 40, ofb, 0, 2, invalid=1))


def test_epoch_mjd_closed_form():
    """getVDIFEpochMJD: epoch 0 = 2000-01-01 = MJD 51544; epoch 1 = 2000-07-01 = MJD 51726;
    epoch 40 = 2020-01-01 = MJD 58849 (calendar arithmetic)."""
    assert vm.epoch_mjd(0) == 51544 and vm.epoch_mjd(1) == 51726 and vm.epoch_mjd(40) == 58849


def test_host_mirror_epoch_matches_the_restatement():
    """difx_amd.vdifmux's epoch MJD (used by VDIFMuxer.initialise) equals the restatement's
    ymd2mjd-based one for every 6-bit epoch."""
    from difx_amd.vdifmux import _epoch_mjd
    for e in range(64):
        assert _epoch_mjd(e) == vm.epoch_mjd(e), e


def test_mux_entry_point_validates_without_a_device():
    """fxg_vdif_mux_device rejects what the corner turn cannot do before touching the device:
    thread counts that are not powers of two, more bits than a word holds, frame lengths that
    are not a multiple of 8, too small output or workspace buffers, misaligned buffers."""
    import ctypes as C
    from pathlib import Path
    from difx_amd import fxg_abi as abi
    lp = Path(__file__).resolve().parent.parent / "difx_amd" / "libfxg.so"
    if not lp.exists():
        pytest.skip("libfxg.so not built")
    from difx_amd.core import load_library
    lib = load_library()
    assert lib.fxg_vdif_mux_workspace_bytes(100, 16) == 64 + 4 * 100 * 16 + 4 * 100

    def call(n=16, bits=2, ifb=8032, nslots=10, out=None, work=None, din=0x1000, dout=0x2000, dwork=0x3000):
        p = abi.fxg_vdif_mux_params()
        p.inputframebytes, p.nthreads, p.bitspersample, p.framespersecond = ifb, n, bits, 2000
        ofb = (ifb - 32) * n + 32
        out = nslots * ofb if out is None else out
        work = lib.fxg_vdif_mux_workspace_bytes(nslots, n) if work is None else work
        return lib.fxg_vdif_mux_device(C.byref(p), C.c_void_p(din), ifb * 4, C.c_void_p(dout), out, nslots,
                                       C.c_void_p(dwork), work, None)
    E = abi.FXG_E_INVALID
    assert call(n=3) == E
    assert call(n=64) == E
    assert call(n=32, bits=2) == E
    assert call(bits=3) == E
    assert call(ifb=8036) == E
    assert call(ifb=32) == E
    assert call(out=10 * ((8032 - 32) * 16 + 32) - 1) == E
    assert call(work=100) == E
    assert call(din=0x1004) == E
    assert call(dout=0x2008) == E
    assert call(nslots=0) == E
