"""Device VDIF corner turn (fxg_vdif_mux_device through difx_amd.vdifmux.VDIFMuxer) against
the restatement of VDIFMuxer (oracle/vdifmux.py): output bytes, output frame count and drop
counters bit-exact on the same inputs, including missing, duplicate, early, late and unmapped
frames; the split -> shuffle -> mux round trip at every (threads, bits) pair; a one-station
segment of the 16-channel 2-bit 32 MHz recording at full size (1000 frame times of 16
threads, 8000-byte thread payloads); and the muxed stream correlated on the device."""
import sys

import numpy as np
import pytest

import vdifmux as vm

pytestmark = pytest.mark.gpu

FILL = 0x5A            # the output buffer's prior contents: bytes the corner turn must not write


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _thread_frames(nslots, n, tfb, fps, tids, rng, sec0=1000, epoch=40, bits=2):
    """frames[s][l]: single-thread frame of thread l at frame time s (random payload)."""
    frames = []
    for s in range(nslots):
        row = []
        for l in range(n):
            f = np.empty(tfb, dtype=np.uint8)
            f[:32] = vm.make_header(sec0 + s // fps, s % fps, epoch, tfb, tids[l], bits).view(np.uint8)
            f[32:] = rng.integers(0, 256, size=tfb - 32, dtype=np.uint8)
            row.append(f)
        frames.append(row)
    return frames


def _run_both(gpu, inbuf, tfb, tids, fps, bits, ref_sec, ref_frame, start, nslots):
    import torch
    from difx_amd.vdifmux import VDIFMuxer
    n = len(tids)
    ofb = (tfb - 32) * n + 32
    ref_out = np.full(nslots * ofb, FILL, dtype=np.uint8)
    ref_out, ref_n, ref_st = vm.mux_segment(inbuf, tfb, tids, fps, bits, ref_sec, ref_frame, start, nslots,
                                            out=ref_out)
    mux = VDIFMuxer(n, tfb, fps, bits, tids)
    mux.set_reference(ref_sec, ref_frame)
    d_in = torch.from_numpy(inbuf).to(gpu)
    d_out = torch.full((nslots * ofb,), FILL, dtype=torch.uint8, device=gpu)
    d_out, nbytes = mux.multiplex(d_in, nslots, out=d_out, start=start)
    got = d_out.cpu().numpy()
    return got, nbytes // ofb, dict(mux.dropped), ref_out, ref_n, ref_st


@pytest.mark.parametrize("n,bits", [(16, 2), (8, 2), (4, 2), (2, 2), (1, 2), (32, 1), (16, 1), (8, 4),
                                    (4, 8), (2, 16), (1, 32)])
def test_split_shuffle_mux_round_trip(gpu, n, bits):
    """An N-channel stream split into N threads (scrambled thread ids), frames shuffled, muxed
    on the device: the original bytes come back, equal to the oracle's output."""
    _say("vdifmux round trip", n, bits)
    rng = np.random.default_rng(100 + 7 * n + bits)
    fps, nfr = 7, 24
    ofb = 32 + 8 * 16 * n                      # thread payload 128 bytes
    buf = np.zeros(nfr * ofb, dtype=np.uint8)
    lg = int(np.log2(n))
    for f in range(nfr):
        buf[f * ofb:f * ofb + 32] = vm.make_header(500 + f // fps, f % fps, 41, ofb, 0, bits, lognchan=lg).view(np.uint8)
        buf[f * ofb + 32:(f + 1) * ofb] = rng.integers(0, 256, size=ofb - 32, dtype=np.uint8)
    tids = [int(x) for x in rng.permutation(1024)[:n]]
    parts = vm.split_threads(buf, ofb, n, bits, tids)
    tfb = (ofb - 32) // n + 32
    frames = [p[i * tfb:(i + 1) * tfb] for p in parts for i in range(len(p) // tfb)]
    inbuf = np.concatenate([frames[i] for i in rng.permutation(len(frames))])
    ref_sec = vm.full_seconds(vm.header_words(buf))
    got, nout, st, ref, ref_n, ref_st = _run_both(gpu, inbuf, tfb, tids, fps, bits, ref_sec, 0, 0, nfr)
    assert nout == ref_n == nfr and st == ref_st
    np.testing.assert_array_equal(got, ref)
    np.testing.assert_array_equal(got, buf)


def test_damaged_segment(gpu):
    """16 threads, fps 25, 150 frame times starting mid-second: 3 leading frame times with
    no thread, 10% of (time, thread) frames missing, 6 frame times with no thread across a
    second boundary, duplicates with other payloads, and frames before the segment, after it
    and of unmapped threads.  Output bytes (the payload of incomplete frames unwritten), frame
    count and drop counters equal the oracle's."""
    _say("vdifmux damaged segment")
    rng = np.random.default_rng(77)
    n, bits, fps, tfb, nslots = 16, 2, 25, 32 + 8 * 40, 150
    tids = [int(x) for x in rng.permutation(1024)[:n]]
    sec0, f0 = 2000, 11
    frames = []
    for s in range(-4, nslots + 5):             # 4 before the segment, 5 after
        t = f0 + s
        for l in range(n):
            f = np.empty(tfb, dtype=np.uint8)
            f[:32] = vm.make_header(sec0 + t // fps, t % fps, 40, tfb, tids[l], bits).view(np.uint8)
            f[32:] = rng.integers(0, 256, size=tfb - 32, dtype=np.uint8)
            if s in (0, 1, 2) or 60 <= s < 66:   # no thread at these frame times
                continue
            if 0 <= s < nslots and rng.random() < 0.10:
                continue
            frames.append(f)
            if rng.random() < 0.03:               # a second copy with another payload
                g = f.copy()
                g[32:] = rng.integers(0, 256, size=tfb - 32, dtype=np.uint8)
                frames.append(g)
    for _ in range(20):                           # unmapped thread ids
        f = frames[int(rng.integers(len(frames)))].copy()
        w = vm.header_words(f)
        vm.set_threadid(w, 1023 if 1023 not in tids else 1022)
        f[:32] = w.view(np.uint8)
        frames.append(f)
    inbuf = np.concatenate([frames[i] for i in rng.permutation(len(frames))])
    ref_sec = vm.epoch_mjd(40) * 86400 + sec0
    got, nout, st, ref, ref_n, ref_st = _run_both(gpu, inbuf, tfb, tids, fps, bits, ref_sec, f0, 0, nslots)
    _say("  nout", nout, "dropped", st)
    assert ref_n == nslots - 3 and ref_st["early"] >= 4 * n and ref_st["late"] > 0 and ref_st["unmapped"] == 20
    assert nout == ref_n and st == ref_st
    np.testing.assert_array_equal(got, ref)


def test_consecutive_segments_and_errors(gpu):
    """Two calls over the same input advance processframenumber by nslots (the second call
    turns the next frame times); a frame whose length field differs raises, as the
    reference aborts there."""
    import torch
    from difx_amd.core import FxgError
    from difx_amd.vdifmux import VDIFMuxer
    _say("vdifmux consecutive segments")
    rng = np.random.default_rng(5)
    n, bits, fps, tfb = 8, 2, 10, 32 + 64
    tids = list(range(8, 0, -1))
    fr = _thread_frames(40, n, tfb, fps, tids, rng)
    inbuf = np.concatenate([f for row in fr for f in row])
    ref_sec = vm.epoch_mjd(40) * 86400 + 1000
    mux = VDIFMuxer(n, tfb, fps, bits, tids)
    mux.initialise(inbuf[:32])
    d_in = torch.from_numpy(inbuf).to(gpu)
    ofb = mux.outputframebytes
    outs = []
    for k in range(2):
        d_out = torch.full((20 * ofb,), FILL, dtype=torch.uint8, device=gpu)
        d_out, nbytes = mux.multiplex(d_in, 20, out=d_out)
        assert nbytes == 20 * ofb
        outs.append(d_out.cpu().numpy())
    for k in range(2):
        ref, ref_n, _ = vm.mux_segment(inbuf, tfb, tids, fps, bits, ref_sec, 0, 20 * k, 20)
        np.testing.assert_array_equal(outs[k], ref)
    # each call drops the other call's 20 frame times of n threads
    assert mux.dropped == {"early": 20 * n, "late": 20 * n, "unmapped": 0}
    bad = inbuf.copy()
    w = vm.header_words(bad[5 * tfb:])
    vm.set_framebytes(w, tfb + 8)
    bad[5 * tfb:5 * tfb + 32] = w.view(np.uint8)
    with pytest.raises(FxgError):
        mux.multiplex(torch.from_numpy(bad).to(gpu), 20, start=0)


def test_full_size_station_segment(gpu):
    """One station's 0.5 s of 8 x 32 MHz dual-polarisation 2-bit data recorded as 16 VDIF
    threads of 8000-byte payloads (1000 frame times, 128 MB), shuffled, 1% of frames
    missing: bit-exact with the oracle, and every complete output frame equals the frame the
    threads were split from."""
    _say("vdifmux full-size station segment")
    rng = np.random.default_rng(2026)
    n, bits, fps, tfb, nslots = 16, 2, 2000, 8032, 1000
    tids = list(range(n))
    hdr = np.stack([vm.make_header(3000 + s // fps, s % fps, 40, tfb, tids[l], bits)
                    for s in range(nslots) for l in range(n)])
    inbuf = rng.integers(0, 256, size=(nslots * n, tfb), dtype=np.uint8)
    inbuf[:, :32] = hdr.view(np.uint8).reshape(-1, 32)
    keep = rng.random(nslots * n) >= 0.01
    keep[:n] = True
    order = rng.permutation(np.flatnonzero(keep))
    shuffled = inbuf[order].reshape(-1)
    ref_sec = vm.epoch_mjd(40) * 86400 + 3000
    got, nout, st, ref, ref_n, ref_st = _run_both(gpu, shuffled, tfb, tids, fps, bits, ref_sec, 0, 0, nslots)
    assert nout == ref_n == nslots and st == ref_st
    np.testing.assert_array_equal(got, ref)
    complete = keep.reshape(nslots, n).all(axis=1)
    ofb = (tfb - 32) * n + 32
    s = int(np.flatnonzero(complete)[len(np.flatnonzero(complete)) // 2])
    tw = np.stack([np.frombuffer(inbuf[s * n + l, 32:].tobytes(), dtype="<u4") for l in range(n)])
    np.testing.assert_array_equal(got[s * ofb + 32:(s + 1) * ofb],
                                  vm.cornerturn_generic(tw, bits).astype("<u4").view(np.uint8))


def test_muxed_stream_correlates_like_the_recording(gpu):
    """A 3-station 8 x 32 MHz dual-polarisation slot, each station's 16-band stream split into
    16 VDIF threads and shuffled, muxed on the device and correlated from the device buffers
    (fxg_process_subint_device): the results equal those of the original streams."""
    import torch
    from difx_amd import config as cf
    from difx_amd import synth
    from difx_amd.core import FxCore
    from difx_amd.vdifmux import VDIFMuxer
    _say("vdifmux -> F/X engine")
    freqs = [8400.0 + 32.0 * i for i in range(8)]
    cfg = cf.standard_config(3, 2, freqs, 32.0, 256, subint_ns=32 * 8000, payloadbytes=16 * 512, seed=31)
    data, ctrl = synth.noise_slot(cfg, [0, 1, 0], seed=32)
    core = FxCore(cfg)
    want = core.processdata([0, 1, 0], data, ctrl)
    rng = np.random.default_rng(33)
    d_muxed, nbytes = [], []
    for d, ds in enumerate(cfg.datastreams):
        fb = ds.framebytes
        tids = [int(x) for x in rng.permutation(1024)[:16]]
        parts = vm.split_threads(data[d], fb, 16, 2, tids)
        tfb = (fb - 32) // 16 + 32
        frames = [p[i * tfb:(i + 1) * tfb] for p in parts for i in range(len(p) // tfb)]
        inbuf = np.concatenate([frames[i] for i in rng.permutation(len(frames))])
        mux = VDIFMuxer(16, tfb, 1000, 2, tids)
        mux.initialise(data[d][:32])
        nfr = len(data[d]) // fb
        out, nb = mux.multiplex(torch.from_numpy(inbuf).to(gpu), nfr)
        assert nb == len(data[d])
        np.testing.assert_array_equal(out.cpu().numpy(), data[d])
        d_muxed.append(out)
        nbytes.append(nb)
    d_res = torch.zeros(2 * core.coreresultlength, dtype=torch.float32, device=gpu)
    core.processdata_device([0, 1, 0], [t.data_ptr() for t in d_muxed], nbytes, ctrl, d_res.data_ptr())
    torch.cuda.synchronize()
    got = d_res.cpu().numpy().view(np.complex64)
    np.testing.assert_array_equal(got, want)


def test_long_gap(gpu):
    """A gap of 600 frame times with no thread (longer than one workgroup's look-back pass)
    between two stretches of data, fps 7 so that the advanced headers cross many seconds."""
    _say("vdifmux long gap")
    rng = np.random.default_rng(91)
    n, bits, fps, tfb, nslots = 4, 2, 7, 32 + 64, 700
    tids = [3, 1, 2, 0]
    fr = _thread_frames(nslots, n, tfb, fps, tids, rng, sec0=100)
    frames = [f for s, row in enumerate(fr) for f in row if not (50 <= s < 650)]
    inbuf = np.concatenate([frames[i] for i in rng.permutation(len(frames))])
    ref_sec = vm.epoch_mjd(40) * 86400 + 100
    got, nout, st, ref, ref_n, ref_st = _run_both(gpu, inbuf, tfb, tids, fps, bits, ref_sec, 0, 0, nslots)
    assert nout == ref_n == nslots and st == ref_st
    np.testing.assert_array_equal(got, ref)
