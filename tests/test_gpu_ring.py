"""The buffer-start choice feeding the correlator: each station's continuous VDIF stream sits
in a 4-segment read ring; fxg_control_params (with the VDIF frame step) picks the subint's
start byte and control words, as DataStream does; the F/X path on those bytes with those
control words equals the FP64 oracle on the same inputs (diffDiFX metric <= 1e-5), and every
FFT block of the subint is valid (the chosen start precedes the first sample needed)."""
import ctypes as C
import sys

import numpy as np
import pytest

import fxoracle as fo
from difx_amd import synth
from fxcases import record_errors

pytestmark = pytest.mark.gpu


def test_ring_start_choice_correlates(gpu):
    from difx_amd import config as cf
    from difx_amd import fxg_abi as abi
    from difx_amd.core import FxCore, load_library
    print("ring start choice -> F/X", flush=True)
    lib = load_library()
    B = 32
    cfg = cf.standard_config(3, 2, [8400.0, 8432.0], 32.0, 256, subint_ns=B * 8000, max_delay_us=200.0,
                             max_rate_us_per_s=0.5, seed=41)
    fa, keep = cfg.to_abi()
    nseg, frames_per_seg = 4, 8
    rng = np.random.default_rng(42)
    offsets = [0, 0, 1_500_000]
    data, ctrl = [], []
    for d, ds in enumerate(cfg.datastreams):
        fb, pay = ds.framebytes, ds.framebytes - ds.headerbytes
        R = frames_per_seg * fb
        nfr = nseg * frames_per_seg
        codes = synth.random_2bit_codes(rng, nfr * pay * 4, ds.format)
        ring = synth.frame_payloads(synth.pack_2bit(codes.reshape(-1, 1)), ds, 0)
        assert ring.size == nseg * R
        db = abi.fxg_databuffer()
        assert lib.fxg_buffer_params(C.byref(fa), d, nseg, nseg * R, C.byref(db)) == 0
        segs = (abi.fxg_segment * nseg)()
        for k in range(nseg):
            assert lib.fxg_segment_params(C.byref(fa), d, nseg, nseg * R, C.byref(segs[k])) == 0
        frame_ns = 1e9 / db.framespersecond
        for k in range(nseg):
            t = int(round(k * frames_per_seg * frame_ns))
            segs[k].scan, segs[k].scanseconds, segs[k].scanns = 0, t // 1_000_000_000, t % 1_000_000_000
            segs[k].validbytes, segs[k].readto = R, 1
        assert segs[0].nsinc == int(round(frames_per_seg * frame_ns))
        db.seg = segs
        db.keepreading, db.lastvalidsegment = 0, nseg - 1
        t0 = offsets[1] + offsets[2] / 1e9
        span = segs[0].numchannels * segs[0].blockspersend * 2 * segs[0].sampletimens
        d1 = synth.eval_delay(ds.model, ds.model_order, t0, cfg.model_inc_secs)
        d2 = synth.eval_delay(ds.model, ds.model_order, t0 + span / 1e9, cfg.model_inc_secs)
        at = C.c_int32(0)
        cw = (C.c_int32 * (3 + (B + 29) // 30))()
        bi = C.c_int32()
        assert lib.fxg_control_params(C.byref(db), C.byref(at), 0, offsets[1], offsets[2], 1, d1, d2,
                                      abi.FXG_ADVANCE_FN(0), None, cw, C.byref(bi)) == 0
        c = np.array(list(cw), dtype=np.int32)
        print(f"  ds {d}: delay {d1:.3f} us, segment {at.value}, byte {bi.value}, start {c[1]} s {c[2]} ns",
              flush=True)
        assert c[1] != abi.FXG_INVALID_SUBINT and bi.value % fb == 0
        assert all(int(w) == (1 << 30) - 1 for w in c[3:3 + B // 30])
        # DataStream sends sendbytes from the chosen byte (datastream.cpp:294-314)
        data.append(ring[bi.value:bi.value + segs[at.value].sendbytes])
        assert data[-1].size == segs[at.value].sendbytes
        ctrl.append(c)
    core = FxCore(cfg)
    got = core.processdata(offsets, data, ctrl)
    ref, stations = fo.process_subint(cfg, offsets, data, ctrl)
    for d in range(len(cfg.datastreams)):
        assert np.all(stations[d].weight > 0.999), (d, stations[d].weight)
        np.testing.assert_allclose(core.getDataWeights(d), stations[d].weight, rtol=0, atol=1e-7)
    lay = fo.result_layout(cfg)
    e = record_errors(cfg, lay, got.astype(np.complex128), ref)
    print("  cross max", max(e["cross"]), "auto max", max(e["auto"]), flush=True)
    assert max(e["cross"]) < 1e-5 and max(e["auto"]) < 1e-5
    core.close()
