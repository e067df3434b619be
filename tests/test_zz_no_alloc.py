"""The allocation-free VDIF corner turn (fxg_vdif_mux_device_static, workspace in the library's
static device array) against fxg_vdif_mux_device with a caller workspace, through the C ABI:
output bytes and outcome bit-exact on two calls in a row, the second with a smaller segment
(a damaged 16-thread segment, then a shorter 4-thread one), and a segment beyond the static
workspace refused on the host."""
import ctypes as C
import sys

import numpy as np
import pytest

import vdifmux as vm

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _segment(rng, nslots, n, tfb, fps, tids, bits=2, damage=True):
    frames = []
    for s in range(nslots):
        for l in range(n):
            f = np.empty(tfb, dtype=np.uint8)
            f[:32] = vm.make_header(1000 + s // fps, s % fps, 40, tfb, tids[l], bits).view(np.uint8)
            f[32:] = rng.integers(0, 256, size=tfb - 32, dtype=np.uint8)
            frames.append(f)
    if damage:                                   # missing, duplicated and unmapped frames
        keep = [i for i in range(len(frames)) if rng.random() > 0.1]
        frames = [frames[i] for i in keep] + [frames[keep[3]].copy()]
        bad = frames[5].copy()
        bad[:32] = vm.make_header(1000, 5 % fps, 40, tfb, 999, bits).view(np.uint8)
        frames.append(bad)
    order = rng.permutation(len(frames))
    return np.concatenate([frames[i] for i in order])


def _params(n, tfb, fps, bits, tids, start):
    from difx_amd import fxg_abi as abi
    p = abi.fxg_vdif_mux_params()
    p.inputframebytes, p.nthreads, p.bitspersample, p.framespersecond = tfb, n, bits, fps
    for j, t in enumerate(tids):
        p.threadmap[j] = t
    p.ref_seconds = vm.epoch_mjd(40) * 86400 + 1000
    p.ref_frame = 0
    p.start = start
    return p


def _run(lib, gpu, inbuf, nslots, n, tfb, fps, bits, tids, start):
    """(output of the workspace form, its outcome, output of the static form, its outcome)"""
    import torch
    ofb = (tfb - 32) * n + 32
    p = _params(n, tfb, fps, bits, tids, start)
    d_in = torch.from_numpy(inbuf).to(gpu)
    st = torch.cuda.current_stream(gpu)
    sh = C.c_void_p(st.cuda_stream)
    out_a = torch.full((nslots * ofb,), 0x5A, dtype=torch.uint8, device=gpu)
    wb = int(lib.fxg_vdif_mux_workspace_bytes(nslots, n))
    work = torch.empty(wb, dtype=torch.uint8, device=gpu)
    assert lib.fxg_vdif_mux_device(C.byref(p), C.c_void_p(d_in.data_ptr()), d_in.numel(),
                                   C.c_void_p(out_a.data_ptr()), out_a.numel(), nslots,
                                   C.c_void_p(work.data_ptr()), wb, sh) == 0
    nout = C.c_int64()
    dr = (C.c_int64 * 3)()
    rc = lib.fxg_vdif_mux_result(C.c_void_p(work.data_ptr()), sh, C.byref(nout), dr)
    out_b = torch.full((nslots * ofb,), 0x5A, dtype=torch.uint8, device=gpu)
    outcome = torch.full((5,), -7, dtype=torch.int64, device=gpu)
    assert lib.fxg_vdif_mux_device_static(C.byref(p), C.c_void_p(d_in.data_ptr()), d_in.numel(),
                                          C.c_void_p(out_b.data_ptr()), out_b.numel(), nslots,
                                          C.c_void_p(outcome.data_ptr()), sh) == 0
    torch.cuda.synchronize()
    want = [nout.value, dr[0], dr[1], dr[2], 0 if rc == 0 else 1]
    return out_a.cpu().numpy(), want, out_b.cpu().numpy(), outcome.cpu().numpy().tolist()


def test_static_workspace_equals_caller_workspace(gpu):
    from difx_amd.core import load_library
    lib = load_library()
    rng = np.random.default_rng(77)
    # first call: a damaged 16-thread 2-bit segment of 60 frame times
    n, tfb, fps = 16, 32 + 256, 25
    tids = [int(x) for x in rng.permutation(64)[:n]]
    seg = _segment(rng, 60, n, tfb, fps, tids)
    _say("static corner turn: 16 threads x 60 slots")
    a, wa, b, wb = _run(lib, gpu, seg, 60, n, tfb, fps, 2, tids, 0)
    assert wa == wb and wa[0] == 60 and wa[3] == 1       # one frame of an unmapped thread
    np.testing.assert_array_equal(a, b)
    # second call, smaller: 4 threads, 12 frame times from frame time 3, no damage (stale
    # contents of the shared array from the first call must not leak into it)
    n2, tfb2 = 4, 32 + 128
    tids2 = [int(x) for x in rng.permutation(64)[:n2]]
    seg2 = _segment(rng, 15, n2, tfb2, fps, tids2, damage=False)
    _say("static corner turn: 4 threads x 12 slots")
    a, wa, b, wb = _run(lib, gpu, seg2, 12, n2, tfb2, fps, 2, tids2, 3)
    assert wa == wb and wa[0] == 12 and wa[1] == 12      # 3 frame times x 4 threads before the segment
    np.testing.assert_array_equal(a, b)


def test_static_workspace_limit_refused(gpu):
    """A segment whose workspace exceeds the static array is refused before any launch."""
    from difx_amd import fxg_abi as abi
    from difx_amd.core import load_library
    import torch
    lib = load_library()
    n, tfb = 32, 32 + 64
    limit = abi.FXG_MUX_STATIC_WORK_BYTES
    nslots = (limit - 64) // (4 * (n + 1)) + 1
    assert lib.fxg_vdif_mux_workspace_bytes(nslots, n) > limit
    p = _params(n, tfb, 25, 1, list(range(n)), 0)
    out = torch.empty(nslots * ((tfb - 32) * n + 32), dtype=torch.uint8, device=gpu)   # a valid output
    outcome = torch.zeros(5, dtype=torch.int64, device=gpu)
    rc = lib.fxg_vdif_mux_device_static(C.byref(p), C.c_void_p(0), 0, C.c_void_p(out.data_ptr()), out.numel(),
                                        nslots, C.c_void_p(outcome.data_ptr()), C.c_void_p(0))
    assert rc == abi.FXG_E_INVALID
    # one slot fewer fits, and runs: one frame of thread 0 at frame time 2
    f = np.zeros(tfb, dtype=np.uint8)
    f[:32] = vm.make_header(1000, 2, 40, tfb, 0, 1).view(np.uint8)
    d_in = torch.from_numpy(f).to(gpu)
    rc = lib.fxg_vdif_mux_device_static(C.byref(p), C.c_void_p(d_in.data_ptr()), d_in.numel(),
                                        C.c_void_p(out.data_ptr()), out.numel(), nslots - 1,
                                        C.c_void_p(outcome.data_ptr()), C.c_void_p(0))
    torch.cuda.synchronize()
    assert rc == abi.FXG_OK
    assert outcome.cpu().tolist() == [nslots - 1 - 2, 0, 0, 0, 0]   # frames from the first present slot
