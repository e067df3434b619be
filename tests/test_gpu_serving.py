"""Serving (fxg_server_*): concurrent jobs with different station counts and bandwidths on one
device, admitted by device footprint, integrated on the device.

The bar: a job's integrated results are bit-identical whether it runs alone or in a mix
(its kernels touch only its own buffers, in the same order), and equal the FP32 sum, in
subint order, of the per-subint results of the single-job path."""
import sys

import numpy as np
import pytest

from difx_amd import synth
from fxcases import case_cfg0, case_cfg1, case_cfg2, case_cfg3

pytestmark = pytest.mark.gpu

NSUB = 3


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _jobs():
    return {
        "cfg0": case_cfg0(B=64),
        "cfg1_fp16": case_cfg1(B=32, nstations=6, nfreq=2, xengine="fp16"),
        "cfg2_bf16": case_cfg2(B=8, nstations=12, xengine="bf16"),
        "cfg3_fp32": case_cfg3(B=32),
    }


def _slots(cfg, gpu, seed):
    """NSUB subints (offsets, device tensors, nbytes, control) of one job."""
    import torch
    out = []
    for k in range(NSUB):
        offsets = (0, 1 + k, 0)
        data, ctrl = synth.noise_slot(cfg, list(offsets), seed=seed + k)
        dev = [torch.from_numpy(d).to(gpu) for d in data]
        out.append((offsets, dev, [len(d) for d in data], ctrl, data))
    return out


def _run(server, jobs, subs, order, data_on_device=True):
    ids = {}
    for name in order:
        ids[name] = server.add_job(jobs[name], NSUB, data_on_device=data_on_device)
    for k in range(NSUB):                      # interleave the jobs' subints as they "arrive"
        for name in order:
            offsets, dev, nb, ctrl, host = subs[name][k]
            ptrs = [t.data_ptr() for t in dev] if data_on_device else [h.ctypes.data for h in host]
            server.push(ids[name], offsets, ptrs, nb, ctrl)
    return {name: server.wait_job(ids[name]) for name in order}


def test_alone_vs_mix_bit_identical(gpu):
    from difx_amd.core import FxCore
    from difx_amd.server import FxServer
    jobs = _jobs()
    subs = {name: _slots(cfg, gpu, 100 * i) for i, (name, cfg) in enumerate(jobs.items())}
    alone = {}
    for name in jobs:
        _say("alone", name)
        with FxServer() as srv:
            alone[name] = _run(srv, jobs, subs, [name])[name][0]
    _say("mix")
    with FxServer() as srv:
        mix = _run(srv, jobs, subs, list(jobs))
        st = srv.stats()
    _say("mix stats", st)
    assert st["jobs_finished"] == len(jobs) and st["max_active"] >= 2
    for name in jobs:
        np.testing.assert_array_equal(mix[name][0].view(np.uint64), alone[name].view(np.uint64))
        assert mix[name][1] > 0 and mix[name][2] >= mix[name][1]
    # the integration equals the FP32 running sum of the single-job path's subint results
    for name in ("cfg0", "cfg1_fp16"):
        core = FxCore(jobs[name])
        acc = np.zeros(core.coreresultlength, np.complex64)
        for offsets, dev, nb, ctrl, host in subs[name]:
            acc = acc + core.processdata(offsets, host, ctrl)
        core.close()
        np.testing.assert_array_equal(acc.view(np.uint64), alone[name].view(np.uint64))


def test_admission_by_footprint(gpu):
    """A budget that holds one job at a time serialises the jobs; host-data jobs (input slots
    counted in the footprint) give the same results as device-data jobs."""
    from difx_amd.server import FxServer
    jobs = _jobs()
    names = ["cfg0", "cfg1_fp16"]
    subs = {name: _slots(jobs[name], gpu, 7 + i) for i, name in enumerate(names)}
    from difx_amd.core import footprint
    fps = []
    for n in names:                    # server footprint: context + 3 result buffers + 2 input slots
        dev, crl, slot = footprint(jobs[n])
        fps.append(dev + 3 * 8 * crl + 2 * slot)
    budget = max(fps) + (1 << 20)
    assert budget < sum(fps)
    with FxServer(hbm_budget_bytes=budget) as srv:
        one = _run(srv, jobs, subs, names, data_on_device=False)
        st = srv.stats()
    _say("budget-limited stats", st)
    assert st["max_active"] == 1 and st["jobs_finished"] == 2
    with FxServer() as srv:
        many = _run(srv, jobs, subs, names, data_on_device=True)
    for n in names:
        np.testing.assert_array_equal(one[n][0].view(np.uint64), many[n][0].view(np.uint64))


def test_footprint_is_what_create_holds(gpu):
    import ctypes as C
    from difx_amd.core import FxCore, footprint
    for cfg in _jobs().values():
        dev, crl, _ = footprint(cfg)
        core = FxCore(cfg)
        held = C.c_int64()
        core.lib.fxg_device_bytes(core.h, C.byref(held))
        core.close()
        assert held.value == dev


def test_serving_mix_against_oracle(gpu):
    """Jobs of the serving stream's shapes (4, 8, 16 and 32 stations dual-pol, 2-bit VDIF,
    1-2 x 32 MHz, nchan 256-4096, FP16 MFMA X-engine, one FP32 job) admitted together and
    interleaved: each job's device integration over its subints against the sum of the FP64
    oracle's subint results, within the bar of its X-engine (1e-3 FP16, 1e-5 FP32)."""
    import fxoracle as fo
    from fxcases import record_errors
    from difx_amd import config as cf
    from difx_amd.server import FxServer
    jobs = {
        "n4_fp32": cf.standard_config(4, 2, [4800.0, 4840.0], 32.0, 256, subint_ns=16 * 8000, xengine="fp32", seed=41),
        "n8": cf.standard_config(8, 2, [4800.0], 32.0, 1024, subint_ns=8 * 32000, xengine="fp16", seed=42),
        "n16": cf.standard_config(16, 2, [4800.0, 4840.0], 32.0, 256, subint_ns=16 * 8000, xengine="fp16", seed=43),
        "n32": cf.standard_config(32, 2, [4800.0], 32.0, 4096, subint_ns=4 * 128000, xengine="fp16", seed=44),
    }
    subs = {name: _slots(cfg, gpu, 300 + 10 * i) for i, (name, cfg) in enumerate(jobs.items())}
    with FxServer() as srv:
        mix = _run(srv, jobs, subs, list(jobs))
        st = srv.stats()
    _say("mix stats", st)
    assert st["jobs_finished"] == len(jobs) and st["max_active"] >= 2
    for name, cfg in jobs.items():
        lay = fo.result_layout(cfg)
        ref = np.zeros(lay.coreresultlength, np.complex128)
        for offsets, dev, nb, ctrl, host in subs[name]:
            r, _ = fo.process_subint(cfg, list(offsets), host, ctrl)
            ref += r
        e = record_errors(cfg, lay, mix[name][0].astype(np.complex128), ref)
        bound = 1e-5 if cfg.xengine == "fp32" else 1e-3
        _say(f"serving {name} vs oracle: cross max {max(e['cross']):.3e}, auto max {max(e['auto']):.3e}")
        assert max(e["cross"]) < bound and max(e["auto"]) < bound
