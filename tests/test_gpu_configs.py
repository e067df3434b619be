"""The benchmark configurations as written, against the oracle on the GPU, through the C ABI:
the uneven-band configuration in one configuration (every FFT length 4096-32768 and six
zooms, both F-engine launch forms, FP32 and FP16); the 16-station MFMA target shape (a full
32-row tile) in FP32, FP16 and BF16; the 64-station complex 8-bit configuration at its full
length of 3906 FFTs, per channel against the FP32 X-engine and, on a subset of baselines,
against the FP64 oracle; and a serving stream shaped like the Poisson benchmark (4-32
stations, 2-16 x 32 MHz, nchan 256-4096) in which every job's integration equals the
oracle's sum of its subints."""
import sys

import numpy as np
import pytest

import fxoracle as fo
from difx_amd import synth
from fxcases import case_cfg2, case_cfg3_full, case_target16, channel_errors, record_errors, slot

pytestmark = pytest.mark.gpu

BOUND = {"fp32": 1e-5, "fp16": 1e-3, "bf16": 1e-2}


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_cfg3_full_band_set(gpu):
    from difx_amd.core import FxCore
    offsets = (0, 1, 0)
    base = case_cfg3_full(B=2)
    data, ctrl = slot(base, offsets, seed=51)
    _say("cfg3 full: oracle")
    ref, _ = fo.process_subint(base, list(offsets), data, ctrl)
    lay = fo.result_layout(base)
    for xe in ("fp32", "fp16"):
        outs = {}
        for mode in (1, 2):
            _say("cfg3 full:", xe, "F-engine form", mode)
            cfg = case_cfg3_full(B=2, xengine=xe, fengine_mode=mode)
            with FxCore(cfg) as core:
                assert core.coreresultlength == lay.coreresultlength
                outs[mode] = core.processdata(offsets, data, ctrl)
            e = record_errors(cfg, lay, outs[mode].astype(np.complex128), ref)
            _say(f"cfg3 full {xe} form {mode}: {len(e['cross'])} cross records, max", max(e["cross"]),
                 "auto max", max(e["auto"]))
            assert max(e["cross"]) < BOUND[xe] and max(e["auto"]) < BOUND[xe]
        np.testing.assert_array_equal(outs[1].view(np.uint64), outs[2].view(np.uint64))


def test_target16_all_precisions(gpu):
    from difx_amd.core import FxCore
    offsets = (0, 1, 0)
    base = case_target16(B=32)
    data, ctrl = slot(base, offsets, seed=52)
    _say("16 stations: oracle")
    ref, _ = fo.process_subint(base, list(offsets), data, ctrl)
    lay = fo.result_layout(base)
    for xe in ("fp32", "fp16", "bf16"):
        with FxCore(case_target16(B=32, xengine=xe)) as core:
            got = core.processdata(offsets, data, ctrl)
        e = record_errors(base, lay, got.astype(np.complex128), ref)
        _say(f"16 stations {xe}: max", max(e["cross"]), "median", float(np.median(e["cross"])),
             "auto max", max(e["auto"]))
        assert max(e["cross"]) < BOUND[xe] and max(e["auto"]) < BOUND[xe]


def _device_cfg2_slot(cfg, offsets, torch, dev, seed):
    """Complex 8-bit VDIF slots generated on the device: Gaussian samples of sigma 20 counts
    about 128, valid headers."""
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    bufs, nbytes, ctrl = [], [], []
    for d, ds in enumerate(cfg.datastreams):
        start_ns = synth.buffer_start_ns(cfg, d, offsets)
        nb = synth.slot_bytes(cfg, d, offsets, start_ns)
        spf = synth.samples_per_frame(ds)
        frame_ns = int(round(spf * cfg.sampletime_us(d) * 1000.0))
        nframes = nb // ds.framebytes
        x = torch.randn((nframes, ds.framebytes), generator=g, device=dev)
        buf = torch.clamp(torch.round(x * 20.0 + 128.0), 0, 255).to(torch.uint8)
        first = start_ns // frame_ns
        hdr = np.stack([synth.vdif_header(0, int(first) + f, ds.framebytes, len(ds.recbands), ds.nbits, True)
                        for f in range(nframes)])
        buf[:, : ds.headerbytes] = torch.from_numpy(hdr).to(dev)
        bufs.append(buf.reshape(-1).contiguous())
        nbytes.append(nframes * ds.framebytes)
        ctrl.append(synth.control_words(cfg, d, offsets, start_ns))
    return bufs, nbytes, ctrl


def test_cfg2_full_length_precision(gpu):
    """64 stations, 16 MHz complex 8-bit, nchan 4096, 3906 FFTs (subint 999 936 000 ns): the
    FP16 and BF16 X-engines per channel against the FP32 X-engine on every baseline, and all
    three against the FP64 oracle on the baselines of four stations."""
    import torch
    from difx_amd.core import FxCore
    offsets = [0, 1, 0]
    mk = lambda xe: case_cfg2(B=3906, xengine=xe, max_delay_us=300.0)
    cfg = mk("fp32")
    assert cfg.blockspersend() == 3906
    bufs, nbytes, ctrl = _device_cfg2_slot(cfg, offsets, torch, gpu, 53)
    ptrs = [b.data_ptr() for b in bufs]
    lay = fo.result_layout(cfg)
    res = {}
    for xe in ("fp32", "fp16", "bf16"):
        _say("cfg2 full length:", xe)
        with FxCore(mk(xe)) as core:
            out = torch.zeros(2 * core.coreresultlength, dtype=torch.float32, device=gpu)
            core.processdata_device(offsets, ptrs, nbytes, ctrl, out.data_ptr(), 0)
            torch.cuda.synchronize()
            res[xe] = out.cpu().numpy().view(np.complex64).astype(np.complex128)
    for xe in ("fp16", "bf16"):
        norm, dphi = channel_errors(cfg, lay, res[xe], res["fp32"])
        e = record_errors(cfg, lay, res[xe], res["fp32"])
        _say(f"cfg2 full {xe} vs FP32 X-engine: |dV|/sqrt(AiAj) median {np.median(norm):.3e} max {np.max(norm):.3e}, "
             f"dphase deg median {np.median(dphi):.4f} max {np.max(dphi):.4f}, record max {max(e['cross']):.3e}")
        assert max(e["cross"]) < BOUND[xe] and np.max(norm) < BOUND[xe]
    # the FP64 oracle on the baselines of stations 0-3
    host = [b.cpu().numpy() for b in bufs[:4]]
    st = []
    for d in range(4):
        _say("cfg2 full length: oracle station", d)
        st.append(fo.process_station(cfg, d, offsets, host[d], ctrl[d]))
    for xe in ("fp32", "fp16", "bf16"):
        worst = 0.0
        for j, bl in enumerate(cfg.baselines):
            if bl.ds2 >= 4:
                continue
            v = np.einsum("tc,tc->c", st[bl.ds1].spectra[0], np.conj(st[bl.ds2].spectra[0]))
            o = lay.baseline_off[(0, j)]
            g = res[xe][o: o + len(v)]
            a1 = np.einsum("tc,tc->c", st[bl.ds1].spectra[0], np.conj(st[bl.ds1].spectra[0])).real
            a2 = np.einsum("tc,tc->c", st[bl.ds2].spectra[0], np.conj(st[bl.ds2].spectra[0])).real
            worst = max(worst, float(np.max(np.abs(g - v) / np.sqrt(a1 * a2))))
            rel = float(np.mean(np.abs(g - v)) / np.mean(np.abs(v)))
            assert rel < BOUND[xe], (xe, j, rel)
        _say(f"cfg2 full {xe} vs FP64 oracle, stations 0-3: max |dV|/sqrt(AiAj) {worst:.3e}")
        assert worst < BOUND[xe]


def _serving_jobs(rng, njobs):
    """Jobs shaped like the Poisson serving benchmark: station count from {4, 8, 16, 32},
    bandwidth m x 32 MHz (m/2 freqs x 2 pols, m from {2, 4, 8, 16}), nchan from
    {256, 1024, 4096}, FP16 X-engine, subints of 4 FFTs."""
    from difx_amd import config as cf
    jobs = []
    shapes = [(32, 16, 4096), (4, 2, 256), (16, 8, 1024), (8, 4, 4096), (32, 2, 256), (4, 16, 1024)][:njobs]
    for n, m, nchan in shapes:
        freqs = [4800.0 + 40.0 * k for k in range(m // 2)]
        cfg = cf.standard_config(n, 2, freqs, 32.0, nchan, subint_ns=4 * nchan * 1000 // 32, xengine="fp16",
                                 seed=int(rng.integers(1 << 30)), max_delay_us=200.0)
        jobs.append(((n, m, nchan), cfg))
    return jobs


def test_serving_stream_against_oracle(gpu):
    from difx_amd.server import FxServer
    import torch
    rng = np.random.default_rng(54)
    jobs = _serving_jobs(rng, 6)
    nsub = 2
    subs = []
    for i, (shape, cfg) in enumerate(jobs):
        s = []
        for k in range(nsub):
            offsets = (0, 1 + k, 0)
            data, ctrl = synth.noise_slot(cfg, list(offsets), seed=1000 * i + k)
            s.append((offsets, [torch.from_numpy(d).to(gpu) for d in data], [len(d) for d in data], ctrl, data))
        subs.append(s)
    with FxServer() as srv:
        ids = [srv.add_job(cfg, nsub, data_on_device=True) for _, cfg in jobs]
        for k in range(nsub):
            for i in range(len(jobs)):
                offsets, dev, nb, ctrl, _ = subs[i][k]
                srv.push(ids[i], offsets, [t.data_ptr() for t in dev], nb, ctrl)
        got = [srv.wait_job(j) for j in ids]
        stats = srv.stats()
    _say("serving stream stats", stats)
    assert stats["jobs_finished"] == len(jobs) and stats["max_active"] >= 2
    for i, (shape, cfg) in enumerate(jobs):
        ref = None
        for offsets, _, _, ctrl, data in subs[i]:
            r, _ = fo.process_subint(cfg, list(offsets), data, ctrl)
            ref = r if ref is None else ref + r
        e = record_errors(cfg, fo.result_layout(cfg), got[i][0].astype(np.complex128), ref)
        _say(f"serving job {shape} (stations, m, nchan): cross max {max(e['cross']):.3e} auto max {max(e['auto']):.3e}, "
             f"p50/p99 {got[i][1]:.2f}/{got[i][2]:.2f} ms")
        assert max(e["cross"]) < 1e-3 and max(e["auto"]) < 1e-3
