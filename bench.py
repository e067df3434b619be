#!/usr/bin/env python3
"""Benchmark of the FX hot path on one MI355X: input Gsamples/s, p99 subint latency, and the
MFMA/batched path against the plain FP32 path.

Workload of `value` (one "step" = one subintegration through the whole per-subint pipeline):
16 stations, dual circular polarisation, 8 subbands x 32 MHz, 2-bit real VDIF with 16 bands
per 32-bit sample word (8000-byte payloads), nchan 1024 (2048-point FFTs), 0.5 s subints
(15625 FFT blocks per station), fringe rotation order 1, seeded quintic delay models.
Synthetic 2-bit data is generated on the device before the timed region (inputs resident
in HBM), so `value` excludes PCIe transfers.

Paths timed on the same inputs, with the same number of steps:
  MFMA      fused F-engine + FP16 MFMA Hermitian X-engine   -> `value`
  plain     fused F-engine (FP32 spectra) + FP32 CMAC X-engine (8 x 8 blocks of each freq's
            input x input matrix per channel, FP32 FMA), with the X-engine's VALU issue
            fraction (plain_fp32.x_engine_valu_issue)
            -> `plain_fp32_value`, and `speedup_mfma_vs_plain` = value / plain
  BF16      the MFMA path with BF16 inputs (section mfma_bf16)

The roofline object describes the dominant kernel (the F-engine) against HBM, timed with HIP
events on its own stream inside the library (fxg_set_timing); `issue_roofline` states the
same kernel against its VALU issue rate, with its per-sample instruction count from the
committed counter profile.  The CPU baseline runs the C restatement in oracle/ (when built)
on a bounded sample of the same workload on the host, with all host threads and with one.

Further sections of the same JSON line (each can be skipped with its --no-* flag):
  cfg1_10_stations    the 10-station array-scale configuration, MFMA and plain
  mfma_bf16           BF16 X-engine inputs on the 16-station workload
  mfma_post_f         the 16-station workload with post-F fringe rotation (order 0), where
                      the F-engine puts the two pols of a freq through one complex FFT
  uneven_bands        mixed bandwidths and zooms, both F-engine launch forms, autotuner report
  serving             Poisson job stream through the server (p50 / p99 subint latency), also
                      under an HBM budget of a third of its peak footprint
  kurtosis            the 10-station subint with the spectral-kurtosis sums enabled
  phase_cal           the 10-station subint with phase-cal extraction on every band (1 MHz comb)
  output_side         device scaling (writedata) and DIFX records of one subint's results
  vdif_corner_turn    10 stations' 16-thread VDIF segments corner-turned on the device, with
                      a device-copy yardstick and the numpy restatement's rate

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-plain] [--no-uneven]
                    [--no-serving] [--no-cfg1] [--no-bf16] [--no-postf] [--no-kurtosis]
                    [--no-pcal] [--no-output] [--no-mux]
    python bench.py --cpu-table      # the C restatement per configuration, 1 and all threads
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def workload_config(xengine: str, nstations: int = 16, **kw):
    """The array workload: nstations dual-pol, 8 x 32 MHz 2-bit VDIF, nchan 1024, 0.5 s."""
    from difx_amd import config as cf
    freqs = [4800.0 + 40.0 * k for k in range(8)]
    return cf.standard_config(nstations, 2, freqs, 32.0, 1024, subint_ns=500_000_000,
                              xengine=xengine, seed=20261202 if nstations == 10 else 20261205, **kw)


def device_slot(cfg, offsets, torch, dev, seed=7):
    """VDIF buffers generated on the device: random 2-bit payloads and valid headers."""
    from difx_amd import synth
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    bufs, ctrl, nbytes = [], [], []
    for d, ds in enumerate(cfg.datastreams):
        start_ns = synth.buffer_start_ns(cfg, d, offsets)
        nb = synth.data_bytes(cfg, d)
        nframes = nb // ds.framebytes
        buf = torch.randint(0, 256, (nframes, ds.framebytes), dtype=torch.uint8, device=dev, generator=g)
        hdr = np.stack([synth.vdif_header(0, f, ds.framebytes, len(ds.recbands), ds.nbits, False)
                        for f in range(nframes)])
        buf[:, : ds.headerbytes] = torch.from_numpy(hdr).to(dev)
        bufs.append(buf.reshape(-1).contiguous())
        nbytes.append(nframes * ds.framebytes)
        ctrl.append(synth.control_words(cfg, d, offsets, start_ns))
    return bufs, nbytes, ctrl


def uneven_config(xengine: str, nstations: int = 16, subint_ns: int = 512_000_000, **kw):
    """Uneven bands: per station a 4 x 64 MHz (nchan 16384, 32768-point FFTs), a 4 x 32 MHz
    (nchan 8192) and an 8 x 8 MHz (nchan 2048) datastream, dual pol, one 256 us FFT duration
    (3.9 kHz channels), plus six zoom windows of 0.5-2 MHz (128-512 channels); 0.512 s subints
    (2000 FFT blocks)."""
    from difx_amd import config as cf
    groups = [(64.0, 16384, [4800.0 + 64.0 * k for k in range(4)]),
              (32.0, 8192, [5200.0 + 32.0 * k for k in range(4)]),
              (8.0, 2048, [5400.0 + 8.0 * k for k in range(8)])]
    zooms = [(0, 0, 10.0, 2.0), (0, 2, 30.0, 0.5), (1, 1, 4.0, 1.0), (1, 3, 20.0, 2.0),
             (2, 0, 1.0, 0.5), (2, 5, 3.0, 2.0)]
    return cf.batched_config(nstations, groups, zooms, subint_ns=subint_ns, xengine=xengine,
                             seed=20261203, **kw)


def samples_per_subint(cfg) -> float:
    tot = 0.0
    for d, ds in enumerate(cfg.datastreams):
        f = cfg.freqs[ds.recfreqs[0]]
        rate = (1.0 if ds.complex_sampling else 2.0) * f.bw_mhz * 1e6
        tot += len(ds.recbands) * rate * cfg.subint_ns * 1e-9
    return tot


def f_engine_bytes_per_launch(cfg, nffts: int, out_bytes_per_chan: int) -> float:
    """Algorithmic HBM bytes of one fused F-engine launch over `nffts` FFT blocks: the 2-bit
    payload read once (L samples x nbands x 2 bits) plus the kept channels written once."""
    tot = 0.0
    for d, ds in enumerate(cfg.datastreams):
        C = cfg.freqs[ds.recfreqs[0]].nchan
        L = C if ds.complex_sampling else 2 * C
        nb = len(ds.recbands)
        bits = ds.nbits * (2 if ds.complex_sampling else 1)
        tot += nffts * (L * nb * bits / 8.0 + nb * C * out_bytes_per_chan)
    return tot


def x_engine_flops(cfg, B: int) -> float:
    """Useful X-engine flops per subint: 8 per complex multiply-accumulate of every
    baseline product and autocorrelation product (core.cpp:963, mode.cpp:1272,1334)."""
    tot = 0
    for b in cfg.baselines:
        for fi, prods in b.freqs:
            tot += len(prods) * cfg.freqs[fi].nchan
    for ds in cfg.datastreams:
        C = cfg.freqs[ds.recfreqs[0]].nchan
        tot += len(ds.recbands) * C * (2 if cfg.write_cross_autocorrs else 1)
    return 8.0 * tot * B


def serving_jobs(rng, window_s: float, rate_per_s: float, subints_per_job: int):
    """Poisson job stream: arrival times with mean rate `rate_per_s` over `window_s`; each job
    draws its station count from {4, 8, 16, 32}, its bandwidth as m x 32 MHz with m from
    {2, 4, 8, 16} (2-bit bands of one sample word must be a power of two) recorded as m/2
    freqs x 2 pols, nchan from {256, 1024, 4096}; 0.512 s subints (0.5 s is not a whole
    number of 4096-channel FFTs at 32 MHz); delay models of up to +-10 ms and 1 us/s."""
    from difx_amd import config as cf
    t, jobs = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate_per_s)
        if t > window_s:
            break
        n = int(rng.choice([4, 8, 16, 32]))
        m = int(rng.choice([2, 4, 8, 16]))
        nchan = int(rng.choice([256, 1024, 4096]))
        freqs = [4800.0 + 40.0 * k for k in range(m // 2)]
        B = 512_000_000 // (nchan * 1000 // 32)
        # spectra chunk of at most ~0.5 GB per job keeps many jobs resident at once
        per_fft = 4 * nchan * m * 32 * ((2 * n + 31) // 32)
        chunk = max(64, min(B, (512 << 20) // per_fft)) // 8 * 8
        cfg = cf.standard_config(n, 2, freqs, 32.0, nchan, subint_ns=512_000_000, xengine="fp16",
                                 seed=int(rng.integers(1 << 30)), chunk_ffts=chunk, max_delay_us=1e4)
        jobs.append((t, cfg, subints_per_job, (n, m, nchan)))
    return jobs


def serving_inputs(cfg, k: int):
    """Inputs of subint k of a serving job: the slot buffer bytes (getDataBytes plus four
    frames) and each station's control words.  Every station's buffer starts one frame before
    the earliest sample it needs (its own delay at the subint start), so the FFT windows of
    every station lie inside its buffer whatever the delay spread."""
    from difx_amd import synth
    ds = cfg.datastreams[0]
    nbytes = synth.data_bytes(cfg, 0) + 4 * ds.framebytes
    offs = [0, 1 + k, 0]
    ctrl = [synth.control_words(cfg, d, offs, synth.buffer_start_ns(cfg, d, offs))
            for d in range(len(cfg.datastreams))]
    return nbytes, ctrl


def serving_footprint(cfg, inflight: int = 2) -> int:
    """The server's admission footprint of a job with device-resident inputs: the context's
    device bytes plus its result buffers (fxg_server.cpp, fxg_server_add_job)."""
    from difx_amd.core import footprint
    dev, crl, _ = footprint(cfg)
    return dev + 8 * max(crl, 1) * (1 + inflight)


def serving_bench(torch, dev, window_s=3.0, rate_per_s=4.0, subints_per_job=59, seed=20261204,
                  hbm_budget_bytes=0):
    """Serving run: jobs of 59 subints of 0.512 s (30.2 s) arrive as a Poisson process of 4 jobs/s
    over `window_s`; each pushes one subint every 0.512 s of wall time from its arrival
    (real-time pacing); the library's scheduler admits them by device footprint (within
    `hbm_budget_bytes`, 0 = the device's free memory) and overlaps them.  The stations of one
    datastream shape read the same device-resident bytes, each through its own control words
    (its own buffer start time for its delay): the content does not affect the timing and every
    FFT window lies inside the buffer, so no work is skipped."""
    from difx_amd.server import FxServer
    rng = np.random.default_rng(seed)
    jobs = serving_jobs(rng, window_s, rate_per_s, subints_per_job)
    pool = {}
    events = []
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    from difx_amd import synth
    for j, (t0, cfg, nsub, shape) in enumerate(jobs):
        key = shape[1:]
        if key not in pool:
            ds = cfg.datastreams[0]
            nbytes, _ = serving_inputs(cfg, 0)
            nframes = nbytes // ds.framebytes
            buf = torch.randint(0, 256, (nframes, ds.framebytes), dtype=torch.uint8, device=dev, generator=g)
            hdr = np.stack([synth.vdif_header(0, f, ds.framebytes, len(ds.recbands), ds.nbits, False)
                            for f in range(nframes)])
            buf[:, : ds.headerbytes] = torch.from_numpy(hdr).to(dev)
            pool[key] = buf.reshape(-1).contiguous()
        for k in range(nsub):
            events.append((t0 + k * 0.512, j, k))
    events.sort()
    inputs = {(j, k): serving_inputs(jobs[j][1], k) for _, j, k in events}   # before the clock starts
    torch.cuda.synchronize()
    ids = {}
    total_samples = sum(samples_per_subint(c) * n for _, c, n, _ in jobs)
    with FxServer(hbm_budget_bytes=hbm_budget_bytes) as srv:
        t_start = time.perf_counter()
        for t_ev, j, k in events:
            delay = t_ev - (time.perf_counter() - t_start)
            if delay > 0:
                time.sleep(delay)
            t0, cfg, nsub, shape = jobs[j]
            if k == 0:
                ids[j] = srv.add_job(cfg, nsub, data_on_device=True)
            nbytes, ctrl = inputs[(j, k)]
            nds = len(cfg.datastreams)
            srv.push(ids[j], (0, 1 + k, 0), [pool[shape[1:]].data_ptr()] * nds, [nbytes] * nds, ctrl)
        per_job = [srv.wait_job(ids[j]) for j in range(len(jobs))]
        t_end = time.perf_counter()
        st = srv.stats()
    busy = st["last_done_s"] - st["first_launch_s"]
    offered = total_samples / (events[-1][0] + 0.512) / 1e9 if events else 0.0
    return {"jobs": len(jobs), "subints": st["subints_done"],
            "arrival_rate_jobs_per_s": rate_per_s, "window_s": window_s, "subints_per_job": subints_per_job,
            "hbm_budget_gb": hbm_budget_bytes / 1e9 if hbm_budget_bytes else None,
            "offered_gsamples_per_s": offered,
            "throughput_gsamples_per_s": st["samples_done"] / busy / 1e9 if busy > 0 else 0.0,
            "p50_subint_latency_ms": st["latency_p50_ms"], "p99_subint_latency_ms": st["latency_p99_ms"],
            "max_subint_latency_ms": st["latency_max_ms"], "subint_length_ms": 512.0,
            "max_concurrent_jobs": st["max_active"],
            "peak_admitted_gb": st["peak_admitted_bytes"] / 1e9,
            "largest_job_footprint_gb": max(serving_footprint(c) for _, c, _, _ in jobs) / 1e9 if jobs else 0.0,
            "worst_job_p99_ms": max(p[2] for p in per_job) if per_job else 0.0,
            "wall_s": t_end - t_start,
            "job_shapes_n_m_nchan": [list(j[3]) for j in jobs]}


def time_path(cfg, steps, warmup, torch, dev, offsets, kurtosis=False, pcal_hz=0):
    from difx_amd.core import FxCore
    core = FxCore(cfg)
    if kurtosis:
        core.set_kurtosis(True)
    for d in range(len(cfg.datastreams) if pcal_hz else 0):
        core.set_pcal(d, pcal_hz)
    bufs, nbytes, ctrl = device_slot(cfg, offsets, torch, dev)
    res = torch.zeros(2 * core.coreresultlength, dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    st = torch.cuda.Stream(device=dev)     # a real stream handle (the null stream is handle 0)
    sh = st.cuda_stream
    ptrs = [b.data_ptr() for b in bufs]
    for _ in range(warmup):
        core.processdata_device(offsets, ptrs, nbytes, ctrl, res.data_ptr(), sh)
    torch.cuda.synchronize()
    core.set_timing(True)
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        evs[k][0].record(st)
        core.processdata_device(offsets, ptrs, nbytes, ctrl, res.data_ptr(), sh)
        evs[k][1].record(st)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    lat = [a.elapsed_time(b) for a, b in evs]
    tm = core.timing()
    out = res.cpu().numpy().view(np.complex64).copy()
    info = {"wall_s": t1 - t0, "lat_ms": lat, "timing": tm, "chunkcap": None,
            "B": core.blockspersend, "results": out}
    core.set_timing(False)
    core.close()
    del bufs, res
    torch.cuda.empty_cache()
    return info


MUX_STATIONS, MUX_THREADS, MUX_SLOTS, MUX_TFB = 10, 16, 1000, 8032


def mux_inputs(torch, dev, seed=20261206):
    """Per station: 0.5 s of 8 x 32 MHz dual-polarisation 2-bit data recorded as 16
    single-channel VDIF threads (8000-byte payloads, 2000 frames/s per thread): 1000 frame
    times x 16 threads, in a random order, generated on the device."""
    sys.path.insert(0, str(ROOT / "oracle"))
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    n, ns, tfb, fps = MUX_THREADS, MUX_SLOTS, MUX_TFB, 2000
    s = np.repeat(np.arange(ns), n)
    l = np.tile(np.arange(n), ns)
    hdr = np.zeros((ns * n, 8), dtype=np.uint32)           # vdif_header fields, epoch 40
    hdr[:, 0] = 5000 + s // fps
    hdr[:, 1] = (s % fps) | (40 << 24)
    hdr[:, 2] = tfb // 8
    hdr[:, 3] = 0x4142 | (l << 16) | (1 << 26)
    hdr_d = torch.from_numpy(hdr.view(np.uint8).reshape(-1, 32)).to(dev)
    bufs = []
    for st in range(MUX_STATIONS):
        fr = torch.randint(0, 256, (ns * n, tfb), dtype=torch.uint8, device=dev, generator=g)
        fr[:, :32] = hdr_d
        perm = torch.randperm(ns * n, device=dev, generator=g)
        bufs.append(fr[perm].reshape(-1).contiguous())
        del fr
    ref_sec = 58849 * 86400 + 5000                         # epoch 40 = MJD 58849
    return bufs, ref_sec, fps


def mux_bench(torch, dev, steps=5, warmup=1):
    """Device corner turn of 10 stations' segments per step (one enqueue per station on one
    stream, then the results).  Bytes per step: every input frame read once, every output
    frame written once."""
    from difx_amd.vdifmux import VDIFMuxer
    bufs, ref_sec, fps = mux_inputs(torch, dev)
    muxers = []
    for st in range(MUX_STATIONS):
        m = VDIFMuxer(MUX_THREADS, MUX_TFB, fps, 2, list(range(MUX_THREADS)))
        m.set_reference(ref_sec, 0)
        muxers.append(m)
    ofb = muxers[0].outputframebytes
    outs = [torch.empty(MUX_SLOTS * ofb, dtype=torch.uint8, device=dev) for _ in range(MUX_STATIONS)]
    stream = torch.cuda.Stream(device=dev)
    def enqueue_all():
        for m, b, o in zip(muxers, bufs, outs):
            m.enqueue(b, MUX_SLOTS, out=o, start=0, stream=stream)
    for _ in range(warmup):
        enqueue_all()
        [m.finish() for m in muxers]
    torch.cuda.synchronize()
    # device time: the enqueued corner turns of all stations; each segment's 64-byte counter
    # readback (finish) follows the end event, as it would overlap the next segment's work
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    t0 = time.perf_counter()
    for k in range(steps):
        evs[k][0].record(stream)
        enqueue_all()
        evs[k][1].record(stream)
        nb = [m.finish() for m in muxers]
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = [a.elapsed_time(b) for a, b in evs]
    assert all(x == MUX_SLOTS * ofb for x in nb), nb
    # yardstick: a plain device copy of the same bytes (each station's input into its output)
    cev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    with torch.cuda.stream(stream):
        for k in range(steps):
            cev[k][0].record(stream)
            for b, o in zip(bufs, outs):
                n = min(b.numel(), o.numel())
                o[:n].copy_(b[:n])
            cev[k][1].record(stream)
    torch.cuda.synchronize()
    copy_ms = float(np.median([a.elapsed_time(b) for a, b in cev]))
    in_bytes = sum(b.numel() for b in bufs)
    out_bytes = MUX_STATIONS * MUX_SLOTS * ofb
    samples = MUX_STATIONS * MUX_SLOTS * MUX_THREADS * (MUX_TFB - 32) * 4
    dev_ms = float(np.median(ms))
    res = {"workload": "VDIF thread corner turn: 10 stations x 1000 frame times x 16 threads of 8032-byte "
                       "2-bit frames (0.5 s of 8 x 32 MHz dual-pol each), shuffled",
           "value": samples / (dev_ms / 1e3) / 1e9, "unit": "Gsamples/s",
           "ms_per_step_device": dev_ms, "ms_per_step_wall": wall * 1e3 / steps,
           "bytes_per_step": in_bytes + out_bytes,
           "achieved_GBps": (in_bytes + out_bytes) / (dev_ms / 1e3) / 1e9, "peak_GBps": HBM_PEAK_GBS,
           "frac": (in_bytes + out_bytes) / (dev_ms / 1e3) / 1e9 / HBM_PEAK_GBS,
           "device_copy_GBps": 2 * sum(min(b.numel(), o.numel()) for b, o in zip(bufs, outs)) / (copy_ms / 1e3) / 1e9}
    del bufs, outs
    torch.cuda.empty_cache()
    return res


def difx_bench(cfg, results: np.ndarray, torch, dev, reps: int = 20):
    """Output side of one headline subint: Visibility::writedata's scaling on the device
    (fxg_normalise_device, one call = cal + apply kernels + its synchronisation) and the DIFX
    records written on the host (fxg_difx_records)."""
    from difx_amd.core import FxCore
    core = FxCore(cfg)
    d = torch.from_numpy(results.view(np.float32).copy()).to(dev)
    core.normalise_device(d.data_ptr(), 1.0)            # warm-up
    t0 = time.perf_counter()
    for _ in range(reps):
        core.normalise_device(d.data_ptr(), 1.0)
    norm_ms = (time.perf_counter() - t0) * 1e3 / reps
    host = d.cpu().numpy().view(np.complex64)
    nb = len(cfg.baselines)
    bl = [256 * (b.ds1 + 1) + b.ds2 + 1 for b in cfg.baselines]
    uvw = np.zeros((nb, 3))
    tel = list(range(len(cfg.datastreams)))
    t0 = time.perf_counter()
    for _ in range(3):
        recs = core.difx_records(host, 60000, 0.25, 0, 0, float(cfg.blockspersend()), bl, uvw, tel)
    rec_ms = (time.perf_counter() - t0) * 1e3 / 3
    core.close()
    vis_bytes = results.nbytes
    return {"workload": "scaling and DIFX records of one subint's results of the headline configuration "
                        f"({results.size} cf32, {vis_bytes / 1e6:.1f} MB)",
            "normalise_ms_per_call": norm_ms, "normalise_GBps_incl_sync": 2 * vis_bytes / (norm_ms / 1e3) / 1e9,
            "records_bytes": len(recs), "records_ms_host": rec_ms,
            "records_GBps_host": len(recs) / (rec_ms / 1e3) / 1e9}


def mux_cpu_baseline(budget_s: float = 6.0):
    """The corner-turn restatement (oracle/vdifmux.py, numpy) on segments of 20 frame times of
    the same 16-thread workload, repeated for about budget_s."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import vdifmux as vm
    rng = np.random.default_rng(5)
    n, ns, tfb, fps = MUX_THREADS, 20, MUX_TFB, 2000
    fr = rng.integers(0, 256, size=(ns * n, tfb), dtype=np.uint8)
    for i in range(ns * n):
        fr[i, :32] = vm.make_header(5000, i // n, 40, tfb, i % n, 2).view(np.uint8)
    buf = fr[rng.permutation(ns * n)].reshape(-1)
    ref_sec = vm.epoch_mjd(40) * 86400 + 5000
    runs, dt = 0, 0.0
    while dt < budget_s and runs < 100:
        t0 = time.perf_counter()
        vm.mux_segment(buf, tfb, list(range(n)), fps, 2, ref_sec, 0, 0, ns)
        dt += time.perf_counter() - t0
        runs += 1
    samples = runs * ns * n * (tfb - 32) * 4
    return {"value": samples / dt / 1e9, "unit": "Gsamples/s", "cores": 1, "kind": "port",
            "sample": f"numpy restatement of VDIFMuxer, {runs} segments of {ns} frame times x {n} threads, "
                      f"{dt:.2f} s"}


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _port_rate(cfg, nthreads: int, budget_s: float, offsets=(0, 1, 0)):
    """Passes of the C restatement over one slot of cfg until about budget_s of wall time:
    (Gsamples/s, per-pass seconds)."""
    import fxport
    from difx_amd import synth
    data, ctrl = synth.noise_slot(cfg, list(offsets), seed=3)
    fxport.process_subint(cfg, list(offsets), data, ctrl, nthreads=nthreads)     # warm-up, plans
    times = []
    while sum(times) < budget_s and len(times) < 200:
        t0 = time.perf_counter()
        fxport.process_subint(cfg, list(offsets), data, ctrl, nthreads=nthreads)
        times.append(time.perf_counter() - t0)
    return samples_per_subint(cfg) * len(times) / sum(times) / 1e9, times


def cpu_baseline(budget_s: float = 12.0, nthreads: int = 0, nstations: int = 16):
    """Time the CPU restatement on a bounded sample of the workload: all stations x 16 bands x
    all baselines of the nstations dual-pol 8 x 32 MHz nchan-1024 workload, for a slot of
    256 FFT blocks per thread (at most 4096), repeated for about budget_s."""
    sys.path.insert(0, str(ROOT / "oracle"))
    from difx_amd import config as cf
    try:
        import fxport  # C restatement (oracle/fxport.c), multithreaded
        have_port = fxport.available()
    except Exception:
        have_port = False
    freqs = [4800.0 + 40.0 * k for k in range(8)]
    offsets = [0, 1, 0]
    if have_port:
        T = nthreads or fxport.default_threads()
        nblk = min(4096, 256 * T)
        cfg = cf.standard_config(nstations, 2, freqs, 32.0, 1024, subint_ns=nblk * 32000, seed=20261202)
        rate, times = _port_rate(cfg, T, budget_s, offsets)
        samples = samples_per_subint(cfg) * len(times)
        return {"value": rate, "unit": "Gsamples/s", "cores": T, "kind": "port",
                "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
                "sample": f"C restatement (oracle/fxport.c, FFT: {fxport.fft_backend()}), {len(times)} passes over "
                          f"{nblk} FFT blocks of the {nstations}-station dual-pol 8x32 MHz nchan-1024 workload "
                          f"({samples/1e6:.0f} Msamples), {T} threads, {sum(times):.2f} s wall"}
    import fxoracle as fo
    from difx_amd import synth
    nblk = 4
    cfg = cf.standard_config(nstations, 2, freqs, 32.0, 1024, subint_ns=nblk * 32000, seed=20261202)
    data, ctrl = synth.noise_slot(cfg, offsets, seed=3)
    t0 = time.perf_counter()
    fo.process_subint(cfg, offsets, data, ctrl)
    dt = time.perf_counter() - t0
    samples = samples_per_subint(cfg)
    return {"value": samples / dt / 1e9, "unit": "Gsamples/s", "cores": 1, "kind": "port",
            "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
            "sample": f"numpy FP64 oracle, {nblk} FFT blocks of the workload "
                      f"({samples/1e6:.1f} Msamples), 1 thread, {dt:.2f} s"}


def cpu_table(budget_s: float = 8.0):
    """The C restatement per benchmark configuration, with one thread and with all host
    threads: Gsamples/s, and the wall time of one whole subint at that rate (p50 / p99 over
    the timed passes of the sample, scaled to the subint)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import fxport
    from difx_amd import config as cf
    T = fxport.default_threads()
    freqs8 = [4800.0 + 40.0 * k for k in range(8)]
    rows = []

    def row(name, cfg_of, full_cfg, T_):
        sample = cfg_of(T_)
        rate, times = _port_rate(sample, T_, budget_s)
        per = np.array(times) * samples_per_subint(full_cfg) / samples_per_subint(sample)
        rows.append({"config": name, "threads": T_, "gsamples_per_s": rate,
                     "subint_s": full_cfg.subint_ns / 1e9,
                     "p50_subint_s": float(np.percentile(per, 50)), "p99_subint_s": float(np.percentile(per, 99)),
                     "passes": len(times), "sample_fft_blocks": sample.blockspersend()})
        print(json.dumps(rows[-1]), file=sys.stderr, flush=True)

    for T_ in (1, T):
        # cfg0: 4 stations, 1 x 32 MHz single pol, nchan 256, 1 s subints
        row("cfg0", lambda t: cf.standard_config(4, 1, [4800.0], 32.0, 256, subint_ns=min(4096, 1024 * t) * 8000,
                                                 seed=1),
            cf.standard_config(4, 1, [4800.0], 32.0, 256, subint_ns=1_000_000_000), T_)
        # cfg1: 10 stations dual pol 8 x 32 MHz nchan 1024, 0.5 s subints
        row("cfg1", lambda t: cf.standard_config(10, 2, freqs8, 32.0, 1024, subint_ns=min(4096, 128 * t) * 32000,
                                                 seed=2),
            cf.standard_config(10, 2, freqs8, 32.0, 1024, subint_ns=500_000_000), T_)
        # cfg2: 64 stations, 16 MHz complex 8-bit, nchan 4096, subint 999 936 000 ns
        row("cfg2", lambda t: cf.standard_config(64, 1, [1400.0], 16.0, 4096, subint_ns=min(64, 4 * t) * 256000,
                                                 nbits=8, complex_sampling=1, seed=3, max_delay_us=200.0),
            cf.standard_config(64, 1, [1400.0], 16.0, 4096, subint_ns=999_936_000, nbits=8, complex_sampling=1),
            T_)
        # cfg3: the uneven 16-station workload of the uneven_bands section, 0.512 s subints
        row("cfg3", lambda t: uneven_config("fp32", subint_ns=min(64, 4 * t) * 256000, max_delay_us=200.0),
            uneven_config("fp32"), T_)
    return {"cpu_model": cpu_model(), "host_cpus": os.cpu_count(), "fft": fxport.fft_backend(), "rows": rows}


def measured_traffic(path: Path, ffts_per_launch: float, stations: int = 16):
    """HBM bytes per F-engine launch from the committed rocprofv3 --pmc summary: FETCH_SIZE
    and WRITE_SIZE (KiB per dispatch) each scaled by what the same counter reads for a plain
    device copy of known size (`bench.py --copy-probe`), rescaled to this launch's FFT count.
    None when the summary was taken on another station count."""
    if not path.exists():
        return None
    kv = {}
    for line in path.read_text().splitlines():
        if ":" in line and not line.startswith("#"):
            k, v = line.split(":", 1)
            try:
                kv[k.strip()] = float(v.split()[0])
            except (ValueError, IndexError):
                pass
    need = ("fetch_kib_per_dispatch", "write_kib_per_dispatch", "copy_fetch_kib", "copy_write_kib",
            "copy_bytes", "ffts_per_dispatch", "stations")
    if not all(k in kv for k in need) or int(kv["stations"]) != stations:
        return None
    fscale = kv["copy_bytes"] / (kv["copy_fetch_kib"] * 1024.0)
    wscale = kv["copy_bytes"] / (kv["copy_write_kib"] * 1024.0)
    per = (kv["fetch_kib_per_dispatch"] * fscale + kv["write_kib_per_dispatch"] * wscale) * 1024.0
    return per * ffts_per_launch / kv["ffts_per_dispatch"]


def copy_probe():
    """A plain 1 GiB device-to-device copy, the yardstick for scaling FETCH_SIZE/WRITE_SIZE."""
    import torch
    x = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
    x.fill_(1.0)
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    y.copy_(x)
    torch.cuda.synchronize()
    print("copy probe: 1073741824 bytes copied", flush=True)


# VALU lane instructions per input sample of k_fengine<2048,false,1,256> (real 2-bit, 16 bands
# per word), from the SQ_INSTS_VALU counter summary committed with the kernel's profile.
FENGINE_VALU_PER_SAMPLE_FILE = ROOT / "profiles" / "k_fengine_2048_valu_counters.txt"
VALU_PEAK_LANE_OPS = 256 * 64 * 2.4e9          # 256 CUs x 64 lanes x 2.4 GHz peak engine clock


def valu_per_sample():
    if not FENGINE_VALU_PER_SAMPLE_FILE.exists():
        return None
    for line in FENGINE_VALU_PER_SAMPLE_FILE.read_text().splitlines():
        if line.startswith("SQ_INSTS_VALU ") and "lane instructions per input sample" in line:
            return float(line.split("->")[1].split()[0])
    return None


def path_summary(r, nsamp, steps):
    lat = np.array(r["lat_ms"])
    tm = r["timing"]
    return {"value": nsamp * steps / r["wall_s"] / 1e9, "unit": "Gsamples/s", "steps": steps,
            "ms_per_step": r["wall_s"] * 1e3 / steps,
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "f_ms_per_step": tm["f_ms"] / steps, "x_ms_per_step": tm["x_ms"] / steps,
            "other_ms_per_step": tm["other_ms"] / steps}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cfg1", action="store_true")
    ap.add_argument("--no-bf16", action="store_true")
    ap.add_argument("--no-postf", action="store_true")
    ap.add_argument("--no-uneven", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-mux", action="store_true")
    ap.add_argument("--no-output", action="store_true")
    ap.add_argument("--no-kurtosis", action="store_true")
    ap.add_argument("--no-pcal", action="store_true")
    ap.add_argument("--copy-probe", action="store_true", help="1 GiB device copy for PMC scaling")
    ap.add_argument("--cpu-table", action="store_true", help="C restatement per configuration (no GPU)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return
    if args.cpu_table:
        print(json.dumps(cpu_table()))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs the FX path on a single MI355X; "
              "multi-GPU sharding (whole subints or subband groups per GPU) is not part of it.",
              file=sys.stderr)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print("no GPU available", file=sys.stderr)
        sys.exit(1)
    dev = torch.device("cuda:0")
    offsets = [0, 1, 0]
    cfg = workload_config("fp16")
    nsamp = samples_per_subint(cfg)
    print(f"[bench] 16 stations, MFMA path (F + FP16 MFMA X), {args.warmup} warmup + {args.steps} steps",
          file=sys.stderr, flush=True)
    head = time_path(cfg, args.steps, args.warmup, torch, dev, offsets)
    B = head["B"]
    hs = path_summary(head, nsamp, args.steps)
    value = hs["value"]
    tm = head["timing"]
    f_avg_ms = tm["f_ms"] / max(tm["f_launches"], 1)
    nchunks_per_step = tm["f_launches"] / args.steps
    ffts_per_launch = B / nchunks_per_step
    fbytes = f_engine_bytes_per_launch(cfg, ffts_per_launch, 4)      # FP16 complex spectra out (half2)
    achieved = fbytes / (f_avg_ms / 1e3) / 1e9
    f_samples_per_launch = ffts_per_launch * sum(2 * cfg.freqs[d.recfreqs[0]].nchan * len(d.recbands)
                                                 for d in cfg.datastreams)
    vps = valu_per_sample()
    issue = None
    if vps is not None:
        ach = vps * f_samples_per_launch / (f_avg_ms / 1e3)
        issue = {"bound": "valu", "achieved": ach, "peak": VALU_PEAK_LANE_OPS, "unit": "lane-instructions/s",
                 "frac": ach / VALU_PEAK_LANE_OPS, "valu_lane_instructions_per_sample": vps,
                 "source": "profiles/k_fengine_2048_valu_counters.txt (SQ_INSTS_VALU x 64 / input samples)"}
    plain = None
    if not args.no_plain:
        pcfg = workload_config("fp32")
        print(f"[bench] 16 stations, plain FP32 path, {args.warmup} warmup + {args.steps} steps",
              file=sys.stderr, flush=True)
        p = time_path(pcfg, args.steps, args.warmup, torch, dev, offsets)
        plain = path_summary(p, nsamp, args.steps)
        # agreement of the two paths on identical inputs (diffDiFX metric over the whole array)
        a, b = head["results"], p["results"]
        plain["mfma_vs_plain_rel_diff"] = float(np.mean(np.abs(a - b)) / max(np.mean(np.abs(b)), 1e-30))
        # the blocked FP32 X-engine against the VALU issue rate: per freq of n inputs it
        # computes nb (nb + 1) / 2 blocks of 8 x 8 products (nb = ceil(n / 8)) per channel and
        # FFT, 4 FMA lane-instructions per product
        lane_ops = 0.0
        for fi, fr in enumerate(pcfg.freqs):
            n = sum(1 for ds in pcfg.datastreams for lf, _ in ds.recbands if ds.recfreqs[lf] == fi)
            nb = (n + 7) // 8
            lane_ops += nb * (nb + 1) / 2 * 64 * 4 * fr.nchan * B
        xs = plain["x_ms_per_step"] / 1e3
        plain["x_engine_valu_issue"] = {
            "kernel": "k_xmac_f32_blk", "achieved": lane_ops / xs, "peak": VALU_PEAK_LANE_OPS,
            "unit": "lane-instructions/s", "frac": lane_ops / xs / VALU_PEAK_LANE_OPS,
            "fma_lane_instructions_per_subint": lane_ops}
    bf16 = None
    if not args.no_bf16:
        bsteps = max(1, min(args.steps, 10))
        print(f"[bench] 16 stations, BF16 MFMA path, 1 warmup + {bsteps} steps", file=sys.stderr, flush=True)
        r = time_path(workload_config("bf16"), bsteps, 1, torch, dev, offsets)
        bf16 = path_summary(r, nsamp, bsteps)
    postf = None
    if not args.no_postf:
        psteps = max(1, min(args.steps, 10))
        print(f"[bench] 16 stations, post-F fringe rotation, MFMA path, 1 warmup + {psteps} steps",
              file=sys.stderr, flush=True)
        r = time_path(workload_config("fp16", fringe_rot_order=0), psteps, 1, torch, dev, offsets)
        postf = path_summary(r, nsamp, psteps)
    cfg1 = None
    if not args.no_cfg1:
        c1 = workload_config("fp16", nstations=10)
        n1 = samples_per_subint(c1)
        csteps = max(1, min(args.steps, 10))
        print(f"[bench] 10 stations, MFMA and plain paths, 1 warmup + {csteps} steps", file=sys.stderr, flush=True)
        r1 = time_path(c1, csteps, 1, torch, dev, offsets)
        r2 = time_path(workload_config("fp32", nstations=10), csteps, 1, torch, dev, offsets)
        cfg1 = {"workload": "10 stations dual-pol, 8 x 32 MHz 2-bit VDIF, nchan 1024, 0.5 s subint",
                "mfma_fp16": path_summary(r1, n1, csteps), "plain_fp32": path_summary(r2, n1, csteps)}
        cfg1["speedup_mfma_vs_plain"] = cfg1["mfma_fp16"]["value"] / cfg1["plain_fp32"]["value"]
    uneven = None
    if not args.no_uneven:
        from difx_amd.core import FxCore
        tcfg = uneven_config("fp16", autotune=1)
        print("[bench] uneven bands: autotune of the F-engine launch form", file=sys.stderr, flush=True)
        core = FxCore(tcfg)
        report = core.autotune_report()
        core.close()
        usamp = samples_per_subint(tcfg)
        forms = {}
        usteps = 5
        for name, mode in (("per_length", 1), ("persistent", 2)):
            ucfg = uneven_config("fp16", fengine_mode=mode)
            print(f"[bench] uneven bands ({name} F-engine launch), 1 warmup + {usteps} steps", file=sys.stderr, flush=True)
            u = time_path(ucfg, usteps, 1, torch, dev, offsets)
            forms[name] = path_summary(u, usamp, usteps)
        best = max(forms, key=lambda n: forms[n]["value"])
        uneven = {"workload": "16 stations dual-pol: 4 x 64 MHz (nchan 16384), 4 x 32 MHz (nchan 8192), "
                              "8 x 8 MHz (nchan 2048) datastreams + 6 zooms of 0.5-2 MHz (128-512 channels), "
                              "0.512 s subint",
                  "value": forms[best]["value"], "unit": "Gsamples/s", "p99_ms": forms[best]["p99_ms"],
                  "subint_length_ms": 512.0, "forms": forms, "autotune_report": report}
    serving = None
    if not args.no_serving:
        print("[bench] serving: Poisson job stream", file=sys.stderr, flush=True)
        serving = serving_bench(torch, dev)
        # the same stream under an HBM budget of a third of what it admitted at its peak (at
        # least the largest job): admission by footprint makes later jobs wait for earlier ones
        budget = int(max(serving["peak_admitted_gb"] / 3, serving["largest_job_footprint_gb"]) * 1e9)
        print("[bench] serving: the same stream under a third of its peak footprint", file=sys.stderr, flush=True)
        serving["under_budget"] = serving_bench(torch, dev, hbm_budget_bytes=budget)
    kcfg = workload_config("fp16", nstations=10)
    kbase = None
    if not (args.no_kurtosis and args.no_pcal):
        print("[bench] 10-station reference run for the side products, 1 warmup + 3 steps", file=sys.stderr, flush=True)
        kbase = time_path(kcfg, 3, 1, torch, dev, offsets)
    kurt = None
    if not args.no_kurtosis:
        print("[bench] 10 stations with spectral kurtosis, 1 warmup + 3 steps", file=sys.stderr, flush=True)
        kr = time_path(kcfg, 3, 1, torch, dev, offsets, kurtosis=True)
        k_ms = kr["wall_s"] * 1e3 / 3
        spec_bytes = float(samples_per_subint(kcfg)) / 2 * 4           # FP16 complex spectra: C = L/2 per FFT
        extra = (kr["timing"]["other_ms"] - kbase["timing"]["other_ms"]) / 3
        kurt = {"workload": "10-station subint with Mode's kurtosis sums (|X|^2, |X|^4 per band and channel)",
                "ms_per_step": k_ms, "extra_ms_vs_without": k_ms - kbase["wall_s"] * 1e3 / 3,
                "spectra_bytes_read_per_step": spec_bytes,
                "kurtosis_pass_GBps": spec_bytes / max(extra / 1e3, 1e-9) / 1e9}
    pcal = None
    if not args.no_pcal:
        print("[bench] 10 stations with phase-cal extraction, 1 warmup + 3 steps", file=sys.stderr, flush=True)
        pr = time_path(kcfg, 3, 1, torch, dev, offsets, pcal_hz=1_000_000)
        p_ms = pr["wall_s"] * 1e3 / 3
        extra_other = (pr["timing"]["other_ms"] - kbase["timing"]["other_ms"]) / 3
        raw_bytes = float(samples_per_subint(kcfg)) * kcfg.datastreams[0].nbits / 8   # sample bytes read once
        pcal = {"workload": "10-station subint with phase-cal extraction, 1 MHz comb on every band",
                "ms_per_step": p_ms, "extra_ms_vs_without": p_ms - kbase["wall_s"] * 1e3 / 3,
                "pcal_ms_per_step": extra_other,
                "sample_bytes_per_step": raw_bytes,
                "pcal_pass_GBps": raw_bytes / max(extra_other / 1e3, 1e-9) / 1e9}
    output = None
    if not args.no_output:
        print("[bench] output side: scaling and DIFX records of one subint", file=sys.stderr, flush=True)
        output = difx_bench(cfg, head["results"], torch, dev)
    mux = None
    if not args.no_mux:
        print("[bench] VDIF thread corner turn, 1 warmup + 5 steps", file=sys.stderr, flush=True)
        mux = mux_bench(torch, dev)
    cpu = cpu1 = None
    if not args.no_cpu:
        print("[bench] CPU baseline, all host threads", file=sys.stderr, flush=True)
        cpu = cpu_baseline()
        print("[bench] CPU baseline, one thread", file=sys.stderr, flush=True)
        cpu1 = cpu_baseline(budget_s=8.0, nthreads=1)
        if mux is not None:
            mux["cpu_baseline"] = mux_cpu_baseline()
    realtime = nsamp / (cfg.subint_ns * 1e-9) / 1e9
    line = {
        "metric": "input Gsamples/s (all stations) + p99 subint latency; MFMA vs plain FP32 path",
        "value": value,
        "unit": "Gsamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": hs["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 F-engine (f64 delay and fringe phase), fp16 X-engine inputs, f32 accumulate",
        "data": "synthetic (device-generated 2-bit VDIF, seeded quintic delay models)",
        "config": {"workload": "16 stations dual-pol, 8 x 32 MHz 2-bit VDIF, nchan 1024, 0.5 s subint "
                               "(the MFMA-vs-plain target shape)",
                   "stations": 16, "pols": 2, "subbands": 8, "bw_mhz": 32, "nchan": 1024,
                   "subint_s": 0.5, "fft_blocks_per_subint": B, "parallelism": "single GPU"},
        "p50_subint_latency_ms": hs["p50_ms"],
        "p99_subint_latency_ms": hs["p99_ms"],
        "subint_length_ms": cfg.subint_ns / 1e6,
        "realtime_factor": value / realtime,
        "plain_fp32_value": plain["value"] if plain else None,
        "plain_fp32_p99_subint_latency_ms": plain["p99_ms"] if plain else None,
        "speedup_mfma_vs_plain": value / plain["value"] if plain else None,
        "kernels": {"f_engine_ms_per_step": hs["f_ms_per_step"],
                    "x_engine_ms_per_step": hs["x_ms_per_step"],
                    "other_ms_per_step": hs["other_ms_per_step"],
                    "f_launches_per_step": nchunks_per_step,
                    "x_engine_useful_tflops": x_engine_flops(cfg, B) / (hs["x_ms_per_step"] / 1e3) / 1e12},
        "roofline": {"kernel": "k_fengine<2048,false,1,256> (fused unpack+rotate+FFT+post-F, FP16 out)",
                     "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(ROOT / "profiles" / "k_fengine_2048_traffic.txt", ffts_per_launch),
                     "avg_launch_ms": f_avg_ms, "algorithmic_bytes_per_launch": fbytes},
        "issue_roofline": issue,
        "plain_fp32": plain,
        "mfma_bf16": bf16,
        "mfma_post_f": postf,
        "cfg1_10_stations": cfg1,
        "uneven_bands": uneven,
        "serving": serving,
        "vdif_corner_turn": mux,
        "output_side": output,
        "kurtosis": kurt,
        "phase_cal": pcal,
        "cpu_baseline": cpu,
        "cpu_baseline_1thread": cpu1,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
