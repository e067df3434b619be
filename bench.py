#!/usr/bin/env python3
"""Benchmark of the FX hot path on one MI355X: input Gsamples/s and subint latency.

Workload (one "step" = one subintegration through the whole per-subint pipeline):
10 stations, dual circular polarisation, 8 subbands x 32 MHz, 2-bit real VDIF with 16 bands
per 32-bit sample word (8000-byte payloads), nchan 1024 (2048-point FFTs), 0.5 s subints
(15625 FFT blocks per station), fringe rotation order 1, seeded quintic delay models.
Synthetic 2-bit data is generated on the device before the timed region (inputs resident
in HBM), so `value` excludes PCIe transfers.

Two paths are timed on the same inputs:
  headline  fused F-engine + FP16 MFMA Hermitian X-engine   -> `value`
  plain     fused F-engine (FP32 spectra) + FP32 CMAC X-engine per baseline product

The roofline object describes the headline path's dominant kernel, timed with HIP events
on its own stream inside the library (fxg_set_timing).  The CPU baseline runs the C
restatement in oracle/ (when built) or the numpy oracle on a bounded sample of the same
workload on the host.

Further sections of the same JSON line (each can be skipped with its --no-* flag):
  plain_fp32          the plain path on the same inputs
  target_16_stations  16 stations, MFMA path vs plain path
  uneven_bands        mixed bandwidths and zooms, both F-engine launch forms, autotuner report
  serving             Poisson job stream through the server (p50 / p99 subint latency)
  kurtosis            the headline subint with the spectral-kurtosis sums enabled
  phase_cal           the headline subint with phase-cal extraction on every band (1 MHz comb)
  output_side         device scaling (writedata) and DIFX records of one subint's results
  vdif_corner_turn    10 stations' 16-thread VDIF segments corner-turned on the device, with
                      a device-copy yardstick and the numpy restatement's rate

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-plain] [--no-uneven]
                    [--no-serving] [--no-target] [--no-kurtosis] [--no-pcal] [--no-output]
                    [--no-mux]
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


def workload_config(xengine: str):
    from difx_amd import config as cf
    freqs = [4800.0 + 40.0 * k for k in range(8)]
    return cf.standard_config(10, 2, freqs, 32.0, 1024, subint_ns=500_000_000,
                              xengine=xengine, seed=20261202)


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


def uneven_config(xengine: str, nstations: int = 16, **kw):
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
    return cf.batched_config(nstations, groups, zooms, subint_ns=512_000_000, xengine=xengine,
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
    number of 4096-channel FFTs at 32 MHz).  Delays stay within +-36 us so that the stations
    of a job can share one input buffer (see serving_inputs)."""
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
                                 seed=int(rng.integers(1 << 30)), chunk_ffts=chunk, max_delay_us=40.0)
        jobs.append((t, cfg, subints_per_job, (n, m, nchan)))
    return jobs


def serving_inputs(cfg, k: int):
    """Shared-buffer inputs of subint k of a serving job: (buffer bytes, control words) valid
    for every station.  The buffer holds the slot capacity (getDataBytes plus four frames) and
    starts one frame before the earliest sample any station needs (its delay at the subint
    start), which leaves room for the +-36 us spread of the stations' delays."""
    import math
    from difx_amd import synth
    ds = cfg.datastreams[0]
    frame_ns = int(round(synth.samples_per_frame(ds) * cfg.sampletime_us(0) * 1000.0))
    t0 = (1 + k) * 1e9
    first = min(t0 - 1000.0 * synth.eval_delay(d.model, d.model_order, t0 / 1e9, cfg.model_inc_secs)
                for d in cfg.datastreams)
    start = int(math.floor(first / frame_ns) - 1) * frame_ns
    nbytes = synth.data_bytes(cfg, 0) + 4 * ds.framebytes
    return nbytes, synth.control_words(cfg, 0, [0, 1 + k, 0], start)


def serving_bench(torch, dev, window_s=6.0, rate_per_s=4.0, subints_per_job=8, seed=20261204):
    """Serving run: jobs arrive as a Poisson process, each pushes one 0.512 s subint every
    0.512 s of wall time from its arrival (real-time pacing), the library's scheduler admits
    them by device footprint and overlaps them.  Input buffers are device-resident and shared
    by every station and job of one datastream shape (their content does not affect the
    timing; every FFT window lies inside them, so no work is skipped)."""
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
    torch.cuda.synchronize()
    ids = {}
    total_samples = sum(samples_per_subint(c) * n for _, c, n, _ in jobs)
    with FxServer() as srv:
        t_start = time.perf_counter()
        for t_ev, j, k in events:
            delay = t_ev - (time.perf_counter() - t_start)
            if delay > 0:
                time.sleep(delay)
            t0, cfg, nsub, shape = jobs[j]
            if k == 0:
                ids[j] = srv.add_job(cfg, nsub, data_on_device=True)
            nbytes, ctrl = serving_inputs(cfg, k)
            nds = len(cfg.datastreams)
            srv.push(ids[j], (0, 1 + k, 0), [pool[shape[1:]].data_ptr()] * nds, [nbytes] * nds, [ctrl] * nds)
        per_job = [srv.wait_job(ids[j]) for j in range(len(jobs))]
        t_end = time.perf_counter()
        st = srv.stats()
    busy = st["last_done_s"] - st["first_launch_s"]
    offered = total_samples / (events[-1][0] + 0.512) / 1e9 if events else 0.0
    return {"jobs": len(jobs), "subints": st["subints_done"],
            "arrival_rate_jobs_per_s": rate_per_s, "window_s": window_s, "subints_per_job": subints_per_job,
            "offered_gsamples_per_s": offered,
            "throughput_gsamples_per_s": st["samples_done"] / busy / 1e9 if busy > 0 else 0.0,
            "p50_subint_latency_ms": st["latency_p50_ms"], "p99_subint_latency_ms": st["latency_p99_ms"],
            "max_subint_latency_ms": st["latency_max_ms"], "subint_length_ms": 512.0,
            "max_concurrent_jobs": st["max_active"],
            "peak_admitted_gb": st["peak_admitted_bytes"] / 1e9,
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


def cpu_baseline(budget_s: float = 12.0):
    """Time the CPU restatement on a bounded sample of the workload: all 10 stations x 16
    bands x all 45 baselines, for as many FFT blocks as fit the time budget."""
    sys.path.insert(0, str(ROOT / "oracle"))
    from difx_amd import config as cf, synth
    try:
        import fxport  # C restatement (oracle/fxport.c), multithreaded
        have_port = fxport.available()
    except Exception:
        have_port = False
    freqs = [4800.0 + 40.0 * k for k in range(8)]
    offsets = [0, 1, 0]
    if have_port:
        # one 4096-FFT-block slot of the workload, processed repeatedly until about
        # budget_s of wall time has been spent (the CPU work is the same every pass)
        nthreads = fxport.default_threads()
        nblk = 4096
        cfg = cf.standard_config(10, 2, freqs, 32.0, 1024, subint_ns=nblk * 32000, seed=20261202)
        data, ctrl = synth.noise_slot(cfg, offsets, seed=3)
        fxport.process_subint(cfg, offsets, data, ctrl, nthreads=nthreads)     # warm-up, plans
        runs, dt = 0, 0.0
        while dt < budget_s and runs < 200:
            t0 = time.perf_counter()
            fxport.process_subint(cfg, offsets, data, ctrl, nthreads=nthreads)
            dt += time.perf_counter() - t0
            runs += 1
        samples = samples_per_subint(cfg) * runs
        return {"value": samples / dt / 1e9, "unit": "Gsamples/s", "cores": nthreads,
                "kind": "port",
                "sample": f"C restatement (oracle/fxport.c, FFT: {fxport.fft_backend()}), {runs} passes over "
                          f"{nblk} FFT blocks of the 10-station dual-pol 8x32 MHz nchan-1024 workload "
                          f"({samples/1e6:.0f} Msamples), {nthreads} threads, {dt:.2f} s wall"}
    import fxoracle as fo
    nblk = 4
    cfg = cf.standard_config(10, 2, freqs, 32.0, 1024, subint_ns=nblk * 32000, seed=20261202)
    data, ctrl = synth.noise_slot(cfg, offsets, seed=3)
    t0 = time.perf_counter()
    fo.process_subint(cfg, offsets, data, ctrl)
    dt = time.perf_counter() - t0
    samples = samples_per_subint(cfg)
    return {"value": samples / dt / 1e9, "unit": "Gsamples/s", "cores": 1, "kind": "port",
            "sample": f"numpy FP64 oracle, {nblk} FFT blocks of the workload "
                      f"({samples/1e6:.1f} Msamples), 1 thread, {dt:.2f} s"}


def measured_traffic(path: Path, ffts_per_launch: float):
    """HBM bytes per F-engine launch from the committed rocprofv3 --pmc summary: FETCH_SIZE
    and WRITE_SIZE (KiB per dispatch) each scaled by what the same counter reads for a plain
    device copy of known size (`bench.py --copy-probe`), rescaled to this launch's FFT count."""
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
            "copy_bytes", "ffts_per_dispatch")
    if not all(k in kv for k in need):
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-uneven", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-target", action="store_true")
    ap.add_argument("--no-mux", action="store_true")
    ap.add_argument("--no-output", action="store_true")
    ap.add_argument("--no-kurtosis", action="store_true")
    ap.add_argument("--no-pcal", action="store_true")
    ap.add_argument("--copy-probe", action="store_true", help="1 GiB device copy for PMC scaling")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
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
    print(f"[bench] headline path (F + FP16 MFMA X), {args.warmup} warmup + {args.steps} steps", file=sys.stderr, flush=True)
    head = time_path(cfg, args.steps, args.warmup, torch, dev, offsets)
    B = head["B"]
    value = nsamp * args.steps / head["wall_s"] / 1e9
    lat = np.array(head["lat_ms"])
    tm = head["timing"]
    f_avg_ms = tm["f_ms"] / max(tm["f_launches"], 1)
    nchunks_per_step = tm["f_launches"] / args.steps
    ffts_per_launch = B / nchunks_per_step
    fbytes = f_engine_bytes_per_launch(cfg, ffts_per_launch, 4)
    achieved = fbytes / (f_avg_ms / 1e3) / 1e9
    x_avg_ms = tm["x_ms"] / max(tm["x_launches"], 1)
    plain = None
    if not args.no_plain:
        pcfg = workload_config("fp32")
        psteps = max(1, min(args.steps, 3))
        print(f"[bench] plain FP32 path, 1 warmup + {psteps} steps", file=sys.stderr, flush=True)
        p = time_path(pcfg, psteps, 1, torch, dev, offsets)
        plat = np.array(p["lat_ms"])
        pval = nsamp * psteps / p["wall_s"] / 1e9
        ptm = p["timing"]
        # agreement of the two paths on identical inputs (diffDiFX metric over the cross records)
        a, b = head["results"], p["results"]
        rel = float(np.mean(np.abs(a - b)) / max(np.mean(np.abs(b)), 1e-30))
        plain = {"value": pval, "unit": "Gsamples/s", "p50_ms": float(np.percentile(plat, 50)),
                 "p99_ms": float(np.percentile(plat, 99)),
                 "f_ms_per_step": ptm["f_ms"] / psteps, "x_ms_per_step": ptm["x_ms"] / psteps,
                 "speedup_headline_vs_plain": value / pval,
                 "headline_vs_plain_rel_diff": rel}
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
        for name, mode in (("per_length", 1), ("persistent", 2)):
            ucfg = uneven_config("fp16", fengine_mode=mode)
            print(f"[bench] uneven bands ({name} F-engine launch), 1 warmup + 3 steps", file=sys.stderr, flush=True)
            u = time_path(ucfg, 3, 1, torch, dev, offsets)
            ulat = np.array(u["lat_ms"])
            forms[name] = {"value": usamp * 3 / u["wall_s"] / 1e9, "p50_ms": float(np.percentile(ulat, 50)),
                           "p99_ms": float(np.percentile(ulat, 99)),
                           "f_ms_per_step": u["timing"]["f_ms"] / 3, "x_ms_per_step": u["timing"]["x_ms"] / 3}
        best = max(forms, key=lambda n: forms[n]["value"])
        uneven = {"workload": "16 stations dual-pol: 4 x 64 MHz (nchan 16384), 4 x 32 MHz (nchan 8192), "
                              "8 x 8 MHz (nchan 2048) datastreams + 6 zooms of 0.5-2 MHz (128-512 channels), "
                              "0.512 s subint",
                  "value": forms[best]["value"], "unit": "Gsamples/s", "p99_ms": forms[best]["p99_ms"],
                  "subint_length_ms": 512.0, "forms": forms, "autotune_report": report}
    target = None
    if not args.no_target:
        from difx_amd import config as cf
        freqs = [4800.0 + 40.0 * k for k in range(8)]
        tg = {}
        for xe, steps_t in (("fp16", 3), ("fp32", 2)):
            tcfg = cf.standard_config(16, 2, freqs, 32.0, 1024, subint_ns=500_000_000, xengine=xe, seed=20261205)
            print(f"[bench] 16-station target shape, {xe}, 1 warmup + {steps_t} steps", file=sys.stderr, flush=True)
            r = time_path(tcfg, steps_t, 1, torch, dev, offsets)
            tl = np.array(r["lat_ms"])
            tg[xe] = {"value": samples_per_subint(tcfg) * steps_t / r["wall_s"] / 1e9,
                      "p99_ms": float(np.percentile(tl, 99)),
                      "f_ms_per_step": r["timing"]["f_ms"] / steps_t, "x_ms_per_step": r["timing"]["x_ms"] / steps_t}
        target = {"workload": "16 stations dual-pol, 8 x 32 MHz, nchan 1024, 0.5 s subint",
                  "mfma_fp16": tg["fp16"], "plain_fp32": tg["fp32"],
                  "speedup_mfma_vs_plain": tg["fp16"]["value"] / tg["fp32"]["value"],
                  "subint_length_ms": 500.0}
    serving = None
    if not args.no_serving:
        print("[bench] serving: Poisson job stream", file=sys.stderr, flush=True)
        serving = serving_bench(torch, dev)
    kurt = None
    if not args.no_kurtosis:
        print("[bench] headline path with spectral kurtosis, 1 warmup + 3 steps", file=sys.stderr, flush=True)
        kr = time_path(cfg, 3, 1, torch, dev, offsets, kurtosis=True)
        k_ms = kr["wall_s"] * 1e3 / 3
        spec_bytes = float(samples_per_subint(cfg)) / 2 * 4           # FP16 complex spectra: C = L/2 per FFT
        kurt = {"workload": "headline subint with Mode's kurtosis sums (|X|^2, |X|^4 per band and channel)",
                "ms_per_step": k_ms, "extra_ms_vs_headline": k_ms - head["wall_s"] * 1e3 / args.steps,
                "other_ms_per_step": kr["timing"]["other_ms"] / 3,
                "spectra_bytes_read_per_step": spec_bytes,
                "kurtosis_pass_GBps": spec_bytes / max((kr["timing"]["other_ms"] - tm["other_ms"] / args.steps * 3) / 3 / 1e3, 1e-9) / 1e9}
    pcal = None
    if not args.no_pcal:
        print("[bench] headline path with phase-cal extraction, 1 warmup + 3 steps", file=sys.stderr, flush=True)
        pr = time_path(cfg, 3, 1, torch, dev, offsets, pcal_hz=1_000_000)
        p_ms = pr["wall_s"] * 1e3 / 3
        extra_other = (pr["timing"]["other_ms"] - tm["other_ms"] / args.steps * 3) / 3
        raw_bytes = float(samples_per_subint(cfg)) * cfg.datastreams[0].nbits / 8   # sample bytes read once
        pcal = {"workload": "headline subint with phase-cal extraction, 1 MHz comb on every band",
                "ms_per_step": p_ms, "extra_ms_vs_headline": p_ms - head["wall_s"] * 1e3 / args.steps,
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
    cpu = None
    if not args.no_cpu:
        print("[bench] CPU baseline", file=sys.stderr, flush=True)
        cpu = cpu_baseline()
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
        "ms_per_step": head["wall_s"] * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 F-engine, fp16 X-engine inputs, f32 accumulate",
        "data": "synthetic (device-generated 2-bit VDIF, seeded quintic delay models)",
        "config": {"workload": "10 stations dual-pol, 8 x 32 MHz 2-bit VDIF, nchan 1024, 0.5 s subint",
                   "stations": 10, "pols": 2, "subbands": 8, "bw_mhz": 32, "nchan": 1024,
                   "subint_s": 0.5, "fft_blocks_per_subint": B, "parallelism": "single GPU"},
        "p50_subint_latency_ms": float(np.percentile(lat, 50)),
        "p99_subint_latency_ms": float(np.percentile(lat, 99)),
        "subint_length_ms": cfg.subint_ns / 1e6,
        "realtime_factor": value / realtime,
        "kernels": {"f_engine_ms_per_step": tm["f_ms"] / args.steps,
                    "x_engine_ms_per_step": tm["x_ms"] / args.steps,
                    "other_ms_per_step": tm["other_ms"] / args.steps,
                    "f_launches_per_step": nchunks_per_step,
                    "x_engine_useful_tflops": x_engine_flops(cfg, B) / (tm["x_ms"] / args.steps / 1e3) / 1e12},
        "roofline": {"kernel": "k_fengine<2048,false,1,256> (fused unpack+rotate+FFT+post-F)",
                     "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(ROOT / "profiles" / "k_fengine_2048_traffic.txt", ffts_per_launch),
                     "avg_launch_ms": f_avg_ms, "algorithmic_bytes_per_launch": fbytes},
        "plain_fp32": plain,
        "target_16_stations": target,
        "uneven_bands": uneven,
        "serving": serving,
        "vdif_corner_turn": mux,
        "output_side": output,
        "kurtosis": kurt,
        "phase_cal": pcal,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
