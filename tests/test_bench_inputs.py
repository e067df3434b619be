"""CPU check of the serving benchmark's shared input buffers: every station of a job reads the
same device bytes through its own control words (its own buffer start for its delay of up to
+-10 ms); each station's every FFT window must lie inside the buffer, or the benchmark would
skip work and overstate its throughput."""
import numpy as np

import fxoracle as fo


def test_serving_inputs_cover_every_fft():
    import bench
    from difx_amd import synth
    rng = np.random.default_rng(20261204)
    jobs = bench.serving_jobs(rng, 6.0, 4.0, 8)
    assert len(jobs) > 5
    checked = 0
    for t0, cfg, nsub, shape in jobs:
        B = cfg.blockspersend()
        L = cfg.fftchannels(0)
        ds0 = cfg.datastreams[0]
        spf = synth.samples_per_frame(ds0)
        for k in (0, nsub - 1):
            nbytes, ctrl = bench.serving_inputs(cfg, k)
            nsamples = (nbytes // ds0.framebytes) * spf
            offsets = [0, 1 + k, 0]
            f0 = cfg.freqs[ds0.recfreqs[0]]
            for d, ds in enumerate(cfg.datastreams):
                interp = fo.delay_interpolator(ds.model, ds.model_order, cfg.model_inc_secs,
                                               offsets[1] + offsets[2] / 1e9, B * f0.nchan / f0.bw_mhz / 1e6, B)
                interp = interp.copy()
                interp[2] -= 1e6 * fo.intclockseconds(ds)
                ns = fo.fft_params(cfg, d, offsets, ctrl[d], interp, B)[0]
                assert ns.min() >= 0 and ns.max() + L <= nsamples, (shape, k, d, ns.min(), ns.max() + L, nsamples)
                checked += 1
    assert checked > 50
