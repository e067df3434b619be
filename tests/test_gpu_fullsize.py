"""The 16-station target shape at its full size (16 stations dual-pol, 8 x 32 MHz 2-bit VDIF,
nchan 1024, one 0.5 s subint of 15625 FFT blocks, data generated on the device), checked
through properties that hold at any size instead of against the oracle, which would take hours
at this size:

- station 15 records exactly what station 0 records (same delay model, same bytes), so each
  parallel-hand product of baseline (0, 15) equals station 0's autocorrelation of that band
  (the X-engine forms both from the same spectra in the same order);
- autocorrelations are real and positive;
- the FP16 and BF16 MFMA X-engines agree with the FP32 X-engine per record within their bars
  (1e-3 and 1e-2, mean |dV| / mean |V|);
- every FFT's data weight is 1 (every frame valid).
"""
import sys

import numpy as np
import pytest

import fxoracle as fo
from fxcases import record_errors

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _twin(cfg):
    cfg.datastreams[15].model = cfg.datastreams[0].model.copy()
    return cfg


def test_target16_full_subint_properties(gpu):
    import torch
    import bench
    from difx_amd.core import FxCore
    offsets = [0, 1, 0]
    base = _twin(bench.workload_config("fp32"))
    B = base.blockspersend()
    assert B == 15625
    bufs, nbytes, ctrl = bench.device_slot(base, offsets, torch, gpu, seed=61)
    bufs[15], nbytes[15], ctrl[15] = bufs[0], nbytes[0], ctrl[0]
    ptrs = [b.data_ptr() for b in bufs]
    lay = fo.result_layout(base)
    res = {}
    for xe in ("fp32", "fp16", "bf16"):
        _say("full-size 16 stations:", xe)
        with FxCore(_twin(bench.workload_config(xe))) as core:
            assert core.coreresultlength == lay.coreresultlength
            out = torch.zeros(2 * core.coreresultlength, dtype=torch.float32, device=gpu)
            core.processdata_device(offsets, ptrs, nbytes, ctrl, out.data_ptr(), 0)
            torch.cuda.synchronize()
            res[xe] = out.cpu().numpy().view(np.complex64).astype(np.complex128)
            if xe == "fp32":
                for d in (0, 7, 15):
                    w = core.getDataWeights(d)
                    assert w.shape[0] == B and np.all(w == 1.0), d
    C = 1024
    r32 = res["fp32"]
    j = [k for k, b in enumerate(base.baselines) if (b.ds1, b.ds2) == (0, 15)][0]
    worst = 0.0
    for fi in range(8):
        off = lay.baseline_off[(fi, j)]
        for pol in range(2):                     # RR, LL: products 0 and 1, bands 2 fi and 2 fi + 1
            cross = r32[off + pol * C: off + (pol + 1) * C]
            auto = r32[lay.autocorr_off[0] + (2 * fi + pol) * C: lay.autocorr_off[0] + (2 * fi + pol + 1) * C]
            worst = max(worst, float(np.max(np.abs(cross - auto) / np.abs(auto))))
    _say("baseline (0, 15) parallel hands vs autocorrelation of station 0: max rel diff", worst)
    assert worst < 1e-6
    for d in (0, 15):
        a = r32[lay.autocorr_off[d]: lay.autocorr_off[d] + 16 * C]
        assert np.all(a.real > 0) and np.max(np.abs(a.imag)) <= 1e-6 * np.max(a.real)
    for xe, bound in (("fp16", 1e-3), ("bf16", 1e-2)):
        e = record_errors(base, lay, res[xe], r32)
        _say(f"full-size {xe} vs FP32 X-engine: {len(e['cross'])} cross records, max {max(e['cross']):.3e}, "
             f"median {float(np.median(e['cross'])):.3e}; autos max {max(e['auto']):.3e}")
        assert max(e["cross"]) < bound and max(e["auto"]) < bound
