"""Device parity of the multi-phase-centre uv shift (Core::uvshiftAndAverageBaselineFreq,
core.cpp:1605-1938) against the oracle, through the C ABI: every phase centre's cross
correlations (FP32 within 1e-5, FP16 within 1e-3 per record in the diffDiFX metric), the
shift-decorrelation words, the autocorrelations and weights, and the device normalisation
of the phase-centre blocks against its restatement."""
import sys

import numpy as np
import pytest

import fxoracle as fo
import visout as vo
from fxcases import case_cfg1, case_cfg3, record_errors, slot, with_phase_centres

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _case(kind, xe):
    if kind == "dualpol_lsb":
        cfg = case_cfg1(B=40, nstations=4, nfreq=2, xengine=xe)
        cfg.freqs[1].lsb = True
        for f in cfg.freqs:
            f.chanavg = 2
        return with_phase_centres(cfg, P=3, interval_ffts=8, threads=3)
    return with_phase_centres(case_cfg3(B=32, xengine=xe), P=2, interval_ffts=8, threads=2)


@pytest.mark.parametrize("xe,bound", [("fp32", 1e-5), ("fp16", 1e-3)])
@pytest.mark.parametrize("kind", ["dualpol_lsb", "uneven_zooms"])
def test_uvshift_parity(gpu, kind, xe, bound):
    from difx_amd.core import FxCore
    _say("uv shift case", kind, xe)
    cfg = _case(kind, xe)
    lay = fo.result_layout(cfg)
    data, ctrl = slot(cfg, seed=41)
    with FxCore(cfg) as core:
        assert core.coreresultlength == lay.coreresultlength
        got = core.processdata((0, 1, 0), data, ctrl)
        import torch
        d = torch.from_numpy(got.view(np.float32).copy()).to(gpu)
        core.normalise_device(d.data_ptr(), 1.0)
        scaled = d.cpu().numpy().view(np.complex64)
    ref, _ = fo.process_subint(cfg, [0, 1, 0], data, ctrl)
    e = record_errors(cfg, lay, got.astype(np.complex128), ref)
    _say(f"uv shift {kind} {xe}: {len(e['cross'])} cross records, max", max(e["cross"]),
         "auto max", max(e["auto"]))
    assert max(e["cross"]) < bound and max(e["auto"]) < bound
    gf, rf = got.view(np.float32), ref.view(np.float64)
    for (fi, j), so in lay.shiftdecorr_off.items():
        np.testing.assert_allclose(gf[2 * so: 2 * so + cfg.nphasecentres], rf[2 * so: 2 * so + cfg.nphasecentres],
                                   rtol=1e-6)
    for (fi, j), wo in lay.bweight_off.items():
        np.testing.assert_allclose(gf[2 * wo], rf[2 * wo], rtol=1e-6)
    # the scaling of the phase-centre blocks equals its restatement on the device's results
    want = vo.normalise(cfg, got, lay, 1.0)
    np.testing.assert_array_equal(scaled.view(np.uint64), want.view(np.uint64))


def test_single_phase_centre_unchanged(gpu):
    """One phase centre in the phase-centre form (P = 1 with the interval fields set) gives the
    results of the plain form bit for bit."""
    from difx_amd.core import FxCore
    cfg = case_cfg1(B=32, nstations=3, nfreq=2)
    data, ctrl = slot(cfg, seed=42)
    with FxCore(cfg) as core:
        a = core.processdata((0, 1, 0), data, ctrl)
    cfg.process_threads = 3
    cfg.max_ns_between_xc_avg = 1000
    with FxCore(cfg) as core:
        b = core.processdata((0, 1, 0), data, ctrl)
    np.testing.assert_array_equal(a.view(np.uint64), b.view(np.uint64))
