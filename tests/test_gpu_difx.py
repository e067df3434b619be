"""Visibility::writedata's scaling on the device (fxg_normalise_device) against the
restatement (oracle/visout.py) on real slot results, and the DIFX records of the scaled
integration (fxg_difx_records of the device context).  Bars: without Tsys the scaled values
are bit-exact (the scale is formed in the reference's float/double order and applied as one
float multiply); with Tsys within 2e-6 relative (the band means are sums in another order)."""
import sys

import numpy as np
import pytest

import fxoracle as fo
import visout as vo
from fxcases import case_cfg1, case_cfg3, slot

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _results(cfg, seeds):
    from difx_amd.core import FxCore
    core = FxCore(cfg)
    tot = None
    for sd in seeds:
        data, ctrl = slot(cfg, seed=sd)
        r = core.processdata([0, 1, 0], data, ctrl)
        tot = r if tot is None else tot + r
    return core, tot


@pytest.mark.parametrize("case,tsys", [("cfg3_zooms", False), ("cfg3_zooms", True), ("cfg1", False), ("cfg1", True)])
def test_normalise_device_matches_the_restatement(gpu, case, tsys):
    import torch
    _say("normalise", case, "tsys" if tsys else "no tsys")
    cfg = case_cfg3(B=32) if case == "cfg3_zooms" else case_cfg1(B=32, nstations=4, nfreq=2)
    core, res = _results(cfg, [21, 22])                     # an integration of two slots
    lay = fo.result_layout(cfg)
    t = [40.0 + 15.0 * d for d in range(len(cfg.datastreams))] if tsys else None
    d = torch.from_numpy(res.view(np.float32).copy()).to(gpu)
    core.normalise_device(d.data_ptr(), 2.0, tsys=t)
    got = d.cpu().numpy().view(np.complex64)
    want = vo.normalise(cfg, res, lay, 2.0, tsys=t)
    assert not np.array_equal(want, res)
    if tsys:
        np.testing.assert_allclose(got.view(np.float32), want.view(np.float32), rtol=2e-6, atol=0)
    else:
        np.testing.assert_array_equal(got, want)


def test_integration_records(gpu):
    """Two slots integrated, scaled on the device, written as DIFX records: equal to the
    restatement's records of the restatement's scaling, byte for byte."""
    import torch
    _say("difx records")
    cfg = case_cfg3(B=32)
    core, res = _results(cfg, [31, 32])
    lay = fo.result_layout(cfg)
    d = torch.from_numpy(res.view(np.float32).copy()).to(gpu)
    core.normalise_device(d.data_ptr(), 2.0)
    scaled = d.cpu().numpy().view(np.complex64)
    rng = np.random.default_rng(3)
    bl = [256 * (b.ds1 + 1) + b.ds2 + 1 for b in cfg.baselines]
    uvw = rng.uniform(-1e6, 1e6, (len(cfg.baselines), 3))
    tel = list(range(len(cfg.datastreams)))
    ffts = 2.0 * cfg.blockspersend()
    got = core.difx_records(scaled, 59500, 3600.5, 0, 0, ffts, bl, uvw, tel)
    want = vo.records(cfg, vo.normalise(cfg, res, lay, 2.0), lay, 59500, 3600.5, 0, 0, ffts, bl, uvw, tel)
    assert got == want
    nchan = {f: cfg.freqs[f].nchan // cfg.freqs[f].chanavg for f in range(len(cfg.freqs))}
    recs = vo.parse(got, nchan)
    autos = [r for r in recs if r["baseline"] % 257 == 0]
    # normalised parallel-hand autocorrelations of unit-variance noise quantised to 2 bits at
    # +-0.95 sigma with levels +-1, +-3.3359: the mean power of the levels,
    # p + (1 - p) 3.3359^2 with p = erf(0.95 / sqrt 2)
    par = [r for r in autos if r["pol"][0] == r["pol"][1]]
    assert par and all(np.isfinite(r["vis"]).all() for r in recs)
    from math import erf, sqrt
    p = erf(0.95 / sqrt(2.0))
    power = p + (1 - p) * 3.3359 ** 2
    mean = float(np.mean([np.mean(r["vis"].real) for r in par]))
    _say("  records", len(recs), "autos", len(autos), "mean parallel auto", mean, "expected", power)
    assert abs(mean - power) < 5e-3 * power
