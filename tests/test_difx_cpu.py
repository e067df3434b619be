"""Visibility::writedata's scaling and the DIFX record writer on the CPU: the restatement's
closed forms (oracle/visout.py), and fxg_difx_records of a host-only layout context
(fxg_create_layout) byte for byte against the restatement, parsed back with the reference
reader's header layout."""
import ctypes as C
from pathlib import Path

import numpy as np
import pytest

import fxoracle as fo
import visout as vo
from fxcases import case_cfg0, case_cfg3

ROOT = Path(__file__).resolve().parent.parent


def _lib():
    if not (ROOT / "difx_amd" / "libfxg.so").exists():
        pytest.skip("libfxg.so not built")
    from difx_amd.core import load_library
    return load_library()


def _results(cfg, lay, rng, zero_some=True):
    """Random visibilities with positive weights in the weight slots (a few zero)."""
    res = (rng.standard_normal(lay.coreresultlength) + 1j * rng.standard_normal(lay.coreresultlength)).astype(np.complex64)
    f = res.view(np.float32)
    B = cfg.blockspersend()
    for (fi, j), off in lay.bweight_off.items():
        n = len(dict(cfg.baselines[j].freqs)[fi])
        f[2 * off:2 * off + n] = rng.uniform(0.5, 1.0, n).astype(np.float32) * B
    for d in range(len(cfg.datastreams)):
        end = lay.acweight_off[d + 1] if d + 1 < len(cfg.datastreams) else lay.coreresultlength
        f[2 * lay.acweight_off[d]:2 * end] = rng.uniform(0.5, 1.0, 2 * (end - lay.acweight_off[d])).astype(np.float32) * B
    if zero_some:
        (fi, j), off = next(iter(lay.bweight_off.items()))
        f[2 * off] = 0.0                                  # one product with no data
    return res


def test_scale_closed_form():
    """weight w: every channel of the product is multiplied by 1 / (w/(msi B) * msi * B * 2 *
    nchan) = 1 / (2 w nchan); weight 0 leaves it unscaled."""
    cfg = case_cfg0(B=128)
    lay = fo.result_layout(cfg)
    rng = np.random.default_rng(1)
    res = _results(cfg, lay, rng, zero_some=False)
    msi = 3.0
    f = res.view(np.float32)
    (fi, j), woff = next(iter(lay.bweight_off.items()))
    f[2 * woff] = 96.0 * msi * cfg.blockspersend() / cfg.blockspersend()    # stored weight: 96 * msi
    out = vo.normalise(cfg, res, lay, msi)
    nch = cfg.freqs[fi].nchan
    w = 96.0 * msi / (msi * cfg.blockspersend())
    want = res[lay.baseline_off[(fi, j)]] / (w * msi * cfg.blockspersend() * 2 * nch)
    assert abs(out[lay.baseline_off[(fi, j)]] - want) <= 1e-6 * abs(want)
    f[2 * woff] = 0.0
    out = vo.normalise(cfg, res, lay, msi)
    assert out[lay.baseline_off[(fi, j)]] == res[lay.baseline_off[(fi, j)]]


def test_zoom_factors():
    """A cross-correlation of two zoom bands carries sqrt(zoom bw / parent bw) per side, an
    autocorrelation of a zoom band the ratio itself (visibility.cpp:525-540, 615-620)."""
    cfg = case_cfg3(B=32)
    ds = cfg.datastreams[0]
    nrec = len(ds.recbands)
    for k in range(len(ds.zoombands)):
        zf, parent, _ = ds.zoomfreqs[ds.zoombands[k][0]]
        ratio = cfg.freqs[zf].bw_mhz / cfg.freqs[ds.recfreqs[parent]].bw_mhz
        assert abs(float(vo.zoom_ratio(cfg, ds, nrec + k)) - ratio) < 1e-7


def test_tsys_closed_form():
    """Tsys on both datastreams, constant parallel autocorrelations c with weights a: the
    cross scale is sqrt(T1 T2 a^2 / (d^2 c^2)) / w, d the 2-bit decorrelation 0.88."""
    cfg = case_cfg0(B=128)
    lay = fo.result_layout(cfg)
    rng = np.random.default_rng(2)
    res = _results(cfg, lay, rng, zero_some=False)
    f = res.view(np.float32)
    B = cfg.blockspersend()
    c, a = 4.0, 0.8 * B
    for d in range(len(cfg.datastreams)):
        off = lay.autocorr_off[d]
        end = lay.acweight_off[0] if d + 1 == len(cfg.datastreams) else lay.autocorr_off[d + 1]
        res[off:end] = c
        wend = lay.acweight_off[d + 1] if d + 1 < len(cfg.datastreams) else lay.coreresultlength
        f[2 * lay.acweight_off[d]:2 * wend] = a
    tsys = [50.0 + 30.0 * d for d in range(len(cfg.datastreams))]
    out = vo.normalise(cfg, res, lay, 1.0, tsys=tsys)
    (fi, j), boff = next(iter(lay.bweight_off.items()))
    T1, T2 = tsys[cfg.baselines[j].ds1], tsys[cfg.baselines[j].ds2]
    w = f[2 * boff] / B
    want = np.sqrt(T1 * T2 * (a / B) ** 2 / (0.88 ** 2 * c * c)) / w
    got = out[lay.baseline_off[(fi, j)]] / res[lay.baseline_off[(fi, j)]]
    assert abs(got - want) <= 2e-6 * want


def _layout_ctx(lib, cfg):
    fa, keep = cfg.to_abi()
    h = C.c_void_p()
    assert lib.fxg_create_layout(C.byref(fa), C.byref(h)) == 0
    return h, keep


def _meta(cfg, rng, ffts):
    from difx_amd import fxg_abi as abi
    nb, nds = len(cfg.baselines), len(cfg.datastreams)
    bl = np.array([256 * (b.ds1 + 1) + b.ds2 + 1 for b in cfg.baselines], dtype=np.int32)
    uvw = rng.uniform(-6e6, 6e6, (nb, 3))
    tel = np.arange(nds, dtype=np.int32) + 3
    m = abi.fxg_difx_meta(59000, 0, 2, 0, 43200.25, ffts, bl.ctypes.data_as(C.POINTER(C.c_int32)),
                          uvw.ctypes.data_as(C.POINTER(C.c_double)), tel.ctypes.data_as(C.POINTER(C.c_int32)))
    return m, (bl, uvw, tel)


@pytest.mark.parametrize("case", ["cfg0", "cfg3_zooms"])
def test_records_equal_the_restatement(case):
    """fxg_difx_records (host-only context) equals the restatement byte for byte; parsed back
    with the reference reader's layout, every header field and visibility is where it should
    be, weight-0 products are absent, and every record is 74 + 8 nchan bytes."""
    lib = _lib()
    cfg = case_cfg0(B=128) if case == "cfg0" else case_cfg3(B=32)
    lay = fo.result_layout(cfg)
    rng = np.random.default_rng(5)
    res = _results(cfg, lay, rng)
    ffts = 2.0 * cfg.blockspersend()
    h, keep = _layout_ctx(lib, cfg)
    try:
        m, arrs = _meta(cfg, rng, ffts)
        nbytes, nrec = C.c_int64(), C.c_int32()
        fres = np.ascontiguousarray(res.view(np.float32))
        assert lib.fxg_difx_records(h, fres.ctypes.data, C.byref(m), None, 0, C.byref(nbytes), C.byref(nrec)) == 0
        out = np.zeros(nbytes.value, dtype=np.uint8)
        assert lib.fxg_difx_records(h, fres.ctypes.data, C.byref(m), out.ctypes.data, out.size, C.byref(nbytes),
                                    C.byref(nrec)) == 0
        small = np.zeros(nbytes.value - 1, dtype=np.uint8)
        assert lib.fxg_difx_records(h, fres.ctypes.data, C.byref(m), small.ctypes.data, small.size,
                                    C.byref(nbytes), C.byref(nrec)) != 0
    finally:
        lib.fxg_destroy(h)
    bl, uvw, tel = arrs
    want = vo.records(cfg, res, lay, 59000, 43200.25, 0, 2, ffts, bl, uvw, tel)
    assert out.tobytes() == want
    nchan = {f: cfg.freqs[f].nchan // cfg.freqs[f].chanavg for f in range(len(cfg.freqs))}
    recs = vo.parse(out.tobytes(), nchan)
    assert len(recs) == nrec.value
    # walk the records in writer order: baseline, freq ascending, product; then the autos
    f32 = res.view(np.float32)
    it = iter(recs)
    for i, bsl in enumerate(cfg.baselines):
        for f in sorted(fi for fi, _ in bsl.freqs):
            prods = dict(bsl.freqs)[f]
            fc = nchan[f]
            for p, (b1, b2) in enumerate(prods):
                w = np.float32(f32[lay.bweight_off[(f, i)] * 2 + p] / ffts)
                if not w > 0:
                    continue
                r = next(it)
                pol = vo.band_pol(cfg.datastreams[bsl.ds1], b1) + vo.band_pol(cfg.datastreams[bsl.ds2], b2)
                assert (r["baseline"], r["freq"], r["pol"], r["weight"], r["uvw"]) == \
                    (bl[i], f, pol, float(w), tuple(uvw[i]))
                off = lay.baseline_off[(f, i)] + p * fc
                np.testing.assert_array_equal(r["vis"], res[off:off + fc])
    rest = list(it)
    assert all(r["baseline"] % 257 == 0 and r["uvw"] == (0.0, 0.0, 0.0) for r in rest)
    assert {r["baseline"] // 257 - 1 for r in rest} == set(int(t) for t in tel)
    for r in recs:
        assert (r["mjd"], r["seconds"], r["config"], r["source"], r["bin"]) == (59000, 43200.25, 0, 2, 0)
        assert r["weight"] > 0
    assert len(out) == sum(74 + 8 * len(r["vis"]) for r in recs)


def test_layout_context_refuses_processing():
    """A host-only context answers layout queries and refuses device work (FXG_E_STATE)."""
    lib = _lib()
    from difx_amd import fxg_abi as abi
    cfg = case_cfg0(B=128)
    h, keep = _layout_ctx(lib, cfg)
    try:
        lay = abi.fxg_layout()
        assert lib.fxg_get_layout(h, C.byref(lay)) == 0
        want = fo.result_layout(cfg)
        assert lay.coreresultlength == want.coreresultlength
        assert lib.fxg_normalise_device(h, C.c_void_p(16), 1.0, None, None) == abi.FXG_E_STATE
        out = (C.c_double * 3)()
        assert lib.fxg_delay_interpolator(h, 0, 0, 0, out) == abi.FXG_E_STATE
    finally:
        lib.fxg_destroy(h)


@pytest.mark.parametrize("dumpchannels", [16, 64, 4096])
def test_sta_autocorr_records(dumpchannels):
    """fxg_sta_autocorr (host-only context) equals the restatement of averageAndSendAutocorrs;
    for a flat autocorrelation of value c and weight w the record is cta c / (2 nchan w) in every
    channel (cta = channels per record channel); a weight below 0.333 of a full span gives no
    record."""
    lib = _lib()
    cfg = case_cfg3(B=32)
    lay = fo.result_layout(cfg)
    rng = np.random.default_rng(11)
    res = _results(cfg, lay, rng, zero_some=False)
    h, keep = _layout_ctx(lib, cfg)
    nswidth = float(cfg.subint_ns)
    try:
        for d, ds in enumerate(cfg.datastreams):
            for band in range(len(ds.recbands)):
                out = np.zeros(16384, dtype=np.float32)
                n = C.c_int32()
                f = res.view(np.float32)
                assert lib.fxg_sta_autocorr(h, f.ctypes.data, d, band, dumpchannels, nswidth,
                                            out.ctypes.data, C.byref(n)) == 0
                want = vo.sta_autocorr(cfg, res, lay, d, band, dumpchannels, nswidth)
                assert want is not None and n.value == want.size
                np.testing.assert_array_equal(out[:n.value], want)
        # closed form on band 0 of datastream 0: flat autocorrelation c, weight w
        ds = cfg.datastreams[0]
        fr = cfg.freqs[vo.band_freq(ds, 0)]
        off, widx = lay.autocorr_off[0], lay.acweight_off[0] * 2
        c, w = 3.0, 0.9 * fr.nchan * 2 * cfg.blockspersend() / (2 * fr.nchan)   # a full span's weight
        r2 = res.copy()
        r2[off:off + fr.nchan // fr.chanavg] = c
        r2.view(np.float32)[widx] = w
        out = np.zeros(16384, dtype=np.float32)
        n = C.c_int32()
        assert lib.fxg_sta_autocorr(h, r2.view(np.float32).ctypes.data, 0, 0, dumpchannels, nswidth,
                                    out.ctypes.data, C.byref(n)) == 0
        fc = fr.nchan // fr.chanavg
        nout = min(dumpchannels, fc)
        np.testing.assert_allclose(out[:n.value], (fc // nout) * c / (2 * fr.nchan * w), rtol=2e-6)
        r2.view(np.float32)[widx] = 0.3 * 0.001 * nswidth * 2 * fr.bw_mhz / (2 * fr.nchan)
        assert lib.fxg_sta_autocorr(h, r2.view(np.float32).ctypes.data, 0, 0, dumpchannels, nswidth,
                                    out.ctypes.data, C.byref(n)) == 0
        assert n.value == 0
    finally:
        lib.fxg_destroy(h)


def test_phase_centre_layout_records_and_scale():
    """Several phase centres: the library's layout equals the restatement's; each phase
    centre's records (its own file) equal the restatement's byte for byte, with the weights
    times the shift decorrelation and the autocorrelations only in the first file; the scale of
    a phase centre's block is divided by its decorrelation over the integration."""
    from fxcases import with_phase_centres
    lib = _lib()
    cfg = with_phase_centres(case_cfg3(B=32), P=3, interval_ffts=8, threads=2)
    lay = fo.result_layout(cfg)
    rng = np.random.default_rng(7)
    res = _results(cfg, lay, rng)
    f32 = res.view(np.float32)
    msi = 2.0
    inttime_ns = msi * cfg.subint_ns
    for (fi, j), so in lay.shiftdecorr_off.items():
        f32[2 * so: 2 * so + 3] = (rng.uniform(0.5, 1.0, 3) * inttime_ns).astype(np.float32)
    ffts = msi * cfg.blockspersend()
    h, keep = _layout_ctx(lib, cfg)
    try:
        from difx_amd import fxg_abi as abi
        lo = abi.fxg_layout()
        assert lib.fxg_get_layout(h, C.byref(lo)) == 0
        assert lo.coreresultlength == lay.coreresultlength
        fres = np.ascontiguousarray(res.view(np.float32))
        outs = []
        for pc in range(3):
            m, arrs = _meta(cfg, rng, ffts)
            m.phasecentre = pc
            nbytes, nrec = C.c_int64(), C.c_int32()
            assert lib.fxg_difx_records(h, fres.ctypes.data, C.byref(m), None, 0, C.byref(nbytes), C.byref(nrec)) == 0
            out = np.zeros(nbytes.value, dtype=np.uint8)
            assert lib.fxg_difx_records(h, fres.ctypes.data, C.byref(m), out.ctypes.data, out.size,
                                        C.byref(nbytes), C.byref(nrec)) == 0
            bl, uvw, tel = arrs
            assert out.tobytes() == vo.records(cfg, res, lay, 59000, 43200.25, 0, 2, ffts, bl, uvw, tel, phasecentre=pc)
            outs.append(vo.parse(out.tobytes(), {f: cfg.freqs[f].nchan // cfg.freqs[f].chanavg
                                                 for f in range(len(cfg.freqs))}))
        m.phasecentre = 3
        assert lib.fxg_difx_records(h, fres.ctypes.data, C.byref(m), None, 0, C.byref(nbytes), C.byref(nrec)) != 0
    finally:
        lib.fxg_destroy(h)
    autos = [sum(1 for r in recs if r["baseline"] % 257 == 0) for recs in outs]
    assert autos[0] > 0 and autos[1] == 0 and autos[2] == 0
    # record weights: baseline weight x decorrelation / (1e9 x integration time)
    (fi, j), so = next(iter(lay.shiftdecorr_off.items()))
    w = np.float32(f32[lay.bweight_off[(fi, j)] * 2 + 1] / ffts)
    for pc in range(3):
        dec = np.float32(f32[2 * so + pc] / inttime_ns)
        assert any(r["weight"] == float(np.float32(w * dec)) for r in outs[pc])
    # scale: the phase-centre block is the single-centre scale divided by its decorrelation
    out = vo.normalise(cfg, res, lay, msi)
    fr = cfg.freqs[fi]
    fc = (fr.nchan + fr.chanavg - 1) // fr.chanavg
    nprod = len(dict(cfg.baselines[j].freqs)[fi])
    for pc in range(3):
        o = lay.baseline_off[(fi, j)] + (pc * nprod + 1) * fc
        wf = np.float32(f32[lay.bweight_off[(fi, j)] * 2 + 1] / ffts)
        sc = np.float32(1.0 / (np.float64(wf) * msi * np.float64(np.float32(cfg.blockspersend() * 2 * fc * fr.chanavg))))
        dec = np.float32(f32[2 * so + pc] / (1e9 * (msi * cfg.subint_ns / 1e9)))
        np.testing.assert_array_equal(out[o:o + fc], (res[o:o + fc].view(np.float32) * np.float32(sc / dec)).view(np.complex64))


def test_usb_lsb_layout_and_autocorr_labels():
    """USB x LSB products (configuration.cpp:3125-3135) on a host-only context: the layout
    equals the restatement's, and the records equal it byte for byte.  Worked by hand: the
    LSB freq 1 (edge 4832 MHz, 32 MHz) covers the sky band of the USB freq 0 (edge 4800 MHz)
    and is correlated against it, so stations 2 and 3's autocorrelations are written under
    freq 0 (visibility.cpp:912-925); the LSB x LSB baseline (2, 3) keeps freq 1."""
    from fxcases import case_usb_lsb
    lib = _lib()
    cfg = case_usb_lsb(B=32)
    assert fo.autocorr_freq_labels(cfg) == [0, 0]
    assert fo.opposite_sideband_freq(cfg, 1) == 0 and fo.opposite_sideband_freq(cfg, 0) == 1
    lay = fo.result_layout(cfg)
    rng = np.random.default_rng(9)
    res = _results(cfg, lay, rng, zero_some=False)
    ffts = 2.0 * cfg.blockspersend()
    h, keep = _layout_ctx(lib, cfg)
    try:
        from difx_amd import fxg_abi as abi
        lt = abi.fxg_layout()
        assert lib.fxg_get_layout(h, C.byref(lt)) == 0
        assert lt.coreresultlength == lay.coreresultlength
        m, arrs = _meta(cfg, rng, ffts)
        nbytes, nrec = C.c_int64(), C.c_int32()
        fres = np.ascontiguousarray(res.view(np.float32))
        assert lib.fxg_difx_records(h, fres.ctypes.data, C.byref(m), None, 0, C.byref(nbytes), C.byref(nrec)) == 0
        out = np.zeros(nbytes.value, dtype=np.uint8)
        assert lib.fxg_difx_records(h, fres.ctypes.data, C.byref(m), out.ctypes.data, out.size, C.byref(nbytes),
                                    C.byref(nrec)) == 0
    finally:
        lib.fxg_destroy(h)
    bl, uvw, tel = arrs
    assert out.tobytes() == vo.records(cfg, res, lay, 59000, 43200.25, 0, 2, ffts, bl, uvw, tel)
    recs = vo.parse(out.tobytes(), {0: 256, 1: 256})
    autos = [r for r in recs if r["baseline"] % 257 == 0]
    assert autos and all(r["freq"] == 0 for r in autos)
    cross = {(r["baseline"], r["freq"]) for r in recs if r["baseline"] % 257 != 0}
    assert (256 * 3 + 4, 1) in cross and (256 * 1 + 3, 0) in cross


@pytest.mark.parametrize("change", ["bandwidth", "edge", "nchan", "first_side", "dsb"])
def test_usb_lsb_mismatches_rejected(change):
    """A product across freqs is accepted only for the opposite sideband of the same sky band,
    recorded on the second datastream, from single-sideband data."""
    from fxcases import case_usb_lsb
    lib = _lib()
    cfg = case_usb_lsb(B=32)
    f1 = cfg.freqs[1]
    if change == "bandwidth":
        f1.bw_mhz, f1.nchan = 16.0, 128
    elif change == "edge":
        f1.freq_mhz = 4840.0
    elif change == "nchan":
        f1.nchan = 512
    elif change == "first_side":
        b = cfg.baselines[1]                       # (0, 2) on freq 0: make it (2, 0) on freq 1
        b.ds1, b.ds2 = b.ds2, b.ds1
        b.freqs = [(1, prods) for _, prods in b.freqs]
        cfg.baselines[1] = b
        for d in (0, 1):
            cfg.datastreams[d].recfreqs = [0]
        # the first datastream's band is of the baseline freq (1, LSB), the second's is USB: accepted
        fa, keep = cfg.to_abi()
        h = C.c_void_p()
        assert lib.fxg_create_layout(C.byref(fa), C.byref(h)) == 0
        lib.fxg_destroy(h)
        b.freqs = [(0, prods) for _, prods in b.freqs]   # baseline freq of the second datastream: refused
    else:
        for d in range(4):
            cfg.datastreams[d].complex_sampling = 2
            cfg.datastreams[d].nbits = 8
    fa, keep = cfg.to_abi()
    h = C.c_void_p()
    assert lib.fxg_create_layout(C.byref(fa), C.byref(h)) != 0
    assert b"opposite sideband" in lib.fxg_last_error(None)
