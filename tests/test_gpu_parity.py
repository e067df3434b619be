"""Device parity of the FX path against the FP64 oracle, through the C ABI.

Bars: unpack bit-exact; delay interpolator to FP64 rounding; FP32 path visibilities within
1e-5 (diffDiFX convention, mean|dV|/mean|V| per record); FP16 MFMA path within 1e-3 and
BF16 within 1e-2 in the same metric; weights exact to FP32 summation.
"""
import sys

import numpy as np
import pytest

import fxoracle as fo
from difx_amd import synth
from fxcases import (case_cfg0, case_cfg1, case_cfg2, case_cfg3, case_complex_fourstep,
                     case_fourstep_ffts, case_long_ffts, case_mark5b, case_offsets, case_real8,
                     case_offsets_fourstep, case_dsb_fraclo,
                     damaged_slot, mark5b_partial_slot, record_errors, slot, channel_errors)

pytestmark = pytest.mark.gpu


def _core(cfg):
    from difx_amd.core import FxCore
    return FxCore(cfg)


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("kind", ["vdif1", "vdif16", "mark5b8", "mark5b_zones", "cplx8"])
def test_unpack_bitexact(gpu, kind):
    import torch
    if kind == "mark5b_zones":
        cfg = case_mark5b(B=32)
    elif kind == "vdif1":
        cfg = case_cfg0(B=128)
    elif kind == "vdif16":
        cfg = case_cfg1(B=32, nstations=2)
    elif kind == "mark5b8":
        cfg = case_cfg1(B=32, nstations=2, nfreq=4)
        for d in cfg.datastreams:
            d.format, d.headerbytes, d.payloadbytes = "MARK5B", 16, 10000
    else:
        cfg = case_cfg2(B=16, nstations=2)
    _say("unpack case", kind)
    core = _core(cfg)
    ds = cfg.datastreams[0]
    fb, hb = ds.framebytes, ds.headerbytes
    if kind == "mark5b_zones":
        # station 1's frames with fill-pattern zones (tail, head, straggler, both ends, middle)
        data, ctrl = mark5b_partial_slot(cfg, seed=5)
        buf = data[1].copy()
    else:
        data, ctrl = slot(cfg, seed=5)
        buf = data[0].copy()
        # invalidate frame 1 (invalid bit / Mark5B flag) and put fill pattern in frame 2's payload
        if ds.format == "VDIF":
            buf[fb + 3] |= 0x80
        else:
            buf[fb + 5] |= 0x80
        fill = np.frombuffer(np.uint64(0x1122334411223344).tobytes(), np.uint8)
        buf[2 * fb + hb: 2 * fb + hb + 8] = fill
    vals, valid = fo.decode_stream(buf, ds)
    nsamp = vals.shape[1]
    start, n = 3, (nsamp - 3 if kind == "mark5b_zones" else min(nsamp - 3, 50000))
    d_buf = torch.from_numpy(buf).to(gpu)
    nb = len(ds.recbands)
    out = torch.zeros(nb * n * (2 if ds.complex_sampling else 1), dtype=torch.float32, device=gpu)
    good = torch.zeros(1, dtype=torch.int32, device=gpu)
    core.unpack_device(0, d_buf.data_ptr(), len(buf), start, n, out.data_ptr(), good.data_ptr())
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    if ds.complex_sampling:
        got = got.view(np.complex64).reshape(nb, n)
    else:
        got = got.reshape(nb, n)
    ref = vals[:, start: start + n]
    assert got.dtype == ref.dtype
    np.testing.assert_array_equal(got.view(np.uint32) if not ds.complex_sampling else got.view(np.uint64),
                                  ref.view(np.uint32) if not ds.complex_sampling else ref.view(np.uint64))
    assert int(good.cpu()[0]) == int(valid[start: start + n].sum())
    core.close()


def test_delay_interpolator(gpu):
    cfg = case_cfg0(B=128)
    core = _core(cfg)
    B = cfg.blockspersend()
    for d, ds in enumerate(cfg.datastreams):
        for sec, ns in [(1, 0), (7, 123456789), (60, 500000000)]:
            got = core.calculateDelayInterpolator(d, sec, ns)
            C = cfg.freqs[ds.recfreqs[0]].nchan
            ref = fo.delay_interpolator(ds.model, ds.model_order, cfg.model_inc_secs,
                                        sec + ns / 1e9, B * C / cfg.freqs[0].bw_mhz / 1e6, B)
            ref[2] -= 1e6 * fo.intclockseconds(ds)
            # a0 and a1 come from differences of delays of size |d0|: FP64 rounding of those
            # differences bounds the agreement (scale-aware absolute tolerance)
            tol = 1e-13 * max(abs(ref[2]), 1.0) * np.array([1.0 / B ** 2, 1.0 / B, 1.0])
            assert np.all(np.abs(got - ref) <= tol), (got, ref)
    core.close()


def _run_and_compare(cfg, seed=11, offsets=(0, 1, 0), spectra_check=True, damaged=False, make_slot=None):
    core = _core(cfg)
    if make_slot is None:
        make_slot = damaged_slot if damaged else slot
    data, ctrl = make_slot(cfg, offsets, seed=seed)
    got = core.processdata(offsets, data, ctrl)
    ref, stations = fo.process_subint(cfg, list(offsets), data, ctrl)
    lay = fo.result_layout(cfg)
    assert core.coreresultlength == lay.coreresultlength
    b, w, a, aw = core.result_offsets()
    for (fi, j), off in lay.baseline_off.items():
        assert b[fi, j] == off
    assert list(a) == lay.autocorr_off and list(aw) == lay.acweight_off
    # per-FFT data weights
    for d in range(len(cfg.datastreams)):
        np.testing.assert_allclose(core.getDataWeights(d), stations[d].weight, rtol=0, atol=1e-7)
    if spectra_check and cfg.xengine == "fp32":
        B = cfg.blockspersend()
        for d in range(min(2, len(cfg.datastreams))):
            for i in (0, B // 2, B - 1):
                F = core.getFreqs(d, 0, i)
                R = stations[d].spectra[0, i]
                err = np.max(np.abs(F - R)) / max(np.max(np.abs(R)), 1e-30)
                assert err < 2e-5, (d, i, err)
    e = record_errors(cfg, lay, got.astype(np.complex128), ref)
    core.close()
    return got, ref, e, lay


def test_fp32_cfg0(gpu):
    cfg = case_cfg0(B=128)
    got, ref, e, lay = _run_and_compare(cfg)
    _say("cfg0 fp32 cross max", max(e["cross"]), "auto max", max(e["auto"]))
    assert max(e["cross"]) < 1e-5
    assert max(e["auto"]) < 1e-5
    # weights: all FFTs valid -> bweight = B, acweight = B
    fres = got.view(np.float32)
    B = cfg.blockspersend()
    for (fi, j), off in lay.bweight_off.items():
        assert abs(fres[2 * off] - B) < 1e-3


def test_fp32_cfg0_lsb(gpu):
    cfg = case_cfg0(B=64, lsb=True)
    got, ref, e, lay = _run_and_compare(cfg)
    _say("cfg0 lsb fp32 cross max", max(e["cross"]))
    assert max(e["cross"]) < 1e-5 and max(e["auto"]) < 1e-5


def test_fp32_cfg1(gpu):
    cfg = case_cfg1(B=32)
    got, ref, e, lay = _run_and_compare(cfg)
    _say("cfg1 fp32 cross max", max(e["cross"]), "auto max", max(e["auto"]))
    assert max(e["cross"]) < 1e-5 and max(e["auto"]) < 1e-5


@pytest.mark.parametrize("xe,bound", [("fp16", 1e-3), ("bf16", 1e-2)])
def test_mfma_cfg1(gpu, xe, bound):
    cfg = case_cfg1(B=32, xengine=xe)
    got, ref, e, lay = _run_and_compare(cfg)
    _say(f"cfg1 {xe} cross max", max(e["cross"]), "median", float(np.median(e["cross"])),
         "auto max", max(e["auto"]))
    assert max(e["cross"]) < bound and max(e["auto"]) < bound


@pytest.mark.parametrize("xe,bound", [("fp16", 1e-3), ("bf16", 1e-2)])
def test_mfma_cfg2(gpu, xe, bound):
    """64 stations, complex 8-bit: record errors, and per-channel amplitude and phase errors
    |dV| / sqrt(A_i A_j) and arg(V conj(V_ref)) (median and max over every channel of every
    baseline)."""
    cfg = case_cfg2(B=16, xengine=xe)
    got, ref, e, lay = _run_and_compare(cfg)
    norm, dphi = channel_errors(cfg, lay, got.astype(np.complex128), ref)
    _say(f"cfg2 {xe} cross max", max(e["cross"]), "median", float(np.median(e["cross"])),
         "auto max", max(e["auto"]), "|dV|/sqrt(AiAj) median", float(np.median(norm)), "max",
         float(np.max(norm)), "dphase deg median", float(np.median(dphi)), "max", float(np.max(dphi)))
    assert max(e["cross"]) < bound and max(e["auto"]) < bound


def test_fp32_cfg2(gpu):
    cfg = case_cfg2(B=8, xengine="fp32", nstations=8)
    got, ref, e, lay = _run_and_compare(cfg)
    _say("cfg2 fp32 cross max", max(e["cross"]))
    assert max(e["cross"]) < 1e-5 and max(e["auto"]) < 1e-5


def test_fp32_device_vs_c_port(gpu):
    """Both FP32: the device path and the C restatement of the reference's FP32 arithmetic."""
    import fxport
    if not fxport.available():
        pytest.skip("libfxport.so not built")
    cfg = case_cfg1(B=64, nstations=6)
    core = _core(cfg)
    data, ctrl = slot(cfg, seed=21)
    got = core.processdata((0, 1, 0), data, ctrl)
    ref, _ = fxport.process_subint(cfg, [0, 1, 0], data, ctrl, nthreads=4)
    e = record_errors(cfg, fo.result_layout(cfg), got.astype(np.complex128), ref.astype(np.complex128))
    _say("device fp32 vs C port: cross max", max(e["cross"]), "auto max", max(e["auto"]))
    assert max(e["cross"]) < 1e-5 and max(e["auto"]) < 1e-5
    core.close()


@pytest.mark.parametrize("xe,bound", [("fp32", 1e-5), ("fp16", 1e-3)])
def test_uneven_bands_and_zooms(gpu, xe, bound):
    """FFT lengths 256, 1024 and 2048 in one configuration, with zoom windows correlated on
    every baseline (the MFMA X-engine takes them as rows of the zoom freq's group)."""
    cfg = case_cfg3(B=32, xengine=xe)
    got, ref, e, lay = _run_and_compare(cfg)
    _say(f"cfg3 {xe} cross max", max(e["cross"]), "auto max", max(e["auto"]))
    assert max(e["cross"]) < bound and max(e["auto"]) < bound


@pytest.mark.parametrize("xe,bound", [("fp32", 1e-5), ("fp16", 1e-3)])
def test_long_ffts(gpu, xe, bound):
    """FFT lengths 4096 and 8192 (1024-thread workgroups) with zoom windows."""
    cfg = case_long_ffts(B=8, xengine=xe)
    got, ref, e, lay = _run_and_compare(cfg)
    _say(f"long ffts {xe} cross max", max(e["cross"]), "auto max", max(e["auto"]))
    assert max(e["cross"]) < bound and max(e["auto"]) < bound


@pytest.mark.parametrize("order", [0, 2])
@pytest.mark.parametrize("xe,bound", [("fp32", 1e-5), ("fp16", 1e-3)])
def test_rotation_orders(gpu, order, xe, bound):
    """Post-F (order 0) and quadratic (order 2) fringe rotation, USB and LSB bands."""
    for lsb in (False, True):
        cfg = case_cfg0(B=64, xengine=xe, lsb=lsb, fringe_rot_order=order)
        got, ref, e, lay = _run_and_compare(cfg)
        _say(f"order {order} lsb {lsb} {xe} cross max", max(e["cross"]), "auto max", max(e["auto"]))
        assert max(e["cross"]) < bound and max(e["auto"]) < bound
    cfg = case_cfg1(B=16, nstations=4, nfreq=2, xengine=xe, fringe_rot_order=order)
    got, ref, e, lay = _run_and_compare(cfg)
    _say(f"order {order} cfg1 {xe} cross max", max(e["cross"]))
    assert max(e["cross"]) < bound and max(e["auto"]) < bound


@pytest.mark.parametrize("name", ["usb", "lsb", "offsets", "zero_delta", "mark5b_zones", "real8", "damaged",
                                  "long_ffts"])
@pytest.mark.parametrize("xe,bound", [("fp32", 1e-5), ("fp16", 1e-3)])
def test_order0_band_pairs(gpu, name, xe, bound):
    """Post-F rotation (order 0) of real data: the two pols of a freq go through one complex
    FFT as a + ib and are split after it.  Dual-pol USB and LSB freqs, pols with their own
    phasor tables (clock deltas and phase offsets: one FFT per band) and without (deltas and
    offsets zero: pairs), Mark5B fill-pattern zones, 8-bit samples, flagged FFTs and invalid
    frames, and 4096/8192-point FFTs with zooms (several butterfly positions per thread).
    Every band's FP32 spectra are checked, the pair's second band included."""
    from difx_amd import config as cf
    make_slot = None
    if name in ("usb", "lsb", "damaged"):
        cfg = case_cfg1(B=16, nstations=4, nfreq=2, xengine=xe, lsb=(name == "lsb"), fringe_rot_order=0)
    elif name in ("offsets", "zero_delta"):
        cfg = case_offsets(B=32, xengine=xe, fringe_rot_order=0)
        if name == "zero_delta":
            for d in cfg.datastreams:
                d.clock_offset_delta_us = [0.0 for _ in d.recfreqs]
                d.phase_offset_deg = [0.0 for _ in d.recfreqs]
    elif name == "mark5b_zones":
        cfg = case_mark5b(B=32, xengine=xe, fringe_rot_order=0)
        make_slot = mark5b_partial_slot
    elif name == "real8":
        cfg = case_real8(B=32, xengine=xe, fringe_rot_order=0)
    else:
        groups = [(64.0, 4096, [6000.0]), (32.0, 2048, [6100.0])]
        zooms = [(0, 0, 5.0, 1.0), (1, 0, 3.0, 1.0)]
        cfg = cf.batched_config(3, groups, zooms, subint_ns=8 * 64000, npol=2, xengine=xe, seed=6,
                                fringe_rot_order=0)
    got, ref, e, lay = _run_and_compare(cfg, damaged=(name == "damaged"), make_slot=make_slot,
                                        spectra_check=name not in ("offsets", "zero_delta"))
    _say(f"order 0 pairs {name} {xe} cross max", max(e["cross"]), "auto max", max(e["auto"]))
    assert max(e["cross"]) < bound and max(e["auto"]) < bound
    if xe != "fp32" or name in ("offsets", "zero_delta"):
        return
    core = _core(cfg)
    offsets = (0, 1, 0)
    data, ctrl = (make_slot or (damaged_slot if name == "damaged" else slot))(cfg, offsets, seed=11)
    core.processdata(offsets, data, ctrl)
    _, stations = fo.process_subint(cfg, list(offsets), data, ctrl)
    B = cfg.blockspersend()
    worst = 0.0
    for d in range(len(cfg.datastreams)):
        for band in range(len(cfg.datastreams[d].recbands)):
            for i in (0, B // 2, B - 1):
                F = core.getFreqs(d, band, i)
                R = stations[d].spectra[band, i]
                worst = max(worst, float(np.max(np.abs(F - R)) / max(np.max(np.abs(R)), 1e-30)))
    core.close()
    _say(f"order 0 pairs {name}: worst spectrum error", worst)
    assert worst < 2e-5


@pytest.mark.parametrize("nst,npol", [(5, 1), (32, 2), (36, 2)])
def test_plain_x_engine_group_sizes(gpu, nst, npol):
    """The blocked plain FP32 X-engine over groups of 5 rows (one partial 8 x 8 block), 64 rows
    (eight row blocks, 36 blocks per channel tile) and 72 rows (above its 64-row limit: every
    product of the group goes to the per-product kernel), nchan 128 (two 64-channel tiles)."""
    from difx_amd import config as cf
    cfg = cf.standard_config(nst, npol, [1600.0], 16.0, 128, subint_ns=8 * 8000, xengine="fp32", seed=30 + nst)
    got, ref, e, lay = _run_and_compare(cfg)
    _say(f"plain X {nst} stations x {npol} pol cross max", max(e["cross"]), "auto max", max(e["auto"]))
    assert max(e["cross"]) < 1e-5 and max(e["auto"]) < 1e-5


@pytest.mark.parametrize("xe,bound", [("fp32", 1e-5), ("fp16", 1e-3), ("bf16", 1e-2)])
def test_usb_lsb_products(gpu, xe, bound):
    """USB x LSB: two stations record a sky band as USB, two as LSB; the cross-sideband
    baselines multiply channel c of the USB band by channel c of the LSB band (both ascending
    in sky frequency, core.cpp:869-871).  The MFMA X-engine contracts both sidebands' rows in
    one group."""
    from fxcases import case_usb_lsb
    cfg = case_usb_lsb(B=32, xengine=xe)
    got, ref, e, lay = _run_and_compare(cfg)
    _say(f"usb x lsb {xe} cross max", max(e["cross"]), "auto max", max(e["auto"]))
    assert max(e["cross"]) < bound and max(e["auto"]) < bound


@pytest.mark.parametrize("cs,lsb", [(1, True), (2, False), (2, True)])
@pytest.mark.parametrize("xe,bound", [("fp32", 1e-5), ("fp16", 1e-3)])
def test_complex_sidebands(gpu, cs, lsb, xe, bound):
    """Complex LSB (DC kept, other channels conj-flipped) and complex double sideband (halves
    swapped; LSB conj-flipped about the band centre), mode.cpp:1186-1222."""
    cfg = case_cfg2(B=8, xengine=xe, nstations=8, complex_sampling=cs, lsb=lsb)
    got, ref, e, lay = _run_and_compare(cfg)
    _say(f"complex {cs} lsb {lsb} {xe} cross max", max(e["cross"]), "auto max", max(e["auto"]))
    assert max(e["cross"]) < bound and max(e["auto"]) < bound


@pytest.mark.parametrize("xe", ["fp32", "fp16"])
def test_fengine_launch_forms_bit_identical(gpu, xe):
    """One launch per FFT length with 256- or 512-thread workgroups, and every length in one
    persistent launch, compute the same arithmetic per FFT: identical results."""
    offsets = (0, 1, 0)
    out = {}
    for mode, threads in ((1, 256), (1, 512), (2, 256), (2, 512)):
        cfg = case_cfg3(B=32, xengine=xe, fengine_mode=mode, fengine_threads=threads)
        core = _core(cfg)
        data, ctrl = slot(cfg, offsets, seed=31)
        out[(mode, threads)] = core.processdata(offsets, data, ctrl)
        core.close()
    ref = out[(1, 256)]
    for key, v in out.items():
        _say("launch form", key, "max |diff|", float(np.max(np.abs(v - ref))))
        np.testing.assert_array_equal(v.view(np.uint64), ref.view(np.uint64))
    lay = fo.result_layout(case_cfg3(B=32, xengine=xe))
    data, ctrl = slot(case_cfg3(B=32, xengine=xe), offsets, seed=31)
    oref, _ = fo.process_subint(case_cfg3(B=32, xengine=xe), list(offsets), data, ctrl)
    e = record_errors(case_cfg3(B=32, xengine=xe), lay, ref.astype(np.complex128), oref)
    assert max(e["cross"]) < (1e-5 if xe == "fp32" else 1e-3)


def test_two_positions_per_thread_form(gpu):
    """L = 4096 at 256 threads (two butterfly positions per thread) against 512 threads (one):
    a different instruction stream, so not bit-identical, but equal to FP32 rounding and both
    within the FP32 bar against the oracle."""
    from difx_amd import config as cf
    offsets = (0, 1, 0)
    mk = lambda t: cf.standard_config(4, 2, [4800.0, 4840.0], 32.0, 2048, subint_ns=8 * 64000, seed=33,
                                      fengine_threads=t, max_delay_us=200.0)
    data, ctrl = slot(mk(0), offsets, seed=34)
    out = {}
    for t in (256, 512):
        core = _core(mk(t))
        out[t] = core.processdata(offsets, data, ctrl).astype(np.complex128)
        core.close()
    ref, _ = fo.process_subint(mk(0), list(offsets), data, ctrl)
    lay = fo.result_layout(mk(0))
    rel = float(np.mean(np.abs(out[256] - out[512])) / np.mean(np.abs(out[512])))
    _say("L 4096, 256 vs 512 threads: mean |diff| / mean |V|", rel)
    assert rel < 1e-6
    for t in (256, 512):
        e = record_errors(mk(t), lay, out[t], ref)
        assert max(e["cross"]) < 1e-5 and max(e["auto"]) < 1e-5


def test_autotune_picks_a_form(gpu):
    cfg = case_cfg3(B=64, xengine="fp16", autotune=1)
    core = _core(cfg)
    rep = core.autotune_report()
    _say(rep)
    assert "F-engine:" in rep and "ms per subint" in rep
    data, ctrl = slot(cfg, (0, 1, 0), seed=32)
    got = core.processdata((0, 1, 0), data, ctrl)
    core.close()
    ref, _ = fo.process_subint(cfg, [0, 1, 0], data, ctrl)
    e = record_errors(cfg, fo.result_layout(cfg), got.astype(np.complex128), ref)
    assert max(e["cross"]) < 1e-3


@pytest.mark.parametrize("xe,bound", [("fp32", 1e-5), ("fp16", 1e-3)])
@pytest.mark.parametrize("kind", ["real", "complex", "real_damaged"])
def test_fourstep_ffts(gpu, kind, xe, bound):
    """FFT lengths 16384 and 32768 (four-step F-engine: 2048-point LDS sub-FFTs, twiddles,
    R-point DFTs) with zoom windows; a 16384-point complex FFT; flagged FFTs, an invalid frame
    and a fill-pattern frame (FFTs with invalid samples take the per-sample decode, the others
    the gathered one)."""
    cfg = case_complex_fourstep(B=4, xengine=xe) if kind == "complex" else case_fourstep_ffts(B=4, xengine=xe)
    got, ref, e, lay = _run_and_compare(cfg, damaged=(kind == "real_damaged"))
    _say(f"four-step {kind} {xe} cross max", max(e["cross"]), "auto max", max(e["auto"]))
    assert max(e["cross"]) < bound and max(e["auto"]) < bound


def test_fourstep_in_persistent_launch_bit_identical(gpu):
    """The persistent launch runs the 16384/32768-point items with the same four-step body."""
    offsets = (0, 1, 0)
    out = {}
    for mode, threads in ((1, 256), (2, 256), (2, 512)):
        cfg = case_fourstep_ffts(B=8, xengine="fp16", fengine_mode=mode, fengine_threads=threads)
        core = _core(cfg)
        data, ctrl = slot(cfg, offsets, seed=33)
        out[(mode, threads)] = core.processdata(offsets, data, ctrl)
        core.close()
    ref = out[(1, 256)]
    for key, v in out.items():
        np.testing.assert_array_equal(v.view(np.uint64), ref.view(np.uint64))


@pytest.mark.parametrize("name", ["offsets", "offsets_fourstep", "dsb_fraclo", "mark5b", "mark5b_zones", "real8",
                                  "damaged"])
@pytest.mark.parametrize("xe,bound", [("fp32", 1e-5), ("fp16", 1e-3)])
def test_more_cases(gpu, name, xe, bound):
    """Per-freq LO offsets (integer and fractional Hz), a fractional-MHz band edge, clock
    offsets, per-pol clock deltas and pol phase offsets with channel averaging 4 (also with L
    recorded before R, and on the four-step F-engine); double-sideband data whose band centre
    is a fractional MHz; Mark5B frames, also with fill-pattern zones inside frames; 8-bit real
    samples; a slot with flagged FFTs, an invalid frame and a fill-pattern frame."""
    if name == "offsets":
        cfg = case_offsets(B=32, xengine=xe)
    elif name == "offsets_fourstep":
        cfg = case_offsets_fourstep(B=4, xengine=xe)
    elif name == "dsb_fraclo":
        cfg = case_dsb_fraclo(B=16, xengine=xe)
    elif name in ("mark5b", "mark5b_zones"):
        cfg = case_mark5b(B=32, xengine=xe)
    elif name == "real8":
        cfg = case_real8(B=32, xengine=xe)
    else:
        cfg = case_cfg1(B=32, nstations=3, nfreq=2, xengine=xe)
    got, ref, e, lay = _run_and_compare(cfg, damaged=(name == "damaged"),
                                        spectra_check=not name.startswith("offsets"),
                                        make_slot=mark5b_partial_slot if name == "mark5b_zones" else None)
    _say(f"{name} {xe} cross max", max(e["cross"]), "auto max", max(e["auto"]))
    assert max(e["cross"]) < bound and max(e["auto"]) < bound
