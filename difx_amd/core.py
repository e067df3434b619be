"""Python mirror of the reference's Core/Mode call surface over libfxg.so.

`FxCore` stands where mpifxcorr's Core stands for one configuration
(mpifxcorr/src/core.h:52-169): `processdata` takes one slot -- the subint offsets
{scan, sec, ns}, each datastream's data buffer and its control buffer
{scan, sec, ns, validflags...} -- and returns the slot's results array in the layout of
Configuration::populateResultLengths (configuration.cpp:2348-2630).  The Mode getters of
mode.h:138-203 (getFreqs, getDataWeight) and Model::calculateDelayInterpolator
(model.cpp:219-289) are exposed for inspection.  Every call goes through the C ABI of
include/fxg.h; there is no host fallback: if libfxg.so is missing or no device is present
the constructor raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path
from typing import List, Optional, Sequence

import numpy as np

from . import fxg_abi as abi
from .config import Config

_LIB = None
LIB_PATH = Path(__file__).resolve().parent / "libfxg.so"


class FxgError(RuntimeError):
    pass


def load_library() -> C.CDLL:
    """Load the in-tree libfxg.so and declare its entry points.  Raises if absent."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not LIB_PATH.exists():
        raise FxgError(f"{LIB_PATH} is not built: run `python -m difx_amd.build` first")
    lib = C.CDLL(str(LIB_PATH))
    P = C.POINTER
    lib.fxg_create.argtypes = [P(abi.fxg_config), P(C.c_void_p)]
    lib.fxg_create_layout.argtypes = [P(abi.fxg_config), P(C.c_void_p)]
    lib.fxg_destroy.argtypes = [C.c_void_p]
    lib.fxg_strerror.argtypes = [C.c_int]
    lib.fxg_strerror.restype = C.c_char_p
    lib.fxg_last_error.argtypes = [C.c_void_p]
    lib.fxg_last_error.restype = C.c_char_p
    lib.fxg_get_layout.argtypes = [C.c_void_p, P(abi.fxg_layout)]
    lib.fxg_data_bytes.argtypes = [C.c_void_p, C.c_int32, P(C.c_int64)]
    lib.fxg_result_offsets.argtypes = [C.c_void_p, P(C.c_int32), P(C.c_int32), P(C.c_int32), P(C.c_int32)]
    lib.fxg_result_pcal_offsets.argtypes = [C.c_void_p, P(C.c_int32)]
    lib.fxg_process_subint.argtypes = [C.c_void_p, P(C.c_int32), P(C.c_void_p), P(C.c_int32),
                                       P(C.c_void_p), C.c_void_p]
    lib.fxg_process_subint_device.argtypes = [C.c_void_p, P(C.c_int32), P(C.c_void_p), P(C.c_int32),
                                              P(C.c_void_p), C.c_void_p, C.c_void_p]
    lib.fxg_ring_length.argtypes = [C.c_void_p]
    lib.fxg_submit_subint.argtypes = [C.c_void_p, P(C.c_int32), P(C.c_void_p), P(C.c_int32),
                                      P(C.c_void_p), P(C.c_int32)]
    lib.fxg_wait.argtypes = [C.c_void_p, C.c_int32, C.c_void_p]
    lib.fxg_get_spectrum.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_void_p]
    lib.fxg_get_data_weights.argtypes = [C.c_void_p, C.c_int32, C.c_void_p]
    lib.fxg_unpack_device.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_int32, C.c_int64,
                                      C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.fxg_delay_interpolator.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, P(C.c_double)]
    lib.fxg_set_timing.argtypes = [C.c_void_p, C.c_int32]
    lib.fxg_get_timing.argtypes = [C.c_void_p, P(abi.fxg_timing)]
    lib.fxg_autotune_report.argtypes = [C.c_void_p, C.c_char_p, C.c_int32]
    lib.fxg_footprint.argtypes = [P(abi.fxg_config), P(C.c_int64), P(C.c_int64), P(C.c_int64)]
    lib.fxg_input_samples.argtypes = [C.c_void_p, P(C.c_int64)]
    lib.fxg_device_bytes.argtypes = [C.c_void_p, P(C.c_int64)]
    lib.fxg_server_create.argtypes = [P(abi.fxg_server_config), P(C.c_void_p)]
    lib.fxg_server_destroy.argtypes = [C.c_void_p]
    lib.fxg_server_add_job.argtypes = [C.c_void_p, P(abi.fxg_config), C.c_int32, C.c_int32, P(C.c_int32)]
    lib.fxg_server_push_subint.argtypes = [C.c_void_p, C.c_int32, P(C.c_int32), P(C.c_void_p), P(C.c_int32),
                                           P(C.c_void_p)]
    lib.fxg_server_wait_job.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, P(C.c_double), P(C.c_double)]
    lib.fxg_server_get_stats.argtypes = [C.c_void_p, P(abi.fxg_server_stats)]
    lib.fxg_server_job_error.argtypes = [C.c_void_p, C.c_int32, C.c_char_p, C.c_int32]
    lib.fxg_vdif_mux_workspace_bytes.argtypes = [C.c_int64, C.c_int32]
    lib.fxg_vdif_mux_workspace_bytes.restype = C.c_int64
    lib.fxg_vdif_mux_device.argtypes = [P(abi.fxg_vdif_mux_params), C.c_void_p, C.c_int64, C.c_void_p, C.c_int64,
                                        C.c_int64, C.c_void_p, C.c_int64, C.c_void_p]
    lib.fxg_vdif_mux_result.argtypes = [C.c_void_p, C.c_void_p, P(C.c_int64), P(C.c_int64)]
    lib.fxg_vdif_mux_device_static.argtypes = [P(abi.fxg_vdif_mux_params), C.c_void_p, C.c_int64, C.c_void_p,
                                               C.c_int64, C.c_int64, C.c_void_p, C.c_void_p]
    lib.fxg_segment_params.argtypes = [P(abi.fxg_config), C.c_int32, C.c_int32, C.c_int32, P(abi.fxg_segment)]
    lib.fxg_buffer_params.argtypes = [P(abi.fxg_config), C.c_int32, C.c_int32, C.c_int32, P(abi.fxg_databuffer)]
    lib.fxg_set_kurtosis.argtypes = [C.c_void_p, C.c_int32]
    lib.fxg_set_pcal.argtypes = [C.c_void_p, C.c_int32, C.c_int64, C.c_int64]
    lib.fxg_get_pcal.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, P(C.c_int32)]
    lib.fxg_pcal_info.argtypes = [C.c_void_p, C.c_int32, C.c_int32, P(C.c_int32), P(C.c_int64), P(C.c_int32)]
    lib.fxg_sta_autocorr.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_double,
                                     C.c_void_p, P(C.c_int32)]
    lib.fxg_get_kurtosis.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_void_p, P(C.c_int32)]
    lib.fxg_normalise_device.argtypes = [C.c_void_p, C.c_void_p, C.c_double, C.c_void_p, C.c_void_p]
    lib.fxg_difx_records.argtypes = [C.c_void_p, C.c_void_p, P(abi.fxg_difx_meta), C.c_void_p, C.c_int64,
                                     P(C.c_int64), P(C.c_int32)]
    lib.fxg_control_params.argtypes = [P(abi.fxg_databuffer), P(C.c_int32), C.c_int32, C.c_int32, C.c_int32,
                                       C.c_int32, C.c_double, C.c_double, abi.FXG_ADVANCE_FN, C.c_void_p,
                                       P(C.c_int32), P(C.c_int32)]
    _LIB = lib
    return lib


def footprint(cfg: Config):
    """(device bytes of a context, coreresultlength, one slot's input bytes) for cfg, computed
    without a device (fxg_footprint)."""
    lib = load_library()
    fc, keep = cfg.to_abi()
    dev, crl, slot = C.c_int64(), C.c_int64(), C.c_int64()
    rc = lib.fxg_footprint(C.byref(fc), C.byref(dev), C.byref(crl), C.byref(slot))
    if rc != abi.FXG_OK:
        raise FxgError(f"fxg_footprint: {lib.fxg_strerror(rc).decode()}: {lib.fxg_last_error(None).decode()}")
    return dev.value, crl.value, slot.value


def _ptr_array(ptrs: Sequence[int]):
    arr = (C.c_void_p * len(ptrs))()
    for i, p in enumerate(ptrs):
        arr[i] = p
    return arr


class FxCore:
    """One configuration's device-resident FX pipeline (Core + its Modes)."""

    def __init__(self, cfg: Config):
        self.lib = load_library()
        self.cfg = cfg
        self._abi, self._keep = cfg.to_abi()
        h = C.c_void_p()
        rc = self.lib.fxg_create(C.byref(self._abi), C.byref(h))
        if rc != abi.FXG_OK:
            raise FxgError(f"fxg_create failed: {self.lib.fxg_strerror(rc).decode()}: "
                           f"{self.lib.fxg_last_error(None).decode()}")
        self.h = h
        lay = abi.fxg_layout()
        self._check(self.lib.fxg_get_layout(self.h, C.byref(lay)))
        self.coreresultlength = int(lay.coreresultlength)
        self.blockspersend = int(lay.blockspersend)
        self.ndatastreams = len(cfg.datastreams)

    def _check(self, rc: int):
        if rc != abi.FXG_OK:
            msg = self.lib.fxg_last_error(self.h).decode() if getattr(self, "h", None) else ""
            raise FxgError(f"{self.lib.fxg_strerror(rc).decode()}: {msg}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.fxg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- layout ----
    def data_bytes(self, ds: int) -> int:
        v = C.c_int64()
        self._check(self.lib.fxg_data_bytes(self.h, ds, C.byref(v)))
        return int(v.value)

    def result_offsets(self):
        nf, nb, nd = len(self.cfg.freqs), len(self.cfg.baselines), self.ndatastreams
        b = np.zeros(max(nf * nb, 1), np.int32)
        w = np.zeros(max(nf * nb, 1), np.int32)
        a = np.zeros(nd, np.int32)
        aw = np.zeros(nd, np.int32)
        P = C.POINTER(C.c_int32)
        self._check(self.lib.fxg_result_offsets(self.h, b.ctypes.data_as(P), w.ctypes.data_as(P),
                                                a.ctypes.data_as(P), aw.ctypes.data_as(P)))
        return b[: nf * nb].reshape(nf, nb), w[: nf * nb].reshape(nf, nb), a, aw

    def result_pcal_offsets(self) -> np.ndarray:
        """Configuration::getCoreResultPCalOffset of every datastream (cf32 units)."""
        p = np.zeros(self.ndatastreams, np.int32)
        self._check(self.lib.fxg_result_pcal_offsets(self.h, p.ctypes.data_as(C.POINTER(C.c_int32))))
        return p

    # ---- Core::processdata ----
    def processdata(self, offsets: Sequence[int], data: Sequence[np.ndarray],
                    control: Sequence[np.ndarray]) -> np.ndarray:
        """Host buffers in, results (complex64[coreresultlength]) out."""
        offs = (C.c_int32 * 3)(*[int(x) for x in offsets])
        data = [np.ascontiguousarray(d, dtype=np.uint8) for d in data]
        ctrl = [np.ascontiguousarray(c, dtype=np.int32) for c in control]
        nbytes = (C.c_int32 * len(data))(*[len(d) for d in data])
        out = np.zeros(self.coreresultlength, np.complex64)
        self._check(self.lib.fxg_process_subint(
            self.h, offs, _ptr_array([d.ctypes.data for d in data]), nbytes,
            _ptr_array([c.ctypes.data for c in ctrl]), out.ctypes.data))
        return out

    def processdata_device(self, offsets, d_data: Sequence[int], nbytes: Sequence[int],
                           control: Sequence[np.ndarray], d_results: int, stream: int = 0):
        """Device-resident variant: d_data are device pointers (e.g. torch tensor
        data_ptr()), d_results a device pointer of coreresultlength cf32; enqueued on
        `stream` (a hipStream_t handle, 0 = the context's own stream)."""
        offs = (C.c_int32 * 3)(*[int(x) for x in offsets])
        ctrl = [np.ascontiguousarray(c, dtype=np.int32) for c in control]
        self._ctrl_keep = ctrl
        nb = (C.c_int32 * len(nbytes))(*[int(x) for x in nbytes])
        self._check(self.lib.fxg_process_subint_device(
            self.h, offs, _ptr_array(list(d_data)), nb, _ptr_array([c.ctypes.data for c in ctrl]),
            C.c_void_p(d_results), C.c_void_p(stream) if stream else None))

    def set_kurtosis(self, enable: bool):
        """Mode::setDumpKurtosis: sum |X|^2 and |X|^4 per channel in every following subint."""
        self._check(self.lib.fxg_set_kurtosis(self.h, int(bool(enable))))

    def getKurtosis(self, ds: int, band: int, maxchannels: int) -> np.ndarray:
        """Mode::calculateAndAverageKurtosis + getKurtosis for the last subint (empty when the
        band had no data)."""
        d = self.cfg.datastreams[ds]
        nchan = self.cfg.freqs[d.recfreqs[d.recbands[band][0]]].nchan
        out = np.zeros(max(min(maxchannels, nchan), 1), dtype=np.float32)
        n = C.c_int32()
        self._check(self.lib.fxg_get_kurtosis(self.h, ds, band, maxchannels, out.ctypes.data, C.byref(n)))
        return out[:n.value].copy()

    def set_pcal(self, ds: int, interval_hz: int, base_hz: int = 0):
        """Phase-cal extraction of datastream ds at a whole-Hz tone interval (0 turns it off)."""
        self._check(self.lib.fxg_set_pcal(self.h, ds, int(interval_hz), int(base_hz)))

    def getPcal(self, ds: int, band: int) -> np.ndarray:
        """Mode::finalisepcal's tones of a recorded band for the last subint (complex64)."""
        d = self.cfg.datastreams[ds]
        bw = self.cfg.freqs[d.recfreqs[d.recbands[band][0]]].bw_mhz
        out = np.zeros(2 * (int(bw * 1e6) + 2), dtype=np.float32)   # at most one tone per Hz
        n = C.c_int32()
        self._check(self.lib.fxg_get_pcal(self.h, ds, band, out.ctypes.data, C.byref(n)))
        return out[: 2 * n.value].view(np.complex64).copy()

    def normalise_device(self, d_results: int, meansubints: float, tsys=None, stream: int = 0):
        """Visibility::writedata's amplitude scaling of a device results array, in place."""
        t = None if tsys is None else (C.c_double * self.ndatastreams)(*[float(x) for x in tsys])
        self._check(self.lib.fxg_normalise_device(self.h, C.c_void_p(d_results), float(meansubints), t,
                                                  C.c_void_p(stream) if stream else None))

    def difx_records(self, results: np.ndarray, mjd: int, seconds: float, configindex: int, sourceindex: int,
                     fftsperintegration: float, baselinenumber, uvw, telescopeindex, phasecentre: int = 0) -> bytes:
        """Visibility::writeSWIN's records of a normalised host results array (complex64); with
        several phase centres those of one phase centre's file."""
        res = np.ascontiguousarray(results, dtype=np.complex64)
        bl = np.ascontiguousarray(baselinenumber, dtype=np.int32)
        uv = np.ascontiguousarray(uvw, dtype=np.float64)
        tel = np.ascontiguousarray(telescopeindex, dtype=np.int32)
        m = abi.fxg_difx_meta(int(mjd), int(configindex), int(sourceindex), int(phasecentre), float(seconds),
                              float(fftsperintegration), bl.ctypes.data_as(C.POINTER(C.c_int32)),
                              uv.ctypes.data_as(C.POINTER(C.c_double)), tel.ctypes.data_as(C.POINTER(C.c_int32)))
        n, nrec = C.c_int64(), C.c_int32()
        self._check(self.lib.fxg_difx_records(self.h, res.ctypes.data, C.byref(m), None, 0, C.byref(n),
                                              C.byref(nrec)))
        out = np.zeros(n.value, dtype=np.uint8)
        self._check(self.lib.fxg_difx_records(self.h, res.ctypes.data, C.byref(m), out.ctypes.data, out.size,
                                              C.byref(n), C.byref(nrec)))
        return out.tobytes()

    def submit(self, offsets, data, control) -> int:
        offs = (C.c_int32 * 3)(*[int(x) for x in offsets])
        data = [np.ascontiguousarray(d, dtype=np.uint8) for d in data]
        ctrl = [np.ascontiguousarray(c, dtype=np.int32) for c in control]
        nbytes = (C.c_int32 * len(data))(*[len(d) for d in data])
        hnd = C.c_int32()
        self._check(self.lib.fxg_submit_subint(
            self.h, offs, _ptr_array([d.ctypes.data for d in data]), nbytes,
            _ptr_array([c.ctypes.data for c in ctrl]), C.byref(hnd)))
        return int(hnd.value)

    def wait(self, handle: int) -> np.ndarray:
        out = np.zeros(self.coreresultlength, np.complex64)
        self._check(self.lib.fxg_wait(self.h, handle, out.ctypes.data))
        return out

    # ---- Mode getters ----
    def getFreqs(self, ds: int, band: int, fft: int) -> np.ndarray:
        C_ = self.cfg.freqs[self.cfg.datastreams[ds].recfreqs[0]].nchan
        out = np.zeros(C_, np.complex64)
        self._check(self.lib.fxg_get_spectrum(self.h, ds, band, fft, out.ctypes.data))
        return out

    def getDataWeights(self, ds: int) -> np.ndarray:
        out = np.zeros(self.blockspersend, np.float32)
        self._check(self.lib.fxg_get_data_weights(self.h, ds, out.ctypes.data))
        return out

    def calculateDelayInterpolator(self, ds: int, sec: int, ns: int) -> np.ndarray:
        out = (C.c_double * 3)()
        self._check(self.lib.fxg_delay_interpolator(self.h, ds, sec, ns, out))
        return np.array(out[:])

    def autotune_report(self) -> str:
        """The F-engine launch form in use and the autotuner's candidate timings."""
        buf = C.create_string_buffer(4096)
        self._check(self.lib.fxg_autotune_report(self.h, buf, 4096))
        return buf.value.decode()

    def set_timing(self, enable: bool):
        self._check(self.lib.fxg_set_timing(self.h, int(enable)))

    def timing(self) -> dict:
        t = abi.fxg_timing()
        self._check(self.lib.fxg_get_timing(self.h, C.byref(t)))
        return {"f_ms": t.f_ms, "f_launches": t.f_launches, "x_ms": t.x_ms,
                "x_launches": t.x_launches, "other_ms": t.other_ms,
                "other_launches": t.other_launches}

    def unpack_device(self, ds: int, d_data: int, nbytes: int, startsample: int, nsamples: int,
                      d_out: int, d_good: int, stream: int = 0):
        self._check(self.lib.fxg_unpack_device(self.h, ds, C.c_void_p(d_data), nbytes, startsample,
                                               nsamples, C.c_void_p(d_out), C.c_void_p(d_good),
                                               C.c_void_p(stream) if stream else None))
