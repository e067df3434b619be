"""Python front end of the serving calls of include/fxg.h (fxg_server_*).

`FxServer` runs many correlation jobs on one device: each job is a Config plus a number of
subints; its subints are pushed as they arrive (device or host data), the library's
scheduler admits jobs by device footprint, overlaps the active jobs on their own streams and
integrates each job's subint results on the device (Visibility::addData,
mpifxcorr/src/visibility.cpp:192-205).  `wait_job` returns a job's integrated (unnormalised)
results in the populateResultLengths layout and its p50 / p99 subint latency.
"""
from __future__ import annotations

import ctypes as C
from typing import Dict, Sequence, Tuple

import numpy as np

from . import fxg_abi as abi
from .config import Config
from .core import FxgError, _ptr_array, footprint, load_library


class FxServer:
    def __init__(self, hbm_budget_bytes: int = 0, max_jobs: int = 0, inflight_per_job: int = 0):
        self.lib = load_library()
        sc = abi.fxg_server_config(hbm_budget_bytes, max_jobs, inflight_per_job)
        h = C.c_void_p()
        rc = self.lib.fxg_server_create(C.byref(sc), C.byref(h))
        if rc != abi.FXG_OK:
            raise FxgError(f"fxg_server_create: {self.lib.fxg_strerror(rc).decode()}")
        self.h = h
        self._crl: Dict[int, int] = {}
        self._keep: Dict[int, list] = {}

    def close(self):
        if self.h:
            self.lib.fxg_server_destroy(self.h)
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

    def add_job(self, cfg: Config, nsubints: int, data_on_device: bool = True) -> int:
        fc, keep = cfg.to_abi()
        _, crl, _ = footprint(cfg)
        jid = C.c_int32()
        rc = self.lib.fxg_server_add_job(self.h, C.byref(fc), nsubints, int(data_on_device), C.byref(jid))
        if rc != abi.FXG_OK:
            raise FxgError(f"fxg_server_add_job: {self.lib.fxg_strerror(rc).decode()}")
        self._crl[jid.value] = crl
        self._keep[jid.value] = []
        return jid.value

    def push(self, job: int, offsets: Sequence[int], data_ptrs: Sequence[int], nbytes: Sequence[int],
             control: Sequence[np.ndarray]):
        """Push one subint; data_ptrs are device (or host) addresses that stay valid until
        wait_job(job) returns."""
        offs = (C.c_int32 * 3)(*offsets)
        nb = (C.c_int32 * len(nbytes))(*nbytes)
        ctl = [np.ascontiguousarray(c, dtype=np.int32) for c in control]
        rc = self.lib.fxg_server_push_subint(self.h, job, offs, _ptr_array(list(data_ptrs)), nb,
                                             _ptr_array([c.ctypes.data for c in ctl]))
        if rc != abi.FXG_OK:
            raise FxgError(f"fxg_server_push_subint: {self.lib.fxg_strerror(rc).decode()}")

    def wait_job(self, job: int) -> Tuple[np.ndarray, float, float]:
        out = np.zeros(self._crl[job], dtype=np.complex64)
        p50, p99 = C.c_double(), C.c_double()
        rc = self.lib.fxg_server_wait_job(self.h, job, out.ctypes.data, C.byref(p50), C.byref(p99))
        if rc != abi.FXG_OK:
            buf = C.create_string_buffer(512)
            self.lib.fxg_server_job_error(self.h, job, buf, 512)
            raise FxgError(f"job {job} failed: {buf.value.decode()}")
        self._keep.pop(job, None)
        return out, p50.value, p99.value

    def stats(self) -> dict:
        st = abi.fxg_server_stats()
        self.lib.fxg_server_get_stats(self.h, C.byref(st))
        return {name: getattr(st, name) for name, _ in st._fields_}
