"""VDIF thread corner turn on the device: the Python mirror of VDIFMuxer
(mpifxcorr/src/datamuxer.h:148-240) over fxg_vdif_mux_device of include/fxg.h.

A station that records one VDIF thread per channel delivers single-channel frames of N
threads interleaved in arbitrary order; `multiplex` turns one segment of them into N-channel
frames (the stream a datastream of N recorded bands decodes), on the device, with the header
rules of VDIFMuxer::multiplex (datamuxer.cpp:670-757) and the slot assignment of
VDIFMuxer::deinterlace (datamuxer.cpp:764-833).  Device buffers are torch uint8 tensors;
there is no host fallback: without libfxg.so the constructor raises.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence, Tuple

import numpy as np

from . import fxg_abi as abi
from .core import FxgError, load_library

HEADER_BYTES = 32


def _epoch_mjd(epoch: int) -> int:
    """getVDIFEpochMJD (vdifio.c:127-131): MJD of 1 January / 1 July of 2000 + epoch / 2."""
    yr = 2000 + epoch // 2
    mo = (epoch % 2) * 6 + 1
    doy = (0 if mo == 1 else 181 + (1 if (yr % 4 == 0 and (yr % 100 != 0 or yr % 400 == 0)) else 0)) + 1
    y1 = yr - 1
    return doy - 678576 + 365 * y1 + y1 // 4 - y1 // 100 + y1 // 400


class VDIFMuxer:
    """One station's corner turner (VDIFMuxer).  `threadmap[j]` is the VDIF thread id of output
    channel j.  `initialise` takes the reference frame from a header (the reference uses the
    first frame of its buffer, datamuxer.cpp:136-150); each `multiplex` call then turns the
    next `nslots` frame times (processframenumber advances by nslots, as per slot in
    multiplex)."""

    def __init__(self, nthreads: int, inputframebytes: int, framespersecond: int, bitspersample: int,
                 threadmap: Sequence[int]):
        if len(threadmap) != nthreads:
            raise ValueError("threadmap must name every thread")
        if nthreads < 1 or nthreads > abi.FXG_MUX_MAX_THREADS or nthreads & (nthreads - 1):
            raise ValueError("nthreads must be a power of two <= %d" % abi.FXG_MUX_MAX_THREADS)
        if (32 // bitspersample) // nthreads == 0:
            # datamuxer.cpp:155-158
            raise ValueError("too many threads/too high bit resolution: no complete timestep in a 32-bit word")
        self.lib = load_library()
        self.p = abi.fxg_vdif_mux_params()
        self.p.inputframebytes = inputframebytes
        self.p.nthreads = nthreads
        self.p.bitspersample = bitspersample
        self.p.framespersecond = framespersecond
        for j, t in enumerate(threadmap):
            self.p.threadmap[j] = int(t)
        self.nthreads = nthreads
        self.inputframebytes = inputframebytes
        self.outputframebytes = (inputframebytes - HEADER_BYTES) * nthreads + HEADER_BYTES
        self.processframenumber = 0
        self.dropped = {"early": 0, "late": 0, "unmapped": 0}
        self._work = None
        self._initialised = False

    def initialise(self, header) -> bool:
        """Reference time from a frame header (at least 32 bytes: numpy, bytes or a torch
        tensor on any device)."""
        try:
            import torch
            if isinstance(header, torch.Tensor):
                header = header[:HEADER_BYTES].cpu().numpy()
        except ImportError:
            pass
        w = np.frombuffer(bytes(np.asarray(header, dtype=np.uint8)[:HEADER_BYTES]), dtype="<u4")
        seconds = int(w[0]) & 0x3FFFFFFF
        epoch = (int(w[1]) >> 24) & 0x3F
        self.p.ref_seconds = _epoch_mjd(epoch) * 86400 + seconds
        self.p.ref_frame = int(w[1]) & 0xFFFFFF
        self.processframenumber = 0
        self._initialised = True
        return True

    def set_reference(self, ref_seconds: int, ref_frame: int) -> None:
        """Reference frame as MJD * 86400 + seconds and its frame number."""
        self.p.ref_seconds = int(ref_seconds)
        self.p.ref_frame = int(ref_frame)
        self.processframenumber = 0
        self._initialised = True

    def enqueue(self, inbuf, nslots: int, out=None, start: Optional[int] = None, stream=None):
        """Enqueue the corner turn of frame times start .. start + nslots - 1 (default: the next
        ones) from the input frames in `inbuf` (a torch uint8 tensor on the device) on `stream`
        (default: torch's current stream) and return the output tensor (nslots output
        frames); `finish` gives the outcome.  One corner turn of a muxer is in flight at a
        time (the workspace holds its slot table)."""
        import torch
        if not self._initialised:
            raise FxgError("initialise() first")
        if inbuf.dtype != torch.uint8 or not inbuf.is_cuda:
            raise ValueError("inbuf must be a uint8 tensor on the device")
        dev = inbuf.device
        if out is None:
            out = torch.empty(nslots * self.outputframebytes, dtype=torch.uint8, device=dev)
        wb = int(self.lib.fxg_vdif_mux_workspace_bytes(nslots, self.nthreads))
        if self._work is None or self._work.numel() < wb or self._work.device != dev:
            self._work = torch.empty(wb, dtype=torch.uint8, device=dev)
        self.p.start = self.processframenumber if start is None else int(start)
        self._stream = C.c_void_p(stream.cuda_stream if stream is not None
                                  else torch.cuda.current_stream(dev).cuda_stream)
        rc = self.lib.fxg_vdif_mux_device(C.byref(self.p), C.c_void_p(inbuf.data_ptr() if inbuf.numel() else 0),
                                          inbuf.numel(), C.c_void_p(out.data_ptr()), out.numel(), nslots,
                                          C.c_void_p(self._work.data_ptr()), self._work.numel(), self._stream)
        if rc != abi.FXG_OK:
            raise FxgError("fxg_vdif_mux_device: %s" % self.lib.fxg_strerror(rc).decode())
        if start is None:
            self.processframenumber += nslots
        return out

    def finish(self) -> int:
        """Wait for the enqueued corner turn; returns the output bytes (what
        VDIFMuxer::multiplex returns) and adds the dropped frames to `self.dropped`."""
        nout = C.c_int64()
        dr = (C.c_int64 * 3)()
        rc = self.lib.fxg_vdif_mux_result(C.c_void_p(self._work.data_ptr()), self._stream, C.byref(nout), dr)
        if rc != abi.FXG_OK:
            raise FxgError("frame length changed in the input (the reference aborts)")
        for k, v in zip(("early", "late", "unmapped"), dr):
            self.dropped[k] += int(v)
        return int(nout.value) * self.outputframebytes

    def multiplex(self, inbuf, nslots: int, out=None, start: Optional[int] = None,
                  stream=None) -> Tuple[object, int]:
        """enqueue + finish: the output tensor and the output bytes."""
        out = self.enqueue(inbuf, nslots, out=out, start=start, stream=stream)
        return out, self.finish()
