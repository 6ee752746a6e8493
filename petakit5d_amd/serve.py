"""Request loop serving decon jobs on one device (continuous batching under an
HBM budget) -- Python front end of pk_serve_* in include/pkdecon.h."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import PkDeconError, PkOpts, PkVol
from .batch import _vol_of
from .decon import _opts_struct, _parse, _psf_array

QUEUED, RUNNING, DONE, FAILED, CANCELLED = range(5)


class PkRequest(ctypes.Structure):
    _fields_ = [
        ("method", ctypes.c_int32),
        ("numit", ctypes.c_int32),
        ("inp", PkVol),
        ("out", PkVol),
        ("psf_f", ctypes.c_void_p),
        ("f_dims", ctypes.c_int64 * 3),
        ("psf_b", ctypes.c_void_p),
        ("b_dims", ctypes.c_int64 * 3),
        ("opts", PkOpts),
    ]


class PkReqInfo(ctypes.Structure):
    _fields_ = [
        ("state", ctypes.c_int32),
        ("status", ctypes.c_int32),
        ("iterations_done", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("latency_ms", ctypes.c_double),
        ("queue_ms", ctypes.c_double),
    ]


class PkServeStats(ctypes.Structure):
    _fields_ = [
        ("submitted", ctypes.c_uint64),
        ("completed", ctypes.c_uint64),
        ("failed", ctypes.c_uint64),
        ("active", ctypes.c_int32),
        ("queued", ctypes.c_int32),
        ("voxels", ctypes.c_double),
    ]


def _pinned_host_vol(x):
    """(keepalive, PkVol, 'numpy') for a pinned, column-major torch CPU tensor
    (the server reads it in place: no staging copy), else None."""
    try:
        import torch
    except ImportError:
        return None
    if not isinstance(x, torch.Tensor) or x.is_cuda or x.dim() != 3 or not x.is_pinned():
        return None
    if x.dtype not in (torch.uint16, torch.float32):
        return None
    d = tuple(int(n) for n in x.shape)
    if tuple(x.stride()) != (1, d[0], d[0] * d[1]):
        return None
    code = 1 if x.dtype == torch.uint16 else 2
    return x.numpy(), PkVol(ctypes.c_void_p(x.data_ptr()), code, _lib.PK_MEM_HOST, (ctypes.c_int64 * 3)(*d)), "numpy"


def _pinned(shape_rev, dtype):
    """A pinned host tensor of the reversed (C-order) shape, or None when no
    GPU runtime is there to pin memory."""
    try:
        import torch
        if not torch.cuda.is_available():
            return None
        tdt = {np.dtype(np.uint16): torch.uint16, np.dtype(np.float32): torch.float32}[np.dtype(dtype)]
        return torch.empty(shape_rev, dtype=tdt, pin_memory=True)
    except Exception:  # noqa: BLE001 -- pinning is an optimisation only
        return None


def _bind(L):
    vp = ctypes.c_void_p
    if getattr(L, "_pk_serve_bound", False):
        return L
    L.pk_serve_start.argtypes = [vp, ctypes.c_int32]
    L.pk_serve_stop.argtypes = [vp]
    L.pk_serve_submit.argtypes = [vp, ctypes.POINTER(PkRequest), ctypes.POINTER(ctypes.c_uint64)]
    L.pk_serve_wait.argtypes = [vp, ctypes.c_uint64, ctypes.c_double, ctypes.POINTER(PkReqInfo)]
    L.pk_serve_cancel.argtypes = [vp, ctypes.c_uint64]
    L.pk_serve_release.argtypes = [vp, ctypes.c_uint64]
    L.pk_serve_stats_get.argtypes = [vp, ctypes.POINTER(PkServeStats)]
    for n in ("pk_serve_start", "pk_serve_stop", "pk_serve_submit", "pk_serve_wait", "pk_serve_cancel",
              "pk_serve_release", "pk_serve_stats_get"):
        getattr(L, n).restype = ctypes.c_int
    L._pk_serve_bound = True
    return L


class Request:
    """A submitted job; keeps its input, output and PSF buffers alive."""

    def __init__(self, server, rid, keep, out):
        self.server, self.id, self._keep, self.out = server, rid, keep, out
        self._done = False

    def info(self, wait_ms=0.0) -> dict:
        i = PkReqInfo()
        self.server.ctx.check(self.server.L.pk_serve_wait(self.server.ctx.handle, self.id, float(wait_ms),
                                                          ctypes.byref(i)))
        if i.state >= DONE:
            self._keep = None  # the server is done with the input (pinned staging goes back to the cache)
            self._done = True
        return {"state": i.state, "status": i.status, "iterations_done": i.iterations_done,
                "latency_ms": i.latency_ms, "queue_ms": i.queue_ms}

    def result(self, timeout_s=None):
        """Wait for the job; returns the deconvolved volume or raises PkDeconError."""
        inf = self.info(-1.0 if timeout_s is None else timeout_s * 1e3)
        if inf["state"] == DONE:
            return self.out
        if inf["state"] in (FAILED, CANCELLED):
            raise PkDeconError(inf["status"], _lib.load().pk_last_error(self.server.ctx.handle).decode())
        raise TimeoutError(f"request {self.id} not finished")

    def cancel(self):
        self.server.ctx.check(self.server.L.pk_serve_cancel(self.server.ctx.handle, self.id))

    def __del__(self):
        # a finished request's record is dropped from the server (pk_serve_release);
        # one not seen finished stays until the server stops
        try:
            srv = self.server
            if self._done and srv is not None and srv.ctx is not None and srv.ctx.handle and not srv.stopped:
                srv.L.pk_serve_release(srv.ctx.handle, self.id)
        except Exception:  # noqa: BLE001 -- interpreter shutdown or a stopped server
            pass


class Server:
    """Continuous-batching decon server on one device.  The context's HBM
    budget (pk_ctx_create) bounds the device memory of requests in flight."""

    def __init__(self, ctx=None, max_active: int = 32):
        self.ctx = ctx or _lib.default_context()
        self.L = _bind(_lib.load())
        self.stopped = False
        self.ctx.check(self.L.pk_serve_start(self.ctx.handle, int(max_active)))

    def submit(self, I, PSF_f, PSF_b=None, NUMIT=2, method="omw", **kwargs) -> Request:
        o = _parse(kwargs)
        m = {"omw": 1, "rl": 0, "simplified": 0}[method]
        opts, _ = _opts_struct(o, (1, 1, 1), False)
        pinned_in = _pinned_host_vol(I)
        obj, vin, kind = pinned_in if pinned_in is not None else _vol_of(I)
        d = tuple(vin.dims)
        if kind == "torch":
            import torch
            out = torch.empty((d[2], d[1], d[0]), dtype=torch.uint16 if o["save16bit"] else torch.float32,
                              device=obj.device).permute(2, 1, 0)
            # the input's producers and the output block's earlier users ran on
            # torch's stream; the server's stream runs after them (a device-wide
            # sync would also wait for the requests already in flight)
            torch.cuda.current_stream(obj.device).synchronize()
            vout = PkVol(ctypes.c_void_p(out.data_ptr()), 1 if o["save16bit"] else 2,
                         _lib.PK_MEM_DEVICE, (ctypes.c_int64 * 3)(*d))
        else:
            # host volumes are staged in pinned memory (torch's caching host
            # allocator): the server's copies then run asynchronously, beside
            # the steps, instead of blocking the scheduler thread
            pin_in = None if pinned_in is not None else _pinned((d[2], d[1], d[0]), obj.dtype)
            if pin_in is not None:
                import torch
                pin_in.permute(2, 1, 0).copy_(torch.from_numpy(obj))  # same layout; torch's parallel copy
                obj = pin_in
                vin = PkVol(ctypes.c_void_p(pin_in.data_ptr()), vin.dtype, _lib.PK_MEM_HOST, (ctypes.c_int64 * 3)(*d))
            odt = np.uint16 if o["save16bit"] else np.float32
            pin_out = _pinned((d[2], d[1], d[0]), odt)
            if pin_out is not None:
                out = pin_out.numpy().T  # a column-major (d0, d1, d2) view of the pinned buffer
                optr = pin_out.data_ptr()
                obj = (obj, pin_out)
            else:
                out = np.empty(d, dtype=odt, order="F")
                optr = out.ctypes.data
            vout = PkVol(ctypes.c_void_p(optr), 1 if o["save16bit"] else 2, _lib.PK_MEM_HOST, (ctypes.c_int64 * 3)(*d))
        pf, fd = _psf_array(PSF_f)
        pb, bd = _psf_array(PSF_b if m == 1 else PSF_f)
        q = PkRequest(m, int(NUMIT), vin, vout, ctypes.c_void_p(pf.ctypes.data), (ctypes.c_int64 * 3)(*fd),
                      ctypes.c_void_p(pb.ctypes.data), (ctypes.c_int64 * 3)(*bd), opts)
        rid = ctypes.c_uint64()
        self.ctx.check(self.L.pk_serve_submit(self.ctx.handle, ctypes.byref(q), ctypes.byref(rid)))
        return Request(self, rid.value, (obj, pf, pb, out), out)

    def stats(self) -> dict:
        s = PkServeStats()
        self.ctx.check(self.L.pk_serve_stats_get(self.ctx.handle, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in PkServeStats._fields_}

    def stop(self):
        if self.ctx is not None and self.ctx.handle:
            self.ctx.check(self.L.pk_serve_stop(self.ctx.handle))
        self.stopped = True
