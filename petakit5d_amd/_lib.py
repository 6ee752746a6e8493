"""ctypes binding of libpkdecon.so (include/pkdecon.h).

The shared library is the product: every decon call below goes through it.
If it is missing, loading fails loudly -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
# PKDECON_LIB selects another build of the library (e.g. one compiled with
# different launch attributes); default: the in-tree build.
LIB_PATH = os.environ.get("PKDECON_LIB") or os.path.join(_HERE, "libpkdecon.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "pkdecon.h")

PK_OK = 0
STATUS = {
    0: "PK_OK",
    1: "PK_ERR_INVALID_ARG",
    2: "PK_ERR_OUT_SIZE",
    3: "PK_ERR_UNSUPPORTED_SIZE",
    4: "PK_ERR_UNSUPPORTED_TYPE",
    5: "PK_ERR_HIP",
    6: "PK_ERR_NO_MEMORY",
    7: "PK_ERR_BUDGET",
    8: "PK_ERR_EMPTY_PSF",
    9: "PK_ERR_EMPTY_MASK",
    10: "PK_ERR_NO_DEVICE",
    11: "PK_ERR_BUSY",
    12: "PK_ERR_CANCELLED",
}
PK_UINT16, PK_SINGLE, PK_UINT8 = 1, 2, 3
PK_MEM_HOST, PK_MEM_DEVICE = 0, 1

# every symbol include/pkdecon.h declares
EXPORTS = [
    "pk_opts_default",
    "pk_ctx_create",
    "pk_ctx_destroy",
    "pk_reset",
    "pk_set_stream",
    "pk_last_error",
    "pk_set_iter_hook",
    "pk_device_bytes",
    "pk_decon_lucy",
    "pk_decon_lucy_omw",
    "pk_decon_lucy_omw_noalloc",
    "pk_decon_batch",
    "pk_edge_erosion_mask",
    "pk_serve_start",
    "pk_serve_stop",
    "pk_serve_submit",
    "pk_serve_wait",
    "pk_serve_cancel",
    "pk_serve_release",
    "pk_serve_stats_get",
    "pk_psf2otf",
    "pk_bp_opts_default",
    "pk_omw_backprojector",
    "pk_wb_backprojector",
    "pk_autotune",
    "pk_tune_set",
    "pk_fft_r2c",
    "pk_fft_c2r",
    "pk_profile_enable",
    "pk_profile_read",
    "pk_probe_copy",
    "pk_affine_warp3d_cubic",
    "pk_deskew_rotate_frame3d_cubic",
    "pk_deskew_rotate_resample_out_dims",
    "pk_deskew_rotate_frame3d_resample",
    "pk_find_good_factor_number",
    "pk_lzw_decode",
    "pk_lzw_encode",
    "pk_lzw_encode_bound",
    "pk_lz4_decode",
    "pk_lz4_encode",
    "pk_lz4_encode_bound",
    "pk_deskew_out_dims",
    "pk_deskew_frame3d",
    "pk_deskew_rotate_out_dims",
    "pk_deskew_rotate_frame3d",
    "pk_affine_warp3d",
]


class PkVol(ctypes.Structure):
    _fields_ = [
        ("data", ctypes.c_void_p),
        ("dtype", ctypes.c_int32),
        ("mem", ctypes.c_int32),
        ("dims", ctypes.c_int64 * 3),
    ]


class PkOpts(ctypes.Structure):
    _fields_ = [
        ("background", ctypes.c_double),
        ("damp_factor", ctypes.c_double),
        ("scale_factor", ctypes.c_double),
        ("decon_offset", ctypes.c_double),
        ("edge_erosion", ctypes.c_int32),
        ("save16bit", ctypes.c_int32),
        ("has_bbox", ctypes.c_int32),
        ("flags", ctypes.c_int32),
        ("bbox", ctypes.c_int64 * 6),
    ]


class PkBpOpts(ctypes.Structure):
    _fields_ = [
        ("otf_cum_thresh", ctypes.c_double),
        ("otf_area_thresh", ctypes.c_double),
        ("hann_win_lo", ctypes.c_double),
        ("hann_win_hi", ctypes.c_double),
        ("min_int_thresh", ctypes.c_double),
    ]


class PkProfEntry(ctypes.Structure):
    _fields_ = [
        ("name", ctypes.c_char * 32),
        ("launches", ctypes.c_int64),
        ("total_ms", ctypes.c_double),
        ("bytes_per_launch", ctypes.c_double),
    ]


class PkDeconError(RuntimeError):
    """A non-OK pk_status; .status holds the code name."""

    def __init__(self, status: int, msg: str):
        self.code = status
        self.status = STATUS.get(status, str(status))
        super().__init__(f"{self.status}: {msg}")


_lib = None
_lock = threading.Lock()


def load():
    """Load libpkdecon.so (raises OSError if it has not been built)."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} is missing: build it with petakit5d_amd/build_lib.sh "
                          "(or __graft_entry__.build()); there is no CPU fallback")
        # torch wheels bundle their own HIP/HSA runtime (NEEDED "libamdhip64.so").
        # Loading torch first lets glibc resolve this library's NEEDED
        # "libamdhip64.so.7" / "libhsa-runtime64.so.1" to that same runtime by
        # SONAME, so the process holds one HSA runtime instead of two competing
        # for the device.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        i64p = ctypes.POINTER(ctypes.c_int64)
        L.pk_opts_default.argtypes = [ctypes.POINTER(PkOpts)]
        L.pk_opts_default.restype = None
        L.pk_ctx_create.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.POINTER(vp)]
        L.pk_ctx_destroy.argtypes = [vp]
        L.pk_ctx_destroy.restype = None
        L.pk_reset.argtypes = [vp]
        L.pk_set_stream.argtypes = [vp, vp]
        L.pk_last_error.argtypes = [vp]
        L.pk_last_error.restype = ctypes.c_char_p
        L.pk_device_bytes.argtypes = [vp]
        L.pk_device_bytes.restype = ctypes.c_uint64
        L.pk_decon_lucy.argtypes = [vp, ctypes.POINTER(PkVol), vp, i64p, ctypes.c_int32, ctypes.POINTER(PkOpts),
                                    ctypes.POINTER(PkVol), vp]
        L.pk_decon_lucy_omw.argtypes = [vp, ctypes.POINTER(PkVol), vp, i64p, vp, i64p, ctypes.c_int32,
                                        ctypes.POINTER(PkOpts), ctypes.POINTER(PkVol), vp]
        L.pk_decon_lucy_omw_noalloc.argtypes = [vp, ctypes.POINTER(PkVol), vp, i64p, vp, i64p, ctypes.c_int32,
                                                ctypes.POINTER(PkOpts), ctypes.POINTER(PkVol), vp]
        L.pk_psf2otf.argtypes = [vp, vp, i64p, i64p, vp]
        L.pk_decon_batch.argtypes = [vp, ctypes.c_int32, vp, ctypes.c_int32, vp, i64p, vp, i64p, ctypes.c_int32,
                                     ctypes.POINTER(PkOpts)]
        L.pk_bp_opts_default.argtypes = [ctypes.POINTER(PkBpOpts)]
        L.pk_bp_opts_default.restype = None
        L.pk_omw_backprojector.argtypes = [vp, vp, i64p, ctypes.c_double, ctypes.c_int32, ctypes.POINTER(PkBpOpts),
                                           vp, vp, vp, vp, ctypes.POINTER(ctypes.c_int32)]
        L.pk_wb_backprojector.argtypes = [vp, vp, i64p, ctypes.c_int32, ctypes.c_double, ctypes.c_double,
                                          ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(ctypes.c_double), vp, vp,
                                          ctypes.POINTER(ctypes.c_double)]
        L.pk_autotune.argtypes = [vp, i64p, ctypes.c_char_p, ctypes.c_size_t]
        L.pk_tune_set.argtypes = [vp, i64p, ctypes.POINTER(ctypes.c_int32)]
        L.pk_fft_r2c.argtypes = [vp, vp, vp, i64p]
        L.pk_fft_c2r.argtypes = [vp, vp, vp, i64p]
        L.pk_profile_enable.argtypes = [vp, ctypes.c_int32]
        L.pk_profile_read.argtypes = [vp, ctypes.POINTER(PkProfEntry), ctypes.c_int32]
        L.pk_profile_read.restype = ctypes.c_int32
        L.pk_probe_copy.argtypes = [vp, ctypes.c_int32, i64p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                    ctypes.c_int32, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
        L.pk_find_good_factor_number.argtypes = [ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]
        L.pk_find_good_factor_number.restype = ctypes.c_int64
        for name in EXPORTS:
            if name not in ("pk_opts_default", "pk_bp_opts_default", "pk_ctx_destroy", "pk_last_error", "pk_device_bytes",
                            "pk_find_good_factor_number", "pk_profile_read", "pk_lzw_encode_bound",
                            "pk_lz4_encode_bound"):
                getattr(L, name).restype = ctypes.c_int
        _lib = L
        return _lib


def i64x3(v):
    return (ctypes.c_int64 * 3)(*[int(x) for x in v])


class Context:
    """A device context: stream, workspace and OTF cache (the reference's
    persistent OTF variables, decon_lucy_function.m:61-77)."""

    def __init__(self, device: int = 0, hbm_budget_bytes: int = 0):
        self._L = load()
        h = ctypes.c_void_p()
        st = self._L.pk_ctx_create(int(device), int(hbm_budget_bytes), ctypes.byref(h))
        if st != PK_OK:
            raise PkDeconError(st, "pk_ctx_create failed")
        self.handle = h
        self.device = device

    def check(self, st: int):
        if st != PK_OK:
            raise PkDeconError(st, self._L.pk_last_error(self.handle).decode(errors="replace"))

    def reset(self):
        self.check(self._L.pk_reset(self.handle))

    def set_stream(self, stream_ptr: int | None):
        self.check(self._L.pk_set_stream(self.handle, ctypes.c_void_p(stream_ptr or 0)))

    def autotune(self, dims) -> str:
        """Measure and keep the fastest launch shapes / pass order for a
        padded size (pk_autotune); returns the report line."""
        buf = ctypes.create_string_buffer(2048)
        self.check(self._L.pk_autotune(self.handle, i64x3(tuple(int(x) for x in dims)), buf, len(buf)))
        return buf.value.decode(errors="replace")

    def tune_set(self, dims, shapes):
        """Set the launch shapes of a padded size without measuring
        (pk_tune_set): (xE, yE, yNT, yT, zE, zNT, zT, conv_axis)."""
        arr = (ctypes.c_int32 * 8)(*[int(x) for x in shapes])
        self.check(self._L.pk_tune_set(self.handle, i64x3(tuple(int(x) for x in dims)), arr))

    def profile(self, enable: bool):
        """Start (clearing) or stop per-pass HIP-event timing."""
        self.check(self._L.pk_profile_enable(self.handle, 1 if enable else 0))

    def profile_read(self) -> dict:
        """{pass name: (launches, total_ms, algorithmic bytes per launch)}."""
        n = self._L.pk_profile_read(self.handle, None, 0)
        buf = (PkProfEntry * max(n, 1))()
        n = self._L.pk_profile_read(self.handle, buf, n)
        return {buf[i].name.decode(): (int(buf[i].launches), float(buf[i].total_ms), float(buf[i].bytes_per_launch))
                for i in range(n)}

    def probe_copy(self, pattern: int, dims, tile: int = 16, vals: int = 16, xcd: bool = True, reps: int = 20,
                   cold: bool = False):
        """(ms per copy, bytes per copy) of pk_probe_copy: pattern 0 sequential,
        1 dim-2 columns, 2 dim-3 columns, 3 dim-3 columns with a second input
        of a dims-sized transform's spectrum; cold: each copy after the
        memory-side cache is flushed (1 GiB read), else back to back."""
        ms, nb = ctypes.c_double(), ctypes.c_double()
        flags = (1 if xcd else 0) | (2 if cold else 0)
        self.check(self._L.pk_probe_copy(self.handle, int(pattern), i64x3(tuple(int(x) for x in dims)), int(tile),
                                         int(vals), flags, int(reps), ctypes.byref(ms), ctypes.byref(nb)))
        return ms.value, nb.value

    def device_bytes(self) -> int:
        return int(self._L.pk_device_bytes(self.handle))

    def close(self):
        if getattr(self, "handle", None):
            self._L.pk_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = None


def default_context() -> Context:
    """Process-wide context on device 0 (created on first use)."""
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


def find_good_factor_number(n: int, direction: int = 1, allow_odd_seven: bool = False, max_prime: int = 7) -> int:
    return int(load().pk_find_good_factor_number(int(n), int(direction), 1 if allow_odd_seven else 0,
                                                 int(max_prime)))
