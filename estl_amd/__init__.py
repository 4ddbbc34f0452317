"""estl_amd -- Python binding of the gfx950 FFT hot path (libestl_hip.so).

The compute path is the HIP library behind the C ABI of ``include/estl_hip.h``;
this module only loads it with ctypes and passes device pointers (from torch
tensors, which serve as device-memory and stream plumbing).  There is no CPU
fallback: if the library is missing, ``lib()`` raises.

The Python names mirror the reference's FFT interface (FFT.h:19-52):
``FFT_HIP(n, forward=True, batch_count=1)`` with ``execute()`` on host arrays, and
``Plan`` for the asynchronous device-resident form.
"""
from __future__ import annotations

import ctypes
import json
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libestl_hip.so")

FORWARD = -1
BACKWARD = +1
FP32 = 0
FP16 = 1
FLAG_NORMALIZE = 0x1
FLAG_AUTOTUNE = 0x2
FLAG_PLANAR = 0x4
CONV_CORRELATE = 0x10
CONV_NORMALIZE_KERNEL = 0x20
CONV_CENTER = 0x40
CONV_FP16 = 0x80
CONV_FULL_LENGTH = 0x100

OK = 0
ERR_INVALID = -1
ERR_HIP = -2
ERR_NOMEM = -3
ERR_UNSUPPORTED = -4
ERR_BUSY = -5

_lib = None


class EstlError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"estl error {code}: {msg}")
        self.code = code


def _sig(L, name, restype, argtypes):
    f = getattr(L, name)
    f.restype = restype
    f.argtypes = argtypes
    return f


def lib():
    """Load libestl_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m estl_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    vp, cl, ci, cu = ctypes.c_void_p, ctypes.c_long, ctypes.c_int, ctypes.c_uint
    ll, u64 = ctypes.c_longlong, ctypes.c_uint64
    _sig(L, "estl_last_error", ctypes.c_char_p, [])
    _sig(L, "estl_version", ci, [])
    _sig(L, "estl_init", ci, [ci])
    _sig(L, "estl_plan_create", ci, [ctypes.POINTER(vp), cl, cl, ci, ci, cu])
    _sig(L, "estl_plan_create_split", ci, [ctypes.POINTER(vp), cl, cl, ci, ci, cu, ci, ctypes.POINTER(ci),
                                              ctypes.POINTER(ci)])
    _sig(L, "estl_plan_destroy", ci, [vp])
    _sig(L, "estl_execute_device", ci, [vp, vp, vp, vp])
    _sig(L, "estl_execute_device_noalloc", ci, [vp, vp, vp, vp])
    _sig(L, "estl_execute_host", ci, [vp, vp, vp])
    _sig(L, "estl_max_batch", ll, [cl, ci, ll])
    _sig(L, "estl_plan_workspace_bytes", ll, [vp])
    _sig(L, "estl_plan_describe", ci, [vp, ctypes.c_char_p, ci])
    _sig(L, "estl_plan_launches", ci, [vp])
    _sig(L, "estl_fill_uniform", ci, [vp, cl, u64, u64, vp])
    _sig(L, "estl_stream_copy", ci, [vp, vp, ll, ci, vp])
    _sig(L, "estl_plan_profile", ci, [vp, vp, vp, ci, ctypes.c_char_p, ci])
    _sig(L, "estl_plan_kernel_flops", ctypes.c_double, [vp, ci])
    _sig(L, "estl_plan_kernel_ms", ctypes.c_double, [vp, ci])
    _sig(L, "estl_plan_kernel_info", ci, [vp, ci, ctypes.c_char_p, ci])
    _sig(L, "estl_tune", ci, [cl, cl, ci, ctypes.c_char_p, ci])
    _sig(L, "estl_tune_precision", ci, [cl, cl, ci, ci, ctypes.c_char_p, ci])
    lp = ctypes.POINTER(ctypes.c_long)
    _sig(L, "estl_ubatch_create", ci, [ctypes.POINTER(vp), ci, lp, lp, ci, cu])
    _sig(L, "estl_ubatch_execute", ci, [vp, vp, vp, vp])
    _sig(L, "estl_ubatch_execute_graph", ci, [vp, vp, vp, vp])
    _sig(L, "estl_ubatch_describe", ci, [vp, ctypes.c_char_p, ci])
    _sig(L, "estl_ubatch_destroy", ci, [vp])
    _sig(L, "estl_plan2d_create", ci, [ctypes.POINTER(vp), cl, cl, cl, ci, cu])
    _sig(L, "estl_plan2d_execute", ci, [vp, vp, vp, vp])
    _sig(L, "estl_plan2d_execute_host", ci, [vp, vp, vp])
    _sig(L, "estl_plan2d_launches", ci, [vp])
    _sig(L, "estl_plan2d_destroy", ci, [vp])
    _sig(L, "estl_pack_fp16", ci, [vp, vp, vp, cl, cl, vp])
    _sig(L, "estl_unpack_fp16", ci, [vp, vp, vp, cl, cl, vp])
    _sig(L, "estl_execute_fp16", ci, [vp, vp, vp, vp, vp, vp])
    _sig(L, "estl_conv_create", ci, [ctypes.POINTER(vp), cl, cl, ci, cu])
    _sig(L, "estl_conv_set_filter", ci, [vp, vp, cl, vp])
    _sig(L, "estl_conv_execute", ci, [vp, vp, vp, vp])
    _sig(L, "estl_conv_execute_fp16", ci, [vp, vp, vp, vp, vp, vp])
    _sig(L, "estl_conv_workspace_bytes", ll, [vp])
    _sig(L, "estl_conv_destroy", ci, [vp])
    llp = ctypes.POINTER(ll)
    _sig(L, "estl_server_create", ci, [ctypes.POINTER(vp), ll, ci])
    _sig(L, "estl_server_submit_fft", ci, [vp, cl, cl, ci, vp, vp, llp])
    _sig(L, "estl_server_submit_conv", ci, [vp, cl, cl, vp, vp, cl, cu, vp, llp])
    _sig(L, "estl_server_register_filter", ci, [vp, cl, vp, cl, cu, vp, llp])
    _sig(L, "estl_server_submit_conv_fp16", ci, [vp, cl, cl, vp, vp, vp, cl, cu, vp, llp])
    _sig(L, "estl_server_submit_conv_filter_fp16", ci, [vp, ll, cl, vp, vp, cu, vp, llp])
    _sig(L, "estl_server_submit_conv_filter", ci, [vp, ll, cl, vp, cu, vp, llp])
    _sig(L, "estl_server_next_done", ci, [vp, llp, ci, ctypes.c_double])
    _sig(L, "estl_server_poll", ci, [vp, ll, ctypes.POINTER(ci), ctypes.POINTER(vp),
                                     ctypes.POINTER(ctypes.c_double)])
    _sig(L, "estl_server_wait", ci, [vp, ll, ctypes.c_double])
    _sig(L, "estl_server_copy_output", ci, [vp, ll, vp, ll, vp])
    _sig(L, "estl_server_release", ci, [vp, ll])
    _sig(L, "estl_server_drain", ci, [vp, ctypes.c_double])
    _sig(L, "estl_server_stats", ci, [vp, ctypes.c_char_p, ci])
    _sig(L, "estl_server_reset_stats", ci, [vp])
    _sig(L, "estl_server_destroy", ci, [vp])
    _lib = L
    return L


def check(rc: int) -> int:
    if rc < 0:
        raise EstlError(rc, lib().estl_last_error().decode())
    return rc


def _stream_ptr(stream) -> int | None:
    if stream is None:
        import torch
        return torch.cuda.current_stream().cuda_stream
    if isinstance(stream, int):
        return stream
    return stream.cuda_stream


def init(device: int = 0) -> None:
    check(lib().estl_init(device))


class Plan:
    """Device plan for `batch` rows of length n (C ABI estl_plan_create)."""

    def __init__(self, n: int, batch: int = 1, direction: int = FORWARD, precision: int = FP32,
                 flags: int = 0, split: list[int] | None = None, tiles: list[int] | None = None):
        self.n, self.batch, self.direction, self.precision, self.flags = n, batch, direction, precision, flags
        self._h = ctypes.c_void_p()
        h = ctypes.c_void_p()
        if split:
            arr = (ctypes.c_int * len(split))(*split)
            tarr = (ctypes.c_int * len(split))(*tiles) if tiles else None
            check(lib().estl_plan_create_split(ctypes.byref(h), n, batch, direction, precision, flags,
                                               len(split), arr, tarr))
        else:
            check(lib().estl_plan_create(ctypes.byref(h), n, batch, direction, precision, flags))
        self._h = h

    @property
    def handle(self) -> int:
        return self._h.value

    def execute(self, d_in, d_out, stream=None) -> None:
        """Run on device tensors (or raw device pointers) on `stream` (torch stream)."""
        pin = d_in if isinstance(d_in, int) else d_in.data_ptr()
        pout = d_out if isinstance(d_out, int) else d_out.data_ptr()
        check(lib().estl_execute_device(self._h, pin, pout, _stream_ptr(stream)))

    def execute_noalloc(self, d_in, d_out, stream=None) -> None:
        """The same with no device allocation: scratch from the library's static 64 MiB
        array of this GPU (plans of at most 2^23 values; calls overlapping on one GPU
        race)."""
        pin = d_in if isinstance(d_in, int) else d_in.data_ptr()
        pout = d_out if isinstance(d_out, int) else d_out.data_ptr()
        check(lib().estl_execute_device_noalloc(self._h, pin, pout, _stream_ptr(stream)))

    def execute_fp16(self, d_in, in_scale, d_out, out_scale, stream=None) -> None:
        """FP16-storage plan: binary16 (re, im) buffers with one FP32 scale per row."""
        p = [t if isinstance(t, int) else t.data_ptr() for t in (d_in, in_scale, d_out, out_scale)]
        check(lib().estl_execute_fp16(self._h, p[0], p[1], p[2], p[3], _stream_ptr(stream)))

    def execute_host(self, h_in, h_out) -> None:
        """Synchronous host-to-host execute (numpy arrays, complex64)."""
        check(lib().estl_execute_host(self._h, h_in.ctypes.data, h_out.ctypes.data))

    def describe(self) -> dict:
        buf = ctypes.create_string_buffer(4096)
        check(lib().estl_plan_describe(self._h, buf, 4096))
        return json.loads(buf.value.decode())

    def profile(self, reps: int = 10, d_in=None, d_out=None) -> dict:
        """Time each launch of the plan (on its own buffers unless given): per-launch
        ms, flops (5 n log2 radix per transform), GFLOP/s and share, as JSON."""
        pin = None if d_in is None else (d_in if isinstance(d_in, int) else d_in.data_ptr())
        pout = None if d_out is None else (d_out if isinstance(d_out, int) else d_out.data_ptr())
        buf = ctypes.create_string_buffer(1 << 14)
        check(lib().estl_plan_profile(self._h, pin, pout, reps, buf, 1 << 14))
        return json.loads(buf.value.decode())

    def launches(self) -> int:
        return check(lib().estl_plan_launches(self._h))

    def workspace_bytes(self) -> int:
        return check(lib().estl_plan_workspace_bytes(self._h))

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().estl_plan_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Plan2D:
    """2-D transforms of `batch` ny x nx arrays (x fastest; C ABI estl_plan2d_*)."""

    def __init__(self, nx: int, ny: int, batch: int = 1, direction: int = FORWARD, flags: int = 0):
        self.nx, self.ny, self.batch, self.direction, self.flags = nx, ny, batch, direction, flags
        self._h = ctypes.c_void_p()
        h = ctypes.c_void_p()
        check(lib().estl_plan2d_create(ctypes.byref(h), nx, ny, batch, direction, flags))
        self._h = h

    def execute(self, d_in, d_out, stream=None) -> None:
        for t in (d_in, d_out):
            if not isinstance(t, int) and t.numel() * t.element_size() < 8 * self.nx * self.ny * self.batch:
                raise ValueError("buffer smaller than batch * ny * nx complex64")
        pin = d_in if isinstance(d_in, int) else d_in.data_ptr()
        pout = d_out if isinstance(d_out, int) else d_out.data_ptr()
        check(lib().estl_plan2d_execute(self._h, pin, pout, _stream_ptr(stream)))

    def execute_host(self, h_in, h_out) -> None:
        check(lib().estl_plan2d_execute_host(self._h, h_in.ctypes.data, h_out.ctypes.data))

    def launches(self) -> int:
        return check(lib().estl_plan2d_launches(self._h))

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().estl_plan2d_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class UnevenBatch:
    """Transforms of different power-of-two sizes in one launch (C ABI estl_ubatch_*)."""

    def __init__(self, sizes, offsets=None, direction: int = FORWARD, flags: int = 0):
        n = len(sizes)
        self.sizes = list(sizes)
        szs = (ctypes.c_long * n)(*self.sizes)
        offs = (ctypes.c_long * n)(*offsets) if offsets is not None else None
        self._h = ctypes.c_void_p()
        h = ctypes.c_void_p()
        check(lib().estl_ubatch_create(ctypes.byref(h), n, szs, offs, direction, flags))
        self._h = h
        self.info = self.describe()

    def describe(self) -> dict:
        buf = ctypes.create_string_buffer(1024)
        check(lib().estl_ubatch_describe(self._h, buf, 1024))
        return json.loads(buf.value.decode())

    def _check_buffers(self, d_in, d_out):
        for t in (d_in, d_out):
            if not isinstance(t, int) and t.numel() < self.info["extent"]:
                raise ValueError(f"buffer holds {t.numel()} values, the batch needs {self.info['extent']}")

    def execute(self, d_in, d_out, stream=None, graph: bool = False) -> None:
        self._check_buffers(d_in, d_out)
        pin = d_in if isinstance(d_in, int) else d_in.data_ptr()
        pout = d_out if isinstance(d_out, int) else d_out.data_ptr()
        fn = lib().estl_ubatch_execute_graph if graph else lib().estl_ubatch_execute
        check(fn(self._h, pin, pout, _stream_ptr(stream)))

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().estl_ubatch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Convolution:
    """Batched circular convolution of real rows (C ABI estl_conv_*)."""

    def __init__(self, n: int, rows: int, flags: int = 0, precision: int = FP32):
        self.n, self.rows, self.flags = n, rows, flags
        self._h = ctypes.c_void_p()
        h = ctypes.c_void_p()
        check(lib().estl_conv_create(ctypes.byref(h), n, rows, precision, flags))
        self._h = h

    def set_filter(self, d_h, k: int | None = None, stream=None) -> None:
        p = d_h if isinstance(d_h, int) else d_h.data_ptr()
        k = k if k is not None else d_h.numel()
        check(lib().estl_conv_set_filter(self._h, p, k, _stream_ptr(stream)))

    def execute(self, d_x, d_y, stream=None) -> None:
        for t in (d_x, d_y):
            if not isinstance(t, int) and t.numel() < self.n * self.rows:
                raise ValueError("buffer smaller than rows * n")
        px = d_x if isinstance(d_x, int) else d_x.data_ptr()
        py = d_y if isinstance(d_y, int) else d_y.data_ptr()
        check(lib().estl_conv_execute(self._h, px, py, _stream_ptr(stream)))

    def execute_fp16(self, d_x16, x_scale, d_y16, y_scale, stream=None) -> None:
        """FP16-storage object: binary16 real rows with one FP32 scale per row."""
        p = [t if isinstance(t, int) else t.data_ptr() for t in (d_x16, x_scale, d_y16, y_scale)]
        check(lib().estl_conv_execute_fp16(self._h, p[0], p[1], p[2], p[3], _stream_ptr(stream)))

    def workspace_bytes(self) -> int:
        return check(lib().estl_conv_workspace_bytes(self._h))

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().estl_conv_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


REQ_QUEUED, REQ_RUNNING, REQ_DONE, REQ_FAILED = 0, 1, 2, 3


class Server:
    """Request loop over the FFT and convolution pipelines (C ABI estl_server_*).

    Requests are admitted in arrival order against an HBM budget (one arena), run
    concurrently on ``streams`` streams, and each result stays in the arena until
    ``release``; ``output(id, ...)`` copies it into a new tensor.
    """

    def __init__(self, budget_bytes: int, streams: int = 4):
        self._h = ctypes.c_void_p()
        h = ctypes.c_void_p()
        check(lib().estl_server_create(ctypes.byref(h), int(budget_bytes), int(streams)))
        self._h = h

    @staticmethod
    def _ptr(t) -> int:
        return t if isinstance(t, int) else t.data_ptr()

    def submit_fft(self, n: int, batch: int, d_in, direction: int = FORWARD, stream=None) -> int:
        if not isinstance(d_in, int) and d_in.numel() * d_in.element_size() < 8 * n * batch:
            raise ValueError("input smaller than batch * n complex values")
        rid = ctypes.c_longlong()
        check(lib().estl_server_submit_fft(self._h, n, batch, direction, self._ptr(d_in), _stream_ptr(stream),
                                           ctypes.byref(rid)))
        return rid.value

    def submit_conv(self, n: int, rows: int, d_x, d_h, k: int | None = None, flags: int = 0, stream=None) -> int:
        if not isinstance(d_x, int) and d_x.numel() < n * rows:
            raise ValueError("input smaller than rows * n")
        k = k if k is not None else d_h.numel()
        rid = ctypes.c_longlong()
        check(lib().estl_server_submit_conv(self._h, n, rows, self._ptr(d_x), self._ptr(d_h), k, flags,
                                            _stream_ptr(stream), ctypes.byref(rid)))
        return rid.value

    def register_filter(self, n: int, d_h, k: int | None = None, flags: int = 0, stream=None) -> int:
        """Keep the spectrum of filter d_h for rows of length n resident; returns its id."""
        k = k if k is not None else d_h.numel()
        fid = ctypes.c_longlong()
        check(lib().estl_server_register_filter(self._h, n, self._ptr(d_h), k, flags, _stream_ptr(stream),
                                                ctypes.byref(fid)))
        return fid.value

    def submit_conv_filter(self, filter_id: int, rows: int, d_x, flags: int = 0, stream=None) -> int:
        rid = ctypes.c_longlong()
        check(lib().estl_server_submit_conv_filter(self._h, filter_id, rows, self._ptr(d_x), flags,
                                                   _stream_ptr(stream), ctypes.byref(rid)))
        return rid.value

    def submit_conv_fp16(self, n: int, rows: int, d_x16, x_scale, d_h, k: int | None = None, flags: int = 0,
                         stream=None) -> int:
        """FP16 storage: binary16 real rows with per-row FP32 scales (see output_fp16)."""
        k = k if k is not None else d_h.numel()
        rid = ctypes.c_longlong()
        check(lib().estl_server_submit_conv_fp16(self._h, n, rows, self._ptr(d_x16), self._ptr(x_scale),
                                                 self._ptr(d_h), k, flags, _stream_ptr(stream), ctypes.byref(rid)))
        return rid.value

    def submit_conv_filter_fp16(self, filter_id: int, rows: int, d_x16, x_scale, flags: int = 0,
                                stream=None) -> int:
        rid = ctypes.c_longlong()
        check(lib().estl_server_submit_conv_filter_fp16(self._h, filter_id, rows, self._ptr(d_x16),
                                                        self._ptr(x_scale), flags, _stream_ptr(stream),
                                                        ctypes.byref(rid)))
        return rid.value

    def output_fp16(self, rid: int, n: int, rows: int, stream=None):
        """(binary16 rows, FP32 row scales) of a DONE FP16 request, as new tensors."""
        import torch
        off = (2 * n * rows + 255) // 256 * 256
        raw = torch.empty(off + 4 * rows, dtype=torch.uint8, device="cuda")
        check(lib().estl_server_copy_output(self._h, rid, raw.data_ptr(), raw.numel(), _stream_ptr(stream)))
        return raw[:2 * n * rows].view(torch.float16), raw[off:off + 4 * rows].view(torch.float32)

    def next_done(self, max_ids: int = 256, timeout: float = 1.0) -> list:
        """Ids of requests that finished since the last call (oldest first)."""
        buf = (ctypes.c_longlong * max_ids)()
        c = check(lib().estl_server_next_done(self._h, buf, max_ids, timeout))
        return list(buf[:c])

    def poll(self, rid: int):
        """(state, device address of the output or 0, latency in s or -1)."""
        st, out, lat = ctypes.c_int(), ctypes.c_void_p(), ctypes.c_double()
        check(lib().estl_server_poll(self._h, rid, ctypes.byref(st), ctypes.byref(out), ctypes.byref(lat)))
        return st.value, out.value or 0, lat.value

    def wait(self, rid: int, timeout: float = 60.0) -> None:
        check(lib().estl_server_wait(self._h, rid, timeout))

    def output(self, rid: int, shape, dtype, stream=None):
        """A copy of the DONE request's output as a new torch tensor (on ``stream``)."""
        import torch
        t = torch.empty(*shape, dtype=dtype, device="cuda")
        check(lib().estl_server_copy_output(self._h, rid, t.data_ptr(), t.numel() * t.element_size(),
                                            _stream_ptr(stream)))
        return t

    def release(self, rid: int) -> None:
        check(lib().estl_server_release(self._h, rid))

    def drain(self, timeout: float = 120.0) -> None:
        check(lib().estl_server_drain(self._h, timeout))

    def stats(self) -> dict:
        buf = ctypes.create_string_buffer(2048)
        check(lib().estl_server_stats(self._h, buf, len(buf)))
        return json.loads(buf.value.decode())

    def reset_stats(self) -> None:
        check(lib().estl_server_reset_stats(self._h))

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().estl_server_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def pack_fp16(x, x16, scale, n: int, batch: int, stream=None) -> None:
    """complex64 device tensor -> binary16 (re, im) pairs + per-row max-abs scale."""
    check(lib().estl_pack_fp16(x.data_ptr(), x16.data_ptr(), scale.data_ptr(), n, batch, _stream_ptr(stream)))


def unpack_fp16(x16, scale, x, n: int, batch: int, stream=None) -> None:
    """binary16 pairs + per-row scale -> complex64 device tensor."""
    check(lib().estl_unpack_fp16(x16.data_ptr(), scale.data_ptr(), x.data_ptr(), n, batch, _stream_ptr(stream)))


def fill_uniform(t, seed: int, offset: int = 0, stream=None) -> None:
    """Fill a float32/complex64 device tensor with the seeded uniform[-1,1) stream."""
    import torch
    count = t.numel() * (2 if t.dtype == torch.complex64 else 1)
    check(lib().estl_fill_uniform(t.data_ptr(), count, seed, offset, _stream_ptr(stream)))


def stream_copy(src, dst, nontemporal: bool = True, stream=None) -> None:
    """Copy device tensor src into dst (same byte size, a multiple of 16) with the
    library's 16-byte streaming-copy kernel."""
    nbytes = src.numel() * src.element_size()
    if dst.numel() * dst.element_size() != nbytes:
        raise ValueError("stream_copy: src and dst differ in size")
    check(lib().estl_stream_copy(src.data_ptr(), dst.data_ptr(), nbytes, 1 if nontemporal else 0, _stream_ptr(stream)))


def tune(n: int, batch: int = 1, direction: int = FORWARD, precision: int = FP32) -> dict:
    """Autotune the plan for (n, batch, direction, precision); returns every candidate's time."""
    buf = ctypes.create_string_buffer(1 << 16)
    check(lib().estl_tune_precision(n, batch, direction, precision, buf, 1 << 16))
    return json.loads(buf.value.decode())


def max_batch(n: int, precision: int = FP32, budget_bytes: int = 0) -> int:
    return check(lib().estl_max_batch(n, precision, budget_bytes))


class FFT_HIP:
    """Mirror of the reference's backend interface (FFT.h:19-36, FFT_FFTW3.h:16-34):
    ``FFT_HIP(n, forward, batch_count)``, ``set_data``/``get_data`` by (index, batch
    index) and a synchronous ``execute()`` over host arrays."""

    @staticmethod
    def get_name() -> str:
        return "HIP FFT (gfx950)"

    @staticmethod
    def get_max_batch_size(n: int) -> int:
        return max_batch(n)

    def __init__(self, n: int, forward: bool = True, batch_count: int = 1):
        import numpy as np
        self.size, self.batch_count = n, batch_count
        self.plan = Plan(n, batch_count, FORWARD if forward else BACKWARD)
        self.data_in = np.zeros(n * batch_count, dtype=np.complex64)
        self.data_out = np.zeros(n * batch_count, dtype=np.complex64)

    def set_data(self, index: int, value: complex, batch_index: int = 0) -> None:
        self.data_in[batch_index * self.size + index] = value

    def get_data(self, index: int, batch_index: int = 0) -> complex:
        return complex(self.data_out[batch_index * self.size + index])

    def execute(self) -> None:
        self.plan.execute_host(self.data_in, self.data_out)
