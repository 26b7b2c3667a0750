"""Host-side front end of the MI355X FMM-FFT (libfmmfft_amd.so, include/fmmfft.h).

The reference exposes a plan/execute pair in C: ``mfti(lq, p, terms, b, w, v)``
builds every operator table (src/fmmfft.c:337-382) and ``mft(lq, qr, dir, p,
myid, terms, b, w, v, ..., plan)`` transforms rank ``myid``'s block in place
(src/fmmfft.c:980-1081).  This module keeps that shape over the C ABI:

* :func:`mfti` (``lq`` = N/P, ``p`` = P, ``terms``, ``b``) returns a :class:`Plan`
  whose operator tables were built on the GPU;
* :func:`mft` transforms all P blocks of a device tensor in place;
* :class:`Plan` adds out-of-place, batched and inverse transforms.

Tensors are torch tensors on the GPU (complex64 or complex128 per the plan);
torch only provides device memory and streams.  There is no CPU fallback:
if the HIP library is missing or no GPU is present every call raises.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfmmfft_amd.so")

C64 = 0
C128 = 1
C64_SPLIT = 2  # complex64 with the M2L flips as FP32 hi + lo split products
FORWARD = 1
INVERSE = -1

OK, EINVAL, ENOMEM, EHIP, ENOTSUP, ETOOBIG = 0, 1, 2, 3, 4, 5
STATIC_SCRATCH_BYTES = 64 << 20  # FMMFFT_STATIC_SCRATCH_BYTES

# Symbols declared in include/fmmfft.h.
EXPORTS = (
    "fmmfft_plan_sizes",
    "fmmfft_plan_create",
    "fmmfft_plan_destroy",
    "fmmfft_execute",
    "fmmfft_execute_batched",
    "fmmfft_workspace_bytes",
    "fmmfft_execute_ws",
    "fmmfft_execute_static",
    "fmmfft_plan_get",
    "fmmfft_read_table",
    "fmmfft_read_stage",
    "fmmfft_profile_begin",
    "fmmfft_profile_end",
    "fmmfft_server_create",
    "fmmfft_server_submit",
    "fmmfft_server_wait",
    "fmmfft_server_drain",
    "fmmfft_server_stats_get",
    "fmmfft_server_reset_stats",
    "fmmfft_server_destroy",
    "fmmfft_strerror",
    "fmmfft_last_hip_error",
)

_lib = None


class ServerStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in ("completed", "batches", "plans_created", "plans_evicted",
                                               "graphs_captured")] + \
               [(n, ctypes.c_double) for n in ("p50_ms", "p99_ms", "mean_ms", "max_ms", "span_s",
                                                "nominal_gflop")] + \
               [(n, ctypes.c_size_t) for n in ("bytes_in_use", "peak_bytes")]


class FmmFftError(RuntimeError):
    def __init__(self, status: int, what: str):
        self.status = status
        msg = _lib.fmmfft_strerror(status).decode() if _lib is not None else str(status)
        hip = _lib.fmmfft_last_hip_error() if _lib is not None else 0
        super().__init__(f"{what}: {msg} (status {status}, last HIP error {hip})")


def lib():
    """Load libfmmfft_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `make -C fmm-fft_amd/csrc`")
        L = ctypes.CDLL(LIB_PATH)
        i32, i64, vp, sz = ctypes.c_int32, ctypes.c_int64, ctypes.c_void_p, ctypes.c_size_t
        L.fmmfft_plan_sizes.argtypes = [i64, i32, i32, i32, i32, ctypes.POINTER(sz), ctypes.POINTER(sz)]
        L.fmmfft_plan_create.argtypes = [ctypes.POINTER(vp), i64, i32, i32, i32, i32, i32]
        L.fmmfft_plan_destroy.argtypes = [vp]
        L.fmmfft_execute.argtypes = [vp, vp, vp, i32, vp]
        L.fmmfft_execute_batched.argtypes = [vp, i64, vp, vp, i32, vp]
        L.fmmfft_workspace_bytes.argtypes = [vp, i64]
        L.fmmfft_workspace_bytes.restype = sz
        L.fmmfft_execute_ws.argtypes = [vp, i64, vp, vp, i32, vp, sz, vp]
        L.fmmfft_execute_static.argtypes = [vp, vp, vp, i32, vp]
        L.fmmfft_plan_get.argtypes = [vp] + [ctypes.c_void_p] * 6
        L.fmmfft_read_table.argtypes = [vp, i32, vp, sz]
        L.fmmfft_read_stage.argtypes = [vp, i32, vp, sz]
        L.fmmfft_profile_begin.argtypes = [vp, i32]
        L.fmmfft_profile_end.argtypes = [vp, ctypes.POINTER(ctypes.c_double), i32, ctypes.POINTER(i64)]
        L.fmmfft_server_create.argtypes = [ctypes.POINTER(vp), sz, i32]
        L.fmmfft_server_submit.argtypes = [vp, i64, i32, i32, i32, i32, i32, vp, vp, vp,
                                           ctypes.POINTER(ctypes.c_uint64)]
        L.fmmfft_server_wait.argtypes = [vp, ctypes.c_uint64, ctypes.POINTER(ctypes.c_double)]
        L.fmmfft_server_drain.argtypes = [vp]
        L.fmmfft_server_stats_get.argtypes = [vp, ctypes.POINTER(ServerStats)]
        L.fmmfft_server_reset_stats.argtypes = [vp]
        L.fmmfft_server_destroy.argtypes = [vp]
        L.fmmfft_strerror.argtypes = [ctypes.c_int]
        L.fmmfft_strerror.restype = ctypes.c_char_p
        L.fmmfft_last_hip_error.restype = ctypes.c_int
        for name in EXPORTS:
            if name not in ("fmmfft_workspace_bytes", "fmmfft_strerror", "fmmfft_last_hip_error"):
                getattr(L, name).restype = ctypes.c_int
        _lib = L
    return _lib


def _check(rc: int, what: str):
    if rc != OK:
        raise FmmFftError(rc, what)


def plan_sizes(n: int, p_blocks: int, terms: int, leaf: int, dtype: int = C128):
    """(table_bytes, work_bytes per signal) -- the reference's getwlen2/getvlen."""
    t, w = ctypes.c_size_t(), ctypes.c_size_t()
    _check(lib().fmmfft_plan_sizes(n, p_blocks, terms, leaf, dtype, ctypes.byref(t), ctypes.byref(w)),
           "fmmfft_plan_sizes")
    return t.value, w.value


def _torch():
    import torch

    return torch


def _stream_ptr(stream):
    torch = _torch()
    if stream is None:
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(stream.cuda_stream)


class Plan:
    """A device-resident FMM-FFT plan for length N = P * M (mfti)."""

    def __init__(self, n: int, p_blocks: int, terms: int, leaf: int, dtype: int = C128, max_batch: int = 1):
        h = ctypes.c_void_p()
        _check(lib().fmmfft_plan_create(ctypes.byref(h), n, p_blocks, terms, leaf, dtype, max_batch),
               f"fmmfft_plan_create(N={n}, P={p_blocks}, terms={terms}, leaf={leaf}, dtype={dtype})")
        self._h = h
        self.n, self.p_blocks, self.terms, self.leaf, self.dtype = n, p_blocks, terms, leaf, dtype
        self.max_batch = max(1, max_batch)
        self.m = n // p_blocks
        depth = ctypes.c_int32()
        lib().fmmfft_plan_get(h, None, None, None, None, None, ctypes.byref(depth))
        self.depth = depth.value

    @property
    def handle(self):
        return self._h

    @property
    def torch_dtype(self):
        torch = _torch()
        return torch.complex128 if self.dtype == C128 else torch.complex64

    def close(self):
        if getattr(self, "_h", None):
            lib().fmmfft_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check_tensor(self, t, count):
        if not t.is_cuda:
            raise ValueError("FMM-FFT tensors must be on the GPU")
        if t.dtype != self.torch_dtype:
            raise ValueError(f"expected {self.torch_dtype}, got {t.dtype}")
        if not t.is_contiguous() or t.numel() != count * self.n:
            raise ValueError(f"expected a contiguous tensor of {count}x{self.n} elements")

    def execute(self, x, out=None, direction: int = FORWARD, stream=None):
        """X = FFT(x) over the last N elements of each signal (out-of-place or in place)."""
        count = x.numel() // self.n if self.n else 0
        self._check_tensor(x, count)
        if out is None:
            out = _torch().empty_like(x)
        self._check_tensor(out, count)
        _check(lib().fmmfft_execute_batched(self._h, count, ctypes.c_void_p(x.data_ptr()),
                                            ctypes.c_void_p(out.data_ptr()), direction, _stream_ptr(stream)),
               "fmmfft_execute_batched")
        return out

    def execute_ws(self, x, out, work, direction: int = FORWARD, stream=None):
        """Execute with a caller-owned workspace tensor (uint8, workspace_bytes(count))."""
        count = x.numel() // self.n
        self._check_tensor(x, count)
        self._check_tensor(out, count)
        _check(lib().fmmfft_execute_ws(self._h, count, ctypes.c_void_p(x.data_ptr()),
                                       ctypes.c_void_p(out.data_ptr()), direction,
                                       ctypes.c_void_p(work.data_ptr()), work.numel() * work.element_size(),
                                       _stream_ptr(stream)),
               "fmmfft_execute_ws")
        return out

    def execute_static(self, x, out=None, direction: int = FORWARD, stream=None):
        """One transform on the library's static device scratch
        (fmmfft_execute_static: allocates nothing; calls that overlap on one
        device race)."""
        self._check_tensor(x, 1)
        if out is None:
            out = _torch().empty_like(x)
        self._check_tensor(out, 1)
        _check(lib().fmmfft_execute_static(self._h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                                           direction, _stream_ptr(stream)), "fmmfft_execute_static")
        return out

    def workspace_bytes(self, count: int = 1) -> int:
        return lib().fmmfft_workspace_bytes(self._h, count)

    STAGES = ("s2m+sums", "upward", "downward", "leaf", "combine", "fft1", "fft2")

    def profile_begin(self, max_calls: int):
        """Record HIP events around every stage of the next max_calls executes."""
        _check(lib().fmmfft_profile_begin(self._h, max_calls), "fmmfft_profile_begin")

    def profile_end(self):
        """(per-stage summed ms dict, number of recorded executes)."""
        ms = (ctypes.c_double * len(self.STAGES))()
        calls = ctypes.c_int64()
        _check(lib().fmmfft_profile_end(self._h, ms, len(self.STAGES), ctypes.byref(calls)), "fmmfft_profile_end")
        return dict(zip(self.STAGES, list(ms))), calls.value

    def read_table(self, which: int, count: int):
        import numpy as np

        a = np.zeros(count, dtype=np.float64)
        _check(lib().fmmfft_read_table(self._h, which, a.ctypes.data_as(ctypes.c_void_p), a.nbytes),
               "fmmfft_read_table")
        return a

    def read_stage(self, which: int, count: int, dtype):
        import numpy as np

        a = np.zeros(count, dtype=dtype)
        _check(lib().fmmfft_read_stage(self._h, which, a.ctypes.data_as(ctypes.c_void_p), a.nbytes),
               "fmmfft_read_stage")
        return a


def mfti(lq: int, p: int, terms: int, b: int, dtype: int = C128, max_batch: int = 1) -> Plan:
    """Plan for N = lq * p (the reference's mfti(lq, p, terms, b, w, v))."""
    return Plan(lq * p, p, terms, b, dtype, max_batch)


def mft(plan: Plan, qr, direction: int = FORWARD, stream=None):
    """In-place transform of all P blocks of ``qr`` (the reference's mft)."""
    return plan.execute(qr, qr, direction, stream)


class Server:
    """Request server (fmmfft_server_*): continuous batching of same-shape
    requests under an HBM budget with cached plans and hipGraph replay.
    Requests pass GPU tensors that must stay alive until wait() returns."""

    def __init__(self, hbm_budget_bytes: int, max_batch: int = 64):
        h = ctypes.c_void_p()
        _check(lib().fmmfft_server_create(ctypes.byref(h), int(hbm_budget_bytes), int(max_batch)),
               "fmmfft_server_create")
        self._h = h

    def submit(self, x, out, n: int, p_blocks: int, terms: int, leaf: int, dtype: int,
               direction: int = FORWARD, stream=None) -> int:
        """Queue FFT(x) -> out; x is read after the work queued so far on
        ``stream`` (torch's current stream by default)."""
        t = ctypes.c_uint64()
        _check(lib().fmmfft_server_submit(self._h, n, p_blocks, terms, leaf, dtype, direction,
                                          ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                                          _stream_ptr(stream), ctypes.byref(t)), "fmmfft_server_submit")
        return t.value

    def wait(self, ticket: int) -> float:
        ms = ctypes.c_double()
        _check(lib().fmmfft_server_wait(self._h, ticket, ctypes.byref(ms)), "fmmfft_server_wait")
        return ms.value

    def drain(self):
        _check(lib().fmmfft_server_drain(self._h), "fmmfft_server_drain")

    def stats(self) -> dict:
        st = ServerStats()
        _check(lib().fmmfft_server_stats_get(self._h, ctypes.byref(st)), "fmmfft_server_stats_get")
        return {n: getattr(st, n) for n, _ in ServerStats._fields_}

    def reset_stats(self):
        _check(lib().fmmfft_server_reset_stats(self._h), "fmmfft_server_reset_stats")

    def close(self):
        if getattr(self, "_h", None):
            lib().fmmfft_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
