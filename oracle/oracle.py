"""ctypes front end of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, as the checker.  It wraps oracle/liborc.so (built from
oracle/fmmfft_oracle.c by oracle/Makefile), the C99 FP64 restatement of the
reference src/fmmfft.c plus an FP64 FFT with long-double twiddles.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liborc.so")
_lib = None

_i64 = ctypes.c_int64
_dp = ctypes.POINTER(ctypes.c_double)


class _Plan(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in ("N", "P", "M", "terms", "b", "s", "n")] + [
        (n, _dp) for n in ("cheb", "initch", "ops", "topflp", "evalm", "evalmh", "kv", "fo")
    ]


def build() -> str:
    """Compile liborc.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.POINTER(_Plan)
        L.orc_plan_init.argtypes = [P, _i64, _i64, _i64, _i64]
        L.orc_plan_init.restype = ctypes.c_int
        L.orc_plan_free.argtypes = [P]
        L.orc_tree_len.argtypes = [P]
        L.orc_tree_len.restype = _i64
        L.orc_level_off.argtypes = [P, ctypes.c_int]
        L.orc_level_off.restype = _i64
        for name in ("orc_section_sums",):
            getattr(L, name).argtypes = [P, _dp, _dp]
        L.orc_upward.argtypes = [P, _dp, _dp]
        L.orc_downward.argtypes = [P, _dp, _dp]
        L.orc_evaluate.argtypes = [P, _dp, _dp, _dp]
        L.orc_combine.argtypes = [P, _dp, _dp, _dp, _dp]
        L.orc_fmmfft.argtypes = [P, _dp, _dp, ctypes.c_int]
        L.orc_fft64.argtypes = [_i64, _dp, ctypes.c_int]
        L.orc_dft_direct.argtypes = [_i64, _dp, _dp, ctypes.c_int]
        L.orc_cot_direct.argtypes = [_i64, _i64, _i64, _dp, _i64, _dp]
        for name in ("orc_rel_l2", "orc_max_abs", "orc_max_rel_pct"):
            getattr(L, name).argtypes = [_i64, _dp, _dp]
            getattr(L, name).restype = ctypes.c_double
        L.orc_fill_uniform.argtypes = [_i64, ctypes.c_uint64, ctypes.c_double, ctypes.c_double, _dp]
        L.orc_fill_normal.argtypes = [_i64, ctypes.c_uint64, ctypes.c_double, _dp]
        L.orc_fill_ramp.argtypes = [_i64, _dp]
        L.orc_num_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    assert a.dtype in (np.float64, np.complex128) and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def _c(n) -> np.ndarray:
    return np.zeros(n, dtype=np.complex128)


class OraclePlan:
    """The reference's mfti: all FP64 operator tables for (N, P, terms, b)."""

    def __init__(self, N: int, P: int, terms: int, b: int):
        self._p = _Plan()
        rc = lib().orc_plan_init(ctypes.byref(self._p), N, P, terms, b)
        if rc != 0:
            raise ValueError(f"invalid oracle plan N={N} P={P} terms={terms} b={b}")
        self.N, self.P, self.M, self.terms, self.b = N, P, N // P, terms, b
        self.s, self.n = self._p.s, self._p.n

    def __del__(self):
        try:
            lib().orc_plan_free(ctypes.byref(self._p))
        except Exception:
            pass

    def _arr(self, name, shape):
        n = int(np.prod(shape))
        return np.ctypeslib.as_array(getattr(self._p, name), shape=(n,)).reshape(shape).copy()

    # operator tables -------------------------------------------------------
    @property
    def cheb(self):
        return self._arr("cheb", (self.terms,))

    @property
    def initch(self):
        return self._arr("initch", (self.b, self.terms))

    @property
    def ops(self):
        return self._arr("ops", (self.n - 2, 8, self.terms, self.terms))

    @property
    def topflp(self):
        return self._arr("topflp", (self.terms, self.terms))

    @property
    def evalm(self):
        return self._arr("evalm", (self.P - 1, self.b, self.terms))

    @property
    def evalmh(self):
        return self._arr("evalmh", (self.P // 2, self.terms))

    @property
    def kv(self):
        return self._arr("kv", (self.P - 1, 4 * self.b))

    @property
    def fo(self):
        a = self._arr("fo", (self.P, self.P - 1, 2))
        return a[..., 0] + 1j * a[..., 1]

    # stages ----------------------------------------------------------------
    def tree_len(self) -> int:
        return lib().orc_tree_len(ctypes.byref(self._p))

    def level_off(self, l: int) -> int:
        return lib().orc_level_off(ctypes.byref(self._p), l)

    def section_sums(self, x):
        x = np.ascontiguousarray(x, dtype=np.complex128)
        out = _c(self.P)
        lib().orc_section_sums(ctypes.byref(self._p), _ptr(x), _ptr(out))
        return out

    def upward(self, x):
        x = np.ascontiguousarray(x, dtype=np.complex128)
        phi = _c(self.tree_len())
        lib().orc_upward(ctypes.byref(self._p), _ptr(x), _ptr(phi))
        return phi

    def downward(self, phi):
        psi = _c(self.tree_len())
        lib().orc_downward(ctypes.byref(self._p), _ptr(np.ascontiguousarray(phi)), _ptr(psi))
        return psi

    def evaluate(self, x, psi):
        x = np.ascontiguousarray(x, dtype=np.complex128)
        out = _c(self.N)
        lib().orc_evaluate(ctypes.byref(self._p), _ptr(x), _ptr(np.ascontiguousarray(psi)), _ptr(out))
        return out

    def combine(self, x, psi_sec, sumsec):
        x = np.ascontiguousarray(x, dtype=np.complex128)
        y = _c(self.N)
        lib().orc_combine(ctypes.byref(self._p), _ptr(x), _ptr(np.ascontiguousarray(psi_sec)),
                          _ptr(np.ascontiguousarray(sumsec)), _ptr(y))
        return y

    def fmmfft(self, x, direction: int = 1):
        x = np.ascontiguousarray(x, dtype=np.complex128)
        X = _c(self.N)
        lib().orc_fmmfft(ctypes.byref(self._p), _ptr(x), _ptr(X), int(direction))
        return X


def fft64(x, sign: int = -1) -> np.ndarray:
    a = np.array(x, dtype=np.complex128, copy=True)
    lib().orc_fft64(a.size, _ptr(a), sign)
    return a


def dft_direct(x, sign: int = -1) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = _c(x.size)
    lib().orc_dft_direct(x.size, _ptr(x), _ptr(out), sign)
    return out


def cot_direct(x_sections, M: int, P: int, sc: int) -> np.ndarray:
    """Exact psi_sec column sc by the O(M^2) cot sum (x in natural order)."""
    x = np.ascontiguousarray(x_sections, dtype=np.complex128)
    out = _c(M)
    lib().orc_cot_direct(M, P, sc, x[sc:].ctypes.data_as(_dp), P, _ptr(out))
    return out


def rel_l2(x, y) -> float:
    x = np.ascontiguousarray(x, dtype=np.complex128)
    y = np.ascontiguousarray(y, dtype=np.complex128)
    return lib().orc_rel_l2(x.size, _ptr(x), _ptr(y))


def max_abs(x, y) -> float:
    x = np.ascontiguousarray(x, dtype=np.complex128)
    y = np.ascontiguousarray(y, dtype=np.complex128)
    return lib().orc_max_abs(x.size, _ptr(x), _ptr(y))


def max_rel_pct(x, y) -> float:
    x = np.ascontiguousarray(x, dtype=np.complex128)
    y = np.ascontiguousarray(y, dtype=np.complex128)
    return lib().orc_max_rel_pct(x.size, _ptr(x), _ptr(y))


def uniform(n: int, seed: int, lo=-1.0, hi=1.0) -> np.ndarray:
    out = _c(n)
    lib().orc_fill_uniform(n, seed, lo, hi, _ptr(out))
    return out


def normal(n: int, seed: int, sigma=1.0) -> np.ndarray:
    out = _c(n)
    lib().orc_fill_normal(n, seed, sigma, _ptr(out))
    return out


def ramp(n: int) -> np.ndarray:
    out = _c(n)
    lib().orc_fill_ramp(n, _ptr(out))
    return out


def num_threads() -> int:
    return lib().orc_num_threads()
