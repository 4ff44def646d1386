"""CPU oracle for the MoFFT 1-D complex FFT path (ctypes wrapper).

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, and only as the
checker (or as the timed CPU restatement of the reference algorithm).  The
product (``mofft_amd``) never imports it.

Functions follow ``oracle/mofft_oracle.c``:

* :func:`dft` - FP64 oracle: direct double-double summation for N <= 4096,
  compensated double-double Stockham FFT above (tests/test_fft.c:20-26 gives
  the transform's definition; include/mofft.h:26 the sign convention).
* :func:`ref_execute` - restatement of the reference executor
  (src/mofft.c:1932-2045) in FP32 or FP64.
* :func:`omega` - exp(sign*2*pi*i*m/M) rounded once (src/mofft.c:1082-1140
  fills the same values).
* :func:`plan_bytes` - byte accounting of src/mofft.c:190-304.
* :class:`PortPlan` - the reference executor on one host core with its
  tables made at plan time (``oracle/mofft_cpu_port.c``), timed by bench.py's
  cpu_baseline.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "libmofft_oracle.so")
_lib = None


def build() -> None:
    """Compile the oracle with oracle/Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    lib = ctypes.CDLL(_LIB_PATH)
    dp = ctypes.POINTER(ctypes.c_double)
    fp = ctypes.POINTER(ctypes.c_float)
    ip = ctypes.POINTER(ctypes.c_int)
    sz = ctypes.c_size_t
    for name in ("oracle_dft_direct", "oracle_fft_dd", "oracle_dft"):
        fn = getattr(lib, name)
        fn.argtypes = [sz, sz, ctypes.c_int, dp, dp]
        fn.restype = ctypes.c_int
    lib.oracle_ref_execute_f32.argtypes = [sz, ctypes.c_int, ip, sz, fp, fp]
    lib.oracle_ref_execute_f32.restype = ctypes.c_int
    lib.oracle_ref_execute_f64.argtypes = [sz, ctypes.c_int, ip, sz, dp, dp]
    lib.oracle_ref_execute_f64.restype = ctypes.c_int
    lib.oracle_omega_f64.argtypes = [ctypes.c_uint64, ctypes.c_uint64,
                                     ctypes.c_int, dp]
    lib.oracle_omega_f32.argtypes = [ctypes.c_uint64, ctypes.c_uint64,
                                     ctypes.c_int, fp]
    lib.oracle_omega_dd.argtypes = [ctypes.c_uint64, ctypes.c_uint64,
                                    ctypes.c_int, dp, dp]
    lib.oracle_omega_many.argtypes = [ctypes.POINTER(ctypes.c_uint64), sz,
                                      ctypes.c_uint64, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_void_p]
    lib.oracle_plan_bytes.argtypes = [sz, ctypes.c_int, ip, sz,
                                      ctypes.c_uint64, ctypes.c_int,
                                      ctypes.c_int, sz,
                                      ctypes.POINTER(ctypes.c_uint64)]
    lib.oracle_plan_bytes.restype = ctypes.c_int
    _lib = lib
    return lib


def _as_c128(x: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(x, dtype=np.complex128)


def _run(fn_name: str, x: np.ndarray, sign: int) -> np.ndarray:
    lib = _load()
    x = _as_c128(x)
    n = x.shape[-1]
    batch = x.size // n if n else 0
    out = np.empty_like(x)
    dp = ctypes.POINTER(ctypes.c_double)
    rc = getattr(lib, fn_name)(n, batch, int(sign),
                               x.ctypes.data_as(dp), out.ctypes.data_as(dp))
    if rc != 0:
        raise ValueError(f"{fn_name} rejected n={n} sign={sign}")
    return out


def dft(x: np.ndarray, sign: int = -1) -> np.ndarray:
    """FP64 oracle DFT along the last axis (direct for N<=4096, else FFT)."""
    return _run("oracle_dft", x, sign)


def dft_direct(x: np.ndarray, sign: int = -1) -> np.ndarray:
    return _run("oracle_dft_direct", x, sign)


def fft_dd(x: np.ndarray, sign: int = -1) -> np.ndarray:
    return _run("oracle_fft_dd", x, sign)


def ref_execute(x: np.ndarray, radices, sign: int = -1,
                precision_bits: int = 32) -> np.ndarray:
    """Reference executor restated (one transform, last axis)."""
    lib = _load()
    radices = np.ascontiguousarray(radices, dtype=np.int32)
    ip = ctypes.POINTER(ctypes.c_int)
    if precision_bits == 32:
        x = np.ascontiguousarray(x, dtype=np.complex64)
        out = np.empty_like(x)
        fp = ctypes.POINTER(ctypes.c_float)
        rc = lib.oracle_ref_execute_f32(x.shape[-1], int(sign),
                                        radices.ctypes.data_as(ip),
                                        radices.size, x.ctypes.data_as(fp),
                                        out.ctypes.data_as(fp))
    else:
        x = np.ascontiguousarray(x, dtype=np.complex128)
        out = np.empty_like(x)
        dp = ctypes.POINTER(ctypes.c_double)
        rc = lib.oracle_ref_execute_f64(x.shape[-1], int(sign),
                                        radices.ctypes.data_as(ip),
                                        radices.size, x.ctypes.data_as(dp),
                                        out.ctypes.data_as(dp))
    if rc != 0:
        raise ValueError("ref_execute rejected the plan")
    return out


def omega(m: int, big_m: int, sign: int = -1, precision_bits: int = 64):
    """exp(sign*2*pi*i*m/M) rounded once to the working precision."""
    lib = _load()
    if precision_bits == 32:
        buf = (ctypes.c_float * 2)()
        lib.oracle_omega_f32(m, big_m, sign, buf)
        return complex(np.float32(buf[0]), np.float32(buf[1]))
    buf = (ctypes.c_double * 2)()
    lib.oracle_omega_f64(m, big_m, sign, buf)
    return complex(buf[0], buf[1])


def omega_table(big_m: int, sign: int = -1, precision_bits: int = 64,
                indices=None) -> np.ndarray:
    """Vector of omega values for m in `indices` (default range(M))."""
    lib = _load()
    if indices is None:
        indices = np.arange(big_m, dtype=np.uint64)
    m = np.ascontiguousarray(np.asarray(indices, dtype=np.uint64))
    out = np.empty(m.size, dtype=np.complex64 if precision_bits == 32 else np.complex128)
    lib.oracle_omega_many(m.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), m.size,
                          big_m, sign, precision_bits, out.ctypes.data_as(ctypes.c_void_p))
    return out


_port = None


def _load_port():
    global _port
    if _port is not None:
        return _port
    path = os.path.join(_HERE, "build", "libmofft_cpu_port.so")
    if not os.path.exists(path):
        build()
    lib = ctypes.CDLL(path)
    for suf in ("f32", "f64"):
        c = getattr(lib, f"port_create_{suf}")
        c.argtypes = [ctypes.c_size_t, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                      ctypes.c_size_t]
        c.restype = ctypes.c_void_p
        e = getattr(lib, f"port_execute_{suf}")
        e.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        e.restype = ctypes.c_int
        getattr(lib, f"port_destroy_{suf}").argtypes = [ctypes.c_void_p]
    _port = lib
    return lib


class PortPlan:
    """The reference executor on one host core with its tables made at plan
    time (oracle/mofft_cpu_port.c): bench.py's timed CPU baseline."""

    def __init__(self, n: int, radices, sign: int = -1, precision_bits: int = 64):
        self._lib = _load_port()
        self._suf = "f32" if precision_bits == 32 else "f64"
        self.dtype = np.complex64 if precision_bits == 32 else np.complex128
        self.n = int(n)
        r = np.ascontiguousarray(radices, dtype=np.int32)
        self._h = getattr(self._lib, f"port_create_{self._suf}")(
            self.n, int(sign), r.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), r.size)
        if not self._h:
            raise ValueError(f"port rejected n={n} radices={list(radices)}")

    def execute(self, x: np.ndarray, out: np.ndarray) -> None:
        assert x.dtype == self.dtype and out.dtype == self.dtype
        assert x.size == self.n and out.size == self.n
        rc = getattr(self._lib, f"port_execute_{self._suf}")(
            self._h, x.ctypes.data, out.ctypes.data)
        if rc:
            raise ValueError("port execute failed")

    def destroy(self) -> None:
        if self._h:
            getattr(self._lib, f"port_destroy_{self._suf}")(self._h)
            self._h = None

    def __del__(self):
        self.destroy()


def plan_bytes(length: int, precision_bits: int, radices, blocked_mask=0,
               direct_input=0, section=False, cache_line_bytes=128):
    """(logical_read, logical_write, transferred_read, transferred_write,
    peak_working_set) or None for invalid input."""
    lib = _load()
    radices = np.ascontiguousarray(radices, dtype=np.int32)
    out = (ctypes.c_uint64 * 5)()
    rc = lib.oracle_plan_bytes(length, precision_bits,
                               radices.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                               radices.size, blocked_mask, direct_input,
                               int(bool(section)), cache_line_bytes, out)
    return None if rc else tuple(int(v) for v in out)


def rel_l2(actual: np.ndarray, expected: np.ndarray) -> float:
    """Relative L2 error ||a - e|| / ||e|| computed in float64."""
    a = np.asarray(actual, dtype=np.complex128)
    e = np.asarray(expected, dtype=np.complex128)
    denom = np.linalg.norm(e.ravel())
    return float(np.linalg.norm((a - e).ravel()) / (denom if denom else 1.0))
