"""ctypes binding of libssr_hip.so (the C ABI declared in include/ssr_hip.h).

The library is required: importing anything that computes raises
``SSRLibraryError`` when it is missing or no GPU is visible -- there is no CPU
fallback.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libssr_hip.so")

_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_int = ctypes.c_int
_dbl = ctypes.c_double
_sz = ctypes.c_size_t


class SSRLibraryError(RuntimeError):
    pass


class SSRError(RuntimeError):
    pass


# name -> argtypes (every function returns int status unless listed in _RESTYPE)
SIGNATURES = {
    "ssr_version": [],
    "ssr_last_error": [],
    "ssr_device_count": [_vp],
    "ssr_set_stream": [_vp],
    "ssr_get_stream": [],
    "ssr_set_thread_stream": [_vp],
    "ssr_stream_create": [_vp],
    "ssr_stream_destroy": [_vp],
    "ssr_dmalloc": [_vp, _sz],
    "ssr_dfree": [_vp],
    "ssr_htod": [_vp, _vp, _sz],
    "ssr_dtoh": [_vp, _vp, _sz],
    "ssr_dtod": [_vp, _vp, _sz],
    "ssr_dzero": [_vp, _sz],
    "ssr_synchronize": [],
    "ssr_event_create": [_vp],
    "ssr_event_destroy": [_vp],
    "ssr_event_record": [_vp],
    "ssr_event_elapsed_ms": [_vp, _vp, _vp],
    "ssr_constitutive_problem_3D_SDS": [_i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "ssr_constitutive_problem_3D_S": [_i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "ssr_constitutive_problem_2D_SDS": [_i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "ssr_constitutive_problem_2D_S": [_i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "ssr_dev_constitutive_3D": [_i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "ssr_dev_constitutive_2D": [_i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "ssr_dev_constitutive_mat": [_int, _i64, _int, _vp, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "ssr_dev_reduction": [_i64, _int, _vp, _vp, _vp, _vp, _int, _int, _vp, _vp],
    "ssr_dev_potential": [_int, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "ssr_set_tuning": [_int, _int],
    "ssr_get_tuning": [_int, _vp],
}
_RESTYPE = {"ssr_last_error": ctypes.c_char_p, "ssr_get_stream": _vp}

_lib = None


def load(path: str = LIB_PATH):
    """Load the library (no GPU needed for loading)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise SSRLibraryError(
            f"{path} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    lib = ctypes.CDLL(path)
    for name, args in SIGNATURES.items():
        if not hasattr(lib, name):
            continue
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = _RESTYPE.get(name, _int)
    _lib = lib
    return lib


def extend_signatures(extra: dict, restypes: dict | None = None):
    """Register additional entry points (used by the other modules)."""
    SIGNATURES.update(extra)
    if restypes:
        _RESTYPE.update(restypes)
    if _lib is not None:
        for name, args in extra.items():
            if hasattr(_lib, name):
                f = getattr(_lib, name)
                f.argtypes = args
                f.restype = _RESTYPE.get(name, _int)


def call(name: str, *args):
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != 0:
        msg = lib.ssr_last_error().decode(errors="replace")
        raise SSRError(f"{name}: {msg} (status {rc})")
    return rc


_gpu_checked = False


def require_gpu():
    """Raise unless the library is loaded and a GPU is visible."""
    global _gpu_checked
    if _gpu_checked:
        return
    lib = load()
    c = ctypes.c_int(0)
    lib.ssr_device_count(ctypes.byref(c))
    if c.value < 1:
        raise SSRLibraryError("no HIP device visible: the SSR kernels need an MI355X (gfx950)")
    _gpu_checked = True


def ptr(a) -> ctypes.c_void_p:
    """Address of a numpy array or DeviceArray (None -> NULL)."""
    if a is None:
        return ctypes.c_void_p(0)
    if isinstance(a, DeviceArray) or hasattr(a, "ptr"):
        return ctypes.c_void_p(a.ptr)
    if isinstance(a, np.ndarray):
        return ctypes.c_void_p(a.ctypes.data)
    raise TypeError(type(a))


class DeviceArray:
    """A device allocation with a numpy-like shape/dtype (C order)."""

    def __init__(self, shape, dtype=np.float64, zero: bool = False):
        require_gpu()
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        p = ctypes.c_void_p(0)
        call("ssr_dmalloc", ctypes.byref(p), max(self.nbytes, 8))
        self.ptr = p.value
        if zero:
            self.zero()

    @classmethod
    def from_host(cls, a, dtype=None):
        a = np.ascontiguousarray(a, dtype=dtype)
        d = cls(a.shape, a.dtype)
        call("ssr_htod", ctypes.c_void_p(d.ptr), ctypes.c_void_p(a.ctypes.data), a.nbytes)
        return d

    def set(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes, (a.shape, self.shape)
        call("ssr_htod", ctypes.c_void_p(self.ptr), ctypes.c_void_p(a.ctypes.data), a.nbytes)

    def get(self):
        out = np.empty(self.shape, dtype=self.dtype)
        call("ssr_dtoh", ctypes.c_void_p(out.ctypes.data), ctypes.c_void_p(self.ptr), self.nbytes)
        return out

    def zero(self):
        call("ssr_dzero", ctypes.c_void_p(self.ptr), self.nbytes)

    def view(self, offset_elems: int, shape):
        """Non-owning view starting `offset_elems` elements in."""
        v = object.__new__(DeviceArray)
        v.shape = tuple(shape) if isinstance(shape, (tuple, list)) else (int(shape),)
        v.dtype = self.dtype
        v.nbytes = int(np.prod(v.shape, dtype=np.int64)) * self.dtype.itemsize
        v.ptr = self.ptr + offset_elems * self.dtype.itemsize
        v._owner = self
        return v

    def free(self):
        if getattr(self, "_owner", None) is None and getattr(self, "ptr", 0):
            if _lib is not None:
                _lib.ssr_dfree(ctypes.c_void_p(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


TUNE_CONSTITUTIVE_VARIANT = 0
TUNE_SPMV_GROUP = 1
TUNE_PERSIST_WAVES = 2
TUNE_AMG_GRAPH = 3
TUNE_AMG_BLOCK_SMOOTH = 4
TUNE_ASSEMBLY_FORM = 5
TUNE_SPMM_GROUP = 6
TUNE_SPMV_UNROLL = 7
TUNE_ELEMENT_KERNEL = 8
TUNE_AMG_COARSE_BLOCKS = 9
TUNE_AMG_DEVICE_SETUP = 10
TUNE_SPMV_STAGE = 11
TUNE_HALVING_SUMS = 12
TUNE_CHEB_PER_DOF = 13
TUNE_SPMV_MIRROR = 14
TUNE_CONSTITUTIVE_2D = 15
TUNE_NEWTON_GRAPH = 16
TUNE_KRYLOV_PREFETCH = 17
TUNE_BRC_PREFETCH = 18


def set_tuning(key: int, value: int):
    call("ssr_set_tuning", int(key), int(value))


def get_tuning(key: int) -> int:
    v = ctypes.c_int(0)
    call("ssr_get_tuning", int(key), ctypes.byref(v))
    return v.value


def synchronize():
    call("ssr_synchronize")


class EventTimer:
    """HIP events recorded on the library's stream (the stream every kernel
    of this library is launched on)."""

    def __init__(self):
        load()
        self._a = ctypes.c_void_p(0)
        self._b = ctypes.c_void_p(0)
        call("ssr_event_create", ctypes.byref(self._a))
        call("ssr_event_create", ctypes.byref(self._b))

    def start(self):
        call("ssr_event_record", self._a)

    def stop(self):
        call("ssr_event_record", self._b)

    def elapsed_ms(self) -> float:
        ms = ctypes.c_float(0)
        call("ssr_event_elapsed_ms", self._a, self._b, ctypes.byref(ms))
        return float(ms.value)

    def __del__(self):
        try:
            _lib.ssr_event_destroy(self._a)
            _lib.ssr_event_destroy(self._b)
        except Exception:
            pass


class ThreadStream:
    """Context manager: the calling thread's library work runs on a stream of
    its own (created on entry, synchronised and destroyed on exit)."""

    def __enter__(self):
        p = ctypes.c_void_p(0)
        call("ssr_stream_create", ctypes.byref(p))
        self.stream = p
        call("ssr_set_thread_stream", p)
        return self

    def __exit__(self, *exc):
        try:
            synchronize()
        finally:
            call("ssr_set_thread_stream", ctypes.c_void_p(0))
            call("ssr_stream_destroy", self.stream)
        return False
