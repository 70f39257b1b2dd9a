"""Minimal HIP runtime plumbing over ctypes: device buffers, copies, events.

It binds the same libamdhip64.so.7 that libxyz_hip.so links, so buffers made
here are valid in the library's kernels.  (PyTorch wheels bundle an older HIP/
HSA runtime with the same sonames; loading both into one process leaves one of
them without a device, so this package does not use torch for device memory.)
"""
import ctypes
import math

import numpy as np

_rt = None


def rt():
    global _rt
    if _rt is None:
        from . import capi
        capi.lib()  # loads the runtime the library links
        R = ctypes.CDLL("libamdhip64.so.7")
        vp = ctypes.c_void_p
        R.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
        R.hipFree.argtypes = [vp]
        R.hipMemcpy.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
        R.hipMemset.argtypes = [vp, ctypes.c_int, ctypes.c_size_t]
        R.hipDeviceSynchronize.argtypes = []
        R.hipGetDeviceCount.argtypes = [ctypes.POINTER(ctypes.c_int)]
        R.hipEventCreate.argtypes = [ctypes.POINTER(vp)]
        R.hipEventDestroy.argtypes = [vp]
        R.hipEventRecord.argtypes = [vp, vp]
        R.hipEventSynchronize.argtypes = [vp]
        R.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), vp, vp]
        R.hipGetErrorString.restype = ctypes.c_char_p
        R.hipGetErrorString.argtypes = [ctypes.c_int]
        R.hipStreamCreate.argtypes = [ctypes.POINTER(vp)]
        R.hipStreamDestroy.argtypes = [vp]
        R.hipStreamSynchronize.argtypes = [vp]
        R.hipGetDeviceProperties = None
        _rt = R
    return _rt


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what}: {rt().hipGetErrorString(rc).decode()}")


def device_count():
    n = ctypes.c_int(0)
    rc = rt().hipGetDeviceCount(ctypes.byref(n))
    return n.value if rc == 0 else 0


def synchronize():
    _check(rt().hipDeviceSynchronize(), "hipDeviceSynchronize")


H2D, D2H, D2D = 1, 2, 3


class DeviceArray:
    """A contiguous device buffer with a numpy-like shape and dtype."""

    def __init__(self, shape, dtype, _ptr=None, _owner=None):
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(math.prod(self.shape)) * self.dtype.itemsize
        self._owner = _owner
        if _ptr is not None:
            self._p = _ptr
        else:
            p = ctypes.c_void_p()
            _check(rt().hipMalloc(ctypes.byref(p), max(self.nbytes, 1)), "hipMalloc")
            self._p = p.value

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        if a.nbytes:
            _check(rt().hipMemcpy(ctypes.c_void_p(d._p), a.ctypes.data_as(ctypes.c_void_p), a.nbytes, H2D),
                   "hipMemcpy H2D")
        return d

    @classmethod
    def zeros(cls, shape, dtype):
        d = cls(shape, dtype)
        d.zero()
        return d

    def zero(self):
        if self.nbytes:
            _check(rt().hipMemset(ctypes.c_void_p(self._p), 0, self.nbytes), "hipMemset")

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            synchronize()
            _check(rt().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(self._p),
                                  self.nbytes, D2H), "hipMemcpy D2H")
        return out

    def copy_from(self, a):
        a = np.ascontiguousarray(a, self.dtype)
        assert a.nbytes == self.nbytes
        _check(rt().hipMemcpy(ctypes.c_void_p(self._p), a.ctypes.data_as(ctypes.c_void_p), a.nbytes, H2D),
               "hipMemcpy H2D")

    def data_ptr(self):
        return self._p

    @property
    def ptr(self):
        return ctypes.c_void_p(self._p)

    def rows(self, start, stop):
        """View of rows [start, stop) along the first axis (no copy)."""
        row = self.nbytes // max(self.shape[0], 1)
        return DeviceArray((stop - start,) + self.shape[1:], self.dtype,
                           _ptr=self._p + start * row, _owner=self)

    def __len__(self):
        return self.shape[0]

    def __del__(self):
        if self._owner is None and getattr(self, "_p", None):
            try:
                rt().hipFree(ctypes.c_void_p(self._p))
            except Exception:
                pass
            self._p = None


class Event:
    def __init__(self):
        e = ctypes.c_void_p()
        _check(rt().hipEventCreate(ctypes.byref(e)), "hipEventCreate")
        self.e = e

    def record(self, stream=None):
        _check(rt().hipEventRecord(self.e, stream), "hipEventRecord")

    def elapsed_ms(self, end):
        _check(rt().hipEventSynchronize(end.e), "hipEventSynchronize")
        ms = ctypes.c_float()
        _check(rt().hipEventElapsedTime(ctypes.byref(ms), self.e, end.e), "hipEventElapsedTime")
        return ms.value

    def __del__(self):
        if getattr(self, "e", None):
            try:
                rt().hipEventDestroy(self.e)
            except Exception:
                pass
            self.e = None
