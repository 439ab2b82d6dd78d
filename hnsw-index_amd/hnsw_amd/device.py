"""Device buffers and streams on the HIP runtime that libhnsw_amd.so runs on.

The *_device entry points of the C ABI take raw device pointers and a HIP
stream.  These helpers allocate and copy through the library itself, so the
pointers and streams always belong to the runtime that launches the
kernels; a Python framework that bundles its own HIP runtime keeps its own
streams, which must not be passed to the library.
"""
import ctypes

import numpy as np

from ._lib import lib


def _check(st):
    if st:
        from . import _check as check
        check(st)


class Buffer:
    """One device allocation of `nbytes` bytes (freed with the object)."""

    def __init__(self, nbytes, device=0):
        self.nbytes = int(nbytes)
        self.device = device
        p = ctypes.c_void_p()
        _check(lib.hh_device_alloc(device, self.nbytes, ctypes.byref(p)))
        self._p = p

    @classmethod
    def from_array(cls, a, device=0, stream=None):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes, device)
        b.upload(a, stream)
        return b

    @property
    def ptr(self):
        return self._p

    def upload(self, a, stream=None):
        a = np.ascontiguousarray(a)
        if a.nbytes > self.nbytes:
            raise ValueError("array larger than the buffer")
        _check(lib.hh_memcpy(self._p, a.ctypes.data, a.nbytes, 0,
                             stream.handle if stream else None))

    def download(self, shape, dtype, stream=None):
        out = np.empty(shape, dtype)
        if out.nbytes > self.nbytes:
            raise ValueError("requested array larger than the buffer")
        _check(lib.hh_memcpy(out.ctypes.data, self._p, out.nbytes, 1,
                             stream.handle if stream else None))
        if stream:
            stream.synchronize()
        return out

    def zero(self):
        _check(lib.hh_memset(self._p, 0, self.nbytes))

    def free(self):
        if getattr(self, "_p", None) and self._p.value:
            lib.hh_device_free(self._p)
            self._p = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    """A non-blocking HIP stream of the library's runtime."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        _check(lib.hh_stream_create(device, ctypes.byref(h)))
        self.handle = h

    def synchronize(self):
        _check(lib.hh_stream_synchronize(self.handle))

    def close(self):
        if getattr(self, "handle", None) and self.handle.value:
            lib.hh_stream_synchronize(self.handle)
            lib.hh_stream_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def synchronize(device=0):
    """Waits for all work the library queued on `device`."""
    _check(lib.hh_device_synchronize(device))
