"""Device buffers, streams and synchronisation through the HIP runtime that
libsbewh_hip.so itself links (libamdhip64.so.7 of /opt/rocm), by ctypes:
what a Python caller of the device-buffer entry points (sbe_rhs_device) needs
to own device memory.

torch bundles a HIP runtime of its own; two runtimes in one process do not
share a device context, so device pointers handed to the library come from
the library's own runtime.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

H2D, D2H, D2D = 1, 2, 3

_hip = None


def hip():
    global _hip
    if _hip is None:
        from ._lib import load
        load()   # the library's dependency is then the process's libamdhip64.so.7
        L = C.CDLL("libamdhip64.so.7")
        L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        L.hipFree.argtypes = [C.c_void_p]
        L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        L.hipDeviceSynchronize.argtypes = []
        L.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        L.hipStreamDestroy.argtypes = [C.c_void_p]
        L.hipStreamSynchronize.argtypes = [C.c_void_p]
        _hip = L
    return _hip


def check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with hipError {rc}")


class DeviceBuffer:
    """`nbytes` of device memory; upload / download of numpy arrays."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        check(hip().hipMalloc(C.byref(p), self.nbytes), "hipMalloc")
        self.ptr = p.value

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        check(hip().hipMemcpy(C.c_void_p(self.ptr), a.ctypes.data_as(C.c_void_p), a.nbytes, H2D), "hipMemcpy H2D")

    def copy_from(self, other: "DeviceBuffer", nbytes: int):
        """device-to-device copy of nbytes"""
        check(hip().hipMemcpy(C.c_void_p(self.ptr), C.c_void_p(other.ptr), int(nbytes), D2D), "hipMemcpy D2D")

    def download(self, dtype, count: int) -> np.ndarray:
        out = np.empty(count, dtype)
        assert out.nbytes <= self.nbytes
        check(hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), out.nbytes, D2H), "hipMemcpy D2H")
        return out

    def free(self):
        if self.ptr:
            check(hip().hipFree(C.c_void_p(self.ptr)), "hipFree")
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        s = C.c_void_p()
        check(hip().hipStreamCreate(C.byref(s)), "hipStreamCreate")
        self.handle = s.value

    def synchronize(self):
        check(hip().hipStreamSynchronize(C.c_void_p(self.handle)), "hipStreamSynchronize")

    def destroy(self):
        if self.handle:
            check(hip().hipStreamDestroy(C.c_void_p(self.handle)), "hipStreamDestroy")
            self.handle = None


def synchronize():
    check(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
