"""Minimal HIP runtime binding (device buffers, copies, sync) for the host
harness; the compute itself is inside librefeff_fms.so."""
from __future__ import annotations

import ctypes as C

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                _hip = C.CDLL(name)
                break
            except OSError:
                continue
        if _hip is None:
            raise RuntimeError("libamdhip64.so not found")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
        _hip.hipGetDeviceCount.argtypes = [C.POINTER(C.c_int)]
        _hip.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
    return _hip


H2D, D2H = 1, 2


def device_count() -> int:
    n = C.c_int(0)
    if hip().hipGetDeviceCount(C.byref(n)) != 0:
        return 0
    return n.value


def mem_info() -> tuple[int, int]:
    """(free, total) bytes of the current device."""
    free, total = C.c_size_t(0), C.c_size_t(0)
    if hip().hipMemGetInfo(C.byref(free), C.byref(total)) != 0:
        raise RuntimeError("hipMemGetInfo failed")
    return free.value, total.value


class DeviceArray:
    def __init__(self, host: np.ndarray):
        self.nbytes = host.nbytes
        self.ptr = C.c_void_p()
        if hip().hipMalloc(C.byref(self.ptr), max(1, self.nbytes)) != 0:
            hip().hipGetLastError()  # a failed allocation must not reach the next caller's error check
            raise MemoryError("hipMalloc failed")
        self.upload(host)

    @classmethod
    def empty(cls, nbytes: int):
        obj = cls.__new__(cls)
        obj.nbytes = nbytes
        obj.ptr = C.c_void_p()
        if hip().hipMalloc(C.byref(obj.ptr), max(1, nbytes)) != 0:
            hip().hipGetLastError()
            raise MemoryError("hipMalloc failed")
        return obj

    def upload(self, host: np.ndarray):
        host = np.ascontiguousarray(host)
        assert host.nbytes == self.nbytes
        if hip().hipMemcpy(self.ptr, host.ctypes.data, self.nbytes, H2D) != 0:
            raise RuntimeError("hipMemcpy H2D failed")

    def download(self, out: np.ndarray):
        assert out.nbytes == self.nbytes and out.flags.c_contiguous
        if hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, D2H) != 0:
            raise RuntimeError("hipMemcpy D2H failed")
        return out

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize():
    if hip().hipDeviceSynchronize() != 0:
        raise RuntimeError("hipDeviceSynchronize failed")
