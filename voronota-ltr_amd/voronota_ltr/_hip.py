"""Minimal ctypes binding of the HIP runtime for device buffers (plumbing for
the device-resident entry points; importing it does not touch the GPU)."""
from __future__ import annotations

import ctypes as C
import os

_hip = None


def _load():
    global _hip
    if _hip is None:
        for cand in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                _hip = C.CDLL(cand)
                break
            except OSError:
                continue
        if _hip is None:
            raise RuntimeError("libamdhip64.so not found")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
        _hip.hipSetDevice.argtypes = [C.c_int]
    return _hip


H2D, D2H, D2D = 1, 2, 3


class DeviceBuffer:
    def __init__(self, nbytes: int, device: int = 0):
        h = _load()
        if h.hipSetDevice(device) != 0:
            raise RuntimeError("hipSetDevice failed")
        self.ptr = C.c_void_p()
        self.nbytes = max(int(nbytes), 1)
        if h.hipMalloc(C.byref(self.ptr), self.nbytes) != 0:
            raise RuntimeError(f"hipMalloc({self.nbytes}) failed")

    def upload(self, array):
        h = _load()
        if array.nbytes > self.nbytes:
            raise ValueError("array larger than buffer")
        if h.hipMemcpy(self.ptr, array.ctypes.data_as(C.c_void_p), array.nbytes, H2D) != 0:
            raise RuntimeError("hipMemcpy H2D failed")

    def download(self, array):
        h = _load()
        if array.nbytes > self.nbytes:
            raise ValueError("array larger than buffer")
        if h.hipMemcpy(array.ctypes.data_as(C.c_void_p), self.ptr, array.nbytes, D2H) != 0:
            raise RuntimeError("hipMemcpy D2H failed")

    def free(self):
        if self.ptr:
            _load().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def copy_d2d(dst: DeviceBuffer, src: DeviceBuffer, nbytes: int):
    if nbytes > min(dst.nbytes, src.nbytes):
        raise ValueError("copy larger than the buffers")
    if _load().hipMemcpy(dst.ptr, src.ptr, nbytes, D2D) != 0:
        raise RuntimeError("hipMemcpy D2D failed")


def synchronize():
    if _load().hipDeviceSynchronize() != 0:
        raise RuntimeError("hipDeviceSynchronize failed")


def visible() -> bool:
    return os.path.exists("/dev/kfd")
