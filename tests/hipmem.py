"""Device buffers for tests, through the HIP runtime that libb40c_hip.so
itself links (ctypes, no other framework): allocate, upload, download."""
import ctypes
import os
import subprocess

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        from back40computing_amd import _native
        _native.load()
        path = None
        out = subprocess.run(["ldd", _native.LIB_PATH], capture_output=True, text=True).stdout
        for line in out.splitlines():
            if "libamdhip64" in line and "=>" in line:
                path = line.split("=>")[1].split("(")[0].strip()
        _hip = ctypes.CDLL(path or "libamdhip64.so")
        vp = ctypes.c_void_p
        _hip.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
        _hip.hipFree.argtypes = [vp]
        _hip.hipMemcpy.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


class DeviceArray:
    """A device allocation holding a copy of a numpy array (or `count` items)."""

    def __init__(self, host=None, dtype=np.uint32, count=None):
        self.dtype = np.dtype(dtype if host is None else host.dtype)
        self.count = int(len(host) if host is not None else count)
        self.nbytes = max(1, self.count) * self.dtype.itemsize
        self.ptr = ctypes.c_void_p()
        assert hip().hipMalloc(ctypes.byref(self.ptr), self.nbytes) == 0
        if host is not None and self.count:
            h = np.ascontiguousarray(host)
            assert hip().hipMemcpy(self.ptr, h.ctypes.data, self.count * self.dtype.itemsize, 1) == 0

    def get(self, count=None):
        n = self.count if count is None else int(count)
        out = np.empty(n, dtype=self.dtype)
        if n:
            assert hip().hipMemcpy(out.ctypes.data, self.ptr, n * self.dtype.itemsize, 2) == 0
        return out

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
