"""Single-pass scan and stream compaction on the device, and CSR
construction on the device (C ABI in include/b40c_scan.h and
include/b40c_graph.h).

    ScanEnactor.scan      <- scan::Enactor::Scan (b40c/scan/enactor.cuh:466-561)
    ScanEnactor.copy_if   <- copy_if::Enactor::CopyIf (b40c/copy_if/enactor.cuh:485-571)
    DeviceCsr.from_coo    <- CsrGraph::FromCoo (b40c/graph/csr_graph.cuh:108-152), on the GPU
    DeviceCsr.rmat        <- the R-MAT builder, sampled and assembled on the GPU

Arrays are device addresses (ints); the caller owns them.
"""
import ctypes

import numpy as np

from . import _native
from .bfs import BfsError, CsrGraph, _check

SELECT_LESS, SELECT_NOT_EQUAL, SELECT_NONNEGATIVE = 0, 1, 2


class ScanEnactor:
    def __init__(self, device=0):
        self._h = ctypes.c_void_p()
        _check(_native.load().b40c_scan_enactor_create(device, ctypes.byref(self._h)), "b40c_scan_enactor_create")

    def scan(self, d_in, d_out, n, exclusive=True):
        """Prefix sum of n uint32 (mod 2^32); returns the total."""
        total = ctypes.c_uint32()
        _check(_native.load().b40c_scan(self._h, d_in, d_out, int(n), int(exclusive), ctypes.byref(total)),
               "b40c_scan")
        return total.value

    def copy_if(self, d_in, d_out, n, op=SELECT_LESS, comparand=0, d_count=None):
        """Stable compaction of n uint32; returns the number kept."""
        k = ctypes.c_int64()
        _check(_native.load().b40c_copy_if(self._h, d_in, d_out, int(n), int(op), int(comparand) & 0xFFFFFFFF,
                                            ctypes.byref(k), d_count), "b40c_copy_if")
        return k.value

    def last_ms(self):
        ms = ctypes.c_double()
        _check(_native.load().b40c_scan_last_ms(self._h, ctypes.byref(ms)), "b40c_scan_last_ms")
        return ms.value

    def destroy(self):
        if self._h:
            _native.load().b40c_scan_enactor_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


NOALLOC_MAX_ITEMS = 1 << 26


def scan_noalloc(d_in, d_out, n, exclusive=True, d_total=None, stream=None, device=0):
    """b40c_scan_noalloc: the scan with no device allocation (scratch in a
    static device array, n <= 2^26), queued on `stream` without waiting."""
    _check(_native.load().b40c_scan_noalloc(device, d_in, d_out, int(n), int(exclusive), d_total, stream),
           "b40c_scan_noalloc")


class DeviceCsr:
    """A CSR graph in device memory, built on the device."""

    def __init__(self):
        self._s = _native.DeviceCsrStruct()

    @classmethod
    def _build(cls, fn, *args):
        g = cls()
        rc = fn(*args, ctypes.byref(g._s))
        if rc != 0:
            raise BfsError(1, "device CSR construction failed with code %d" % rc)
        return g

    @classmethod
    def from_coo(cls, nodes, src, dst, symmetrize=False, dedup=False, drop_self_loops=False, device=0):
        src = np.ascontiguousarray(src, dtype=np.int32)
        dst = np.ascontiguousarray(dst, dtype=np.int32)
        assert src.shape == dst.shape
        z = np.zeros(1, np.int32)
        return cls._build(_native.load().b40c_device_csr_from_coo, device, nodes, len(src),
                          _native.i32ptr(src if len(src) else z), _native.i32ptr(dst if len(dst) else z),
                          int(symmetrize), int(dedup), int(drop_self_loops))

    @classmethod
    def rmat(cls, scale, edge_factor, a=0.57, b=0.19, c=0.19, seed=0x5EED0022, scramble=True, device=0):
        return cls._build(_native.load().b40c_device_csr_rmat, device, scale, edge_factor, a, b, c, seed,
                          int(scramble))

    def transpose(self):
        t = DeviceCsr()
        rc = _native.load().b40c_device_csr_transpose(ctypes.byref(self._s), ctypes.byref(t._s))
        if rc != 0:
            raise BfsError(1, "device transpose failed with code %d" % rc)
        return t

    @property
    def nodes(self):
        return int(self._s.nodes)

    @property
    def edges(self):
        return int(self._s.edges)

    @property
    def symmetric(self):
        return bool(self._s.symmetric)

    @property
    def build_ms(self):
        return float(self._s.build_ms)

    def to_host(self):
        g = CsrGraph()
        rc = _native.load().b40c_device_csr_to_host(ctypes.byref(self._s), ctypes.byref(g._s))
        if rc != 0:
            raise BfsError(1, "device CSR copy failed with code %d" % rc)
        g._owned = True
        return g

    def free(self):
        if self._s.row_offsets:
            _native.load().b40c_device_csr_free(ctypes.byref(self._s))

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
