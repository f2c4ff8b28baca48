"""The allocation-free scan (b40c_scan_noalloc) against the allocating one
(b40c_scan) through the C ABI: two calls in a row on one GPU, the second
smaller, so leftovers of the first call's scratch would show."""
import ctypes

import numpy as np
import pytest

from back40computing_amd.primitives import NOALLOC_MAX_ITEMS, ScanEnactor, scan_noalloc
from hipmem import DeviceArray, hip

pytestmark = pytest.mark.gpu


def run_both(e, x, exclusive):
    n = len(x)
    d_in = DeviceArray(x)
    d_ref, d_new = DeviceArray(count=n), DeviceArray(count=n)
    d_total = DeviceArray(np.array([0xDEADBEEF], dtype=np.uint32))
    ref_total = e.scan(d_in.ptr, d_ref.ptr, n, exclusive)
    scan_noalloc(d_in.ptr, d_new.ptr, n, exclusive, d_total=d_total.ptr)
    assert hip().hipDeviceSynchronize() == 0
    return d_ref.get(), d_new.get(), ref_total, int(d_total.get()[0])


@pytest.mark.parametrize("exclusive", [True, False], ids=["exclusive", "inclusive"])
def test_noalloc_scan_matches_allocating_scan(gpu, exclusive):
    e = ScanEnactor()
    rng = np.random.default_rng(0x5CA7)
    # a long chain (4097 tiles), then a short ragged one: stale status words
    # of the first call would feed the second's look-back
    for n in [4096 * 4096 + 77, 4096 * 3 + 5, 1]:
        x = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
        ref, got, ref_total, got_total = run_both(e, x, exclusive)
        assert np.array_equal(got, ref), n
        assert got_total == ref_total, n
        incl = np.cumsum(x.astype(np.uint64)) & 0xFFFFFFFF
        assert got_total == int(incl[-1]), n
    e.destroy()


def test_noalloc_scan_limits(gpu):
    from back40computing_amd._native import load
    lib = load()
    d = DeviceArray(count=1)
    # past the static scratch: rejected before any launch
    assert lib.b40c_scan_noalloc(0, d.ptr, d.ptr, NOALLOC_MAX_ITEMS + 1, 1, None, None) == 1
    # the largest accepted size, all ones: exclusive scan = 0, 1, 2, ...
    n = NOALLOC_MAX_ITEMS
    x = DeviceArray(np.ones(n, dtype=np.uint32))
    tot = DeviceArray(count=1)
    scan_noalloc(x.ptr, x.ptr, n, True, d_total=tot.ptr)
    assert hip().hipDeviceSynchronize() == 0
    assert int(tot.get()[0]) == n
    assert np.array_equal(x.get(), np.arange(n, dtype=np.uint32))
    # empty input: total cleared, nothing launched on the arrays
    scan_noalloc(None, None, 0, True, d_total=tot.ptr)
    assert hip().hipDeviceSynchronize() == 0
    assert int(tot.get()[0]) == 0
