"""Single-pass scan and stable compaction (include/b40c_scan.h) against numpy
on the host: the reference's scan and copy-if drivers compare with a serial
CPU result the same way (test/scan/test_scan.cu, test/copy-if/
test_copy_if.cu:141-157).  Closed forms pin the long look-back chains."""
import numpy as np
import pytest

from back40computing_amd.primitives import (SELECT_LESS, SELECT_NONNEGATIVE, SELECT_NOT_EQUAL, ScanEnactor)
from hipmem import DeviceArray

pytestmark = pytest.mark.gpu

SIZES = [0, 1, 2, 63, 64, 65, 4095, 4096, 4097, 3 * 65536 + 17, 1_000_003]


def expected_scan(x, exclusive):
    incl = np.cumsum(x.astype(np.uint64)) & 0xFFFFFFFF
    return ((incl - x) & 0xFFFFFFFF if exclusive else incl).astype(np.uint32), int(incl[-1]) if len(x) else 0


@pytest.mark.parametrize("exclusive", [True, False], ids=["exclusive", "inclusive"])
def test_scan_matches_cumsum(gpu, exclusive):
    e = ScanEnactor()
    rng = np.random.default_rng(11)
    for n in SIZES:
        for hi in (2, 1 << 32):  # small values, and values that wrap the 32-bit sum
            x = rng.integers(0, hi, n, dtype=np.uint64).astype(np.uint32)
            d_in, d_out = DeviceArray(x), DeviceArray(count=n)
            total = e.scan(d_in.ptr, d_out.ptr, n, exclusive)
            want, want_total = expected_scan(x, exclusive)
            assert np.array_equal(d_out.get(), want), (n, hi)
            assert total == want_total, (n, hi)
    e.destroy()


def test_scan_in_place_and_unaligned(gpu):
    e = ScanEnactor()
    x = np.arange(70001, dtype=np.uint32) * 7 + 3
    d = DeviceArray(x)
    e.scan(d.ptr, d.ptr, len(x), True)
    assert np.array_equal(d.get(), expected_scan(x, True)[0])
    # 4-byte offsets: the scalar (unvectorised) path
    d2 = DeviceArray(np.concatenate([[0], x]).astype(np.uint32))
    out = DeviceArray(count=len(x) + 1)
    e.scan(d2.ptr.value + 4, out.ptr.value + 4, len(x), False)
    assert np.array_equal(out.get()[1:], expected_scan(x, False)[0])
    e.destroy()


def test_scan_long_chain_closed_form(gpu):
    # 2^26 ones: 16384 tiles chained; exclusive scan = 0, 1, 2, ...
    n = 1 << 26
    e = ScanEnactor()
    d = DeviceArray(np.ones(n, dtype=np.uint32))
    assert e.scan(d.ptr, d.ptr, n, True) == n
    got = d.get()
    assert np.array_equal(got, np.arange(n, dtype=np.uint32))
    print("scan of 2^26 uint32: %.3f ms (%.1f GB/s)" % (e.last_ms(), 8 * n / (e.last_ms() * 1e-3) / 1e9))
    e.destroy()


@pytest.mark.parametrize("op", [SELECT_LESS, SELECT_NOT_EQUAL, SELECT_NONNEGATIVE], ids=["less", "ne", "nonneg"])
def test_copy_if_is_stable_compaction(gpu, op):
    e = ScanEnactor()
    rng = np.random.default_rng(12 + op)
    for n in SIZES:
        x = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
        if op == SELECT_NOT_EQUAL:
            x[rng.random(n) < 0.3] = 77
        for comparand in ((0, 1 << 31, 0xFFFFFFFF) if op == SELECT_LESS else (77,)):
            if op == SELECT_LESS:
                mask = x < np.uint32(comparand)
            elif op == SELECT_NOT_EQUAL:
                mask = x != np.uint32(comparand)
            else:
                mask = x.view(np.int32) >= 0
            d_in, d_out, d_cnt = DeviceArray(x), DeviceArray(count=n), DeviceArray(count=1)
            k = e.copy_if(d_in.ptr, d_out.ptr, n, op, comparand, d_count=d_cnt.ptr)
            assert k == int(mask.sum()), (n, comparand)
            assert int(d_cnt.get()[0]) == k
            assert np.array_equal(d_out.get(k), x[mask]), (n, comparand)
    e.destroy()


def test_copy_if_deterministic_across_runs(gpu):
    e = ScanEnactor()
    x = np.random.default_rng(3).integers(0, 1 << 32, 5_000_000, dtype=np.uint64).astype(np.uint32)
    d_in = DeviceArray(x)
    outs = []
    for _ in range(3):
        d_out = DeviceArray(count=len(x))
        k = e.copy_if(d_in.ptr, d_out.ptr, len(x), SELECT_LESS, 1 << 31)
        outs.append(d_out.get(k))
    assert all(np.array_equal(outs[0], o) for o in outs[1:])
    e.destroy()
