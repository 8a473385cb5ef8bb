"""RadixSortNoAlloc (the sort entry point that allocates nothing) against RadixSort through
the C ABI: two calls in a row, the second with a smaller input, bit-identical to the
allocating sort and to a stable sort on the host; sizes past its fixed scratch refused."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _keys(n, key_max, seed):
    """(cell key, index) pairs in index order -- what calcHash hands the sort."""
    rng = np.random.default_rng(seed)
    kv = np.empty((n, 2), dtype=np.uint32)
    kv[:, 0] = rng.integers(0, key_max, n, dtype=np.uint32)
    kv[:, 1] = np.arange(n, dtype=np.uint32)
    return kv


def _sort_both(dem, kv, bits):
    L = dem.lib()
    D = dem.DeviceArray
    a0, a1 = D.from_host(kv), D(kv.shape, np.uint32)
    b0, b1 = D.from_host(kv), D(kv.shape, np.uint32)
    n = kv.shape[0]
    L.RadixSort(a0.p, a1.p, n, bits)
    assert L.dem_last_error() == 0
    rc = L.RadixSortNoAlloc(b0.p, b1.p, n, bits)
    assert rc == 0, dem.DemError(rc, "RadixSortNoAlloc")
    assert L.dem_device_sync() == 0
    return a0.host(), b0.host()


def test_no_alloc_sort_equals_radix_sort(dem, gpu_ok):
    cases = [(3_000_000, 1 << 20, 32, 1), (200_000, 50_000, 16, 2),   # second call smaller
             (5_000, 300, 32, 3)]
    for n, key_max, bits, seed in cases:
        say(f"[no-alloc sort] n={n} keys<{key_max} bits={bits}")
        kv = _keys(n, key_max, seed)
        ref, got = _sort_both(dem, kv, bits)
        assert np.array_equal(got, ref), n
        order = np.argsort(kv[:, 0], kind="stable")
        assert np.array_equal(got, kv[order]), n


def test_no_alloc_sort_refuses_sizes_past_its_scratch(dem, gpu_ok):
    """More than 2^23 elements: DEM_ERR_CAPACITY before any device work (the buffers are
    never touched), and the entry point keeps working afterwards."""
    L = dem.lib()
    D = dem.DeviceArray
    a, b = D(16, np.uint32), D(16, np.uint32)
    assert L.RadixSortNoAlloc(a.p, b.p, (1 << 23) + 1, 32) == dem.DEM_ERR_CAPACITY
    assert L.dem_last_error() == dem.DEM_ERR_CAPACITY
    assert L.dem_last_error() == 0
    assert L.RadixSortNoAlloc(a.p, a.p, 8, 32) == dem.DEM_ERR_INVALID
    assert L.dem_last_error() == dem.DEM_ERR_INVALID
    kv = _keys(1000, 64, 9)
    ref, got = _sort_both(dem, kv, 32)
    assert np.array_equal(got, ref)
