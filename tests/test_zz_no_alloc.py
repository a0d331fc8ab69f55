"""sfp_find_event_index_noalloc (no counterpart in the reference; for callers
that cannot allocate device memory) against sfp_find_event_index, the entry
point it mirrors (cpu.py:193-216): the same ascending indices and count, bit
for bit, on two calls in a row -- the second with fewer flags, so that it
runs on scratch still holding the first call's block counts -- and sizes
beyond its stated maximum are refused before any GPU work."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _flags(n, p, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    return (rng.random(n) < p).astype(np.uint8)


def _with_alloc(L, f):
    n = f.size
    df = L.DeviceArray.of(f)
    out = L.DeviceArray(8 * max(n, 1))
    cnt = C.c_int64(-1)
    L.call("sfp_find_event_index", df.ptr, n, out.ptr, C.byref(cnt), None)
    idx = np.empty(cnt.value, np.int64)
    if cnt.value:
        out.download(idx)
    return idx


def _no_alloc(L, f, out, cnt_dev):
    """Inputs and outputs preallocated by the caller; the call itself
    allocates nothing."""
    n = f.size
    df = L.DeviceArray.of(f)
    L.call("sfp_find_event_index_noalloc", df.ptr, n, out.ptr, cnt_dev.ptr, None)
    L.call("sfp_synchronize")
    cnt = np.empty(1, np.int64)
    cnt_dev.download(cnt)
    idx = np.empty(int(cnt[0]), np.int64)
    if cnt[0]:
        out.download(idx)
    return idx


def test_noalloc_matches_find_event_index_twice():
    from sfparticles_amd import _lib as L
    big = 3_000_000 + 517        # 2931 blocks of 1024 flags, the last one partial
    small = 70_001               # 69 blocks: the scratch keeps the first call's counts beyond
    out = L.DeviceArray(8 * big)
    cnt_dev = L.DeviceArray(8)
    for n, p, seed in ((big, 0.3, 1), (small, 0.02, 2)):
        f = _flags(n, p, seed)
        want = _with_alloc(L, f)
        got = _no_alloc(L, f, out, cnt_dev)
        assert np.array_equal(want, np.flatnonzero(f))
        assert got.size == want.size and np.array_equal(got, want), (n, got.size, want.size)
    # edge cases: no flags at all, no set flag, every flag set
    for f in (np.zeros(0, np.uint8), np.zeros(5000, np.uint8), np.ones(4096, np.uint8)):
        assert np.array_equal(_no_alloc(L, f, out, cnt_dev), np.flatnonzero(f))


def test_noalloc_refuses_more_than_its_maximum():
    from sfparticles_amd import _lib as L
    cnt_dev = L.DeviceArray(8)
    with pytest.raises(L.SFPError, match="without allocation"):
        L.call("sfp_find_event_index_noalloc", cnt_dev.ptr, (1 << 26) + 1, cnt_dev.ptr,
               cnt_dev.ptr, None)
