"""The context's cache of freed device blocks (xh_set_device_cache / xh_device_cache_stats).  Needs an MI355X.

Freed blocks are kept in size classes (1 KiB, then four per power of two) and handed out again in
stream order.  Checks:

* a request of the same class gets the freed block back (a hit), another class reaches the runtime;
* a block freed while a queued launch still reads it can serve the next request at once: the
  launch sees the old contents, later work the new ones (one stream orders both);
* a zero bound empties the cache and keeps nothing afterwards.
"""
import ctypes as C

import numpy as np
import pytest

from xped_amd._lib import lib
from xped_amd.tensor import Runtime

pytestmark = pytest.mark.gpu


def _alloc(L, ctx, nbytes):
    p = C.c_void_p()
    assert L.xh_alloc(ctx, nbytes, C.byref(p)) == 0
    return p


def test_freed_blocks_are_reused_by_size_class():
    rt = Runtime(0)
    L, ctx = lib(), rt.ctx()
    s0 = rt.device_cache()
    a = _alloc(L, ctx, 8000)  # class 8192 (a multiple of 1024 between 4096 and 8192)
    assert L.xh_free(ctx, a) == 0
    assert rt.device_cache()["idle_bytes"] == s0["idle_bytes"] + 8192
    b = _alloc(L, ctx, 8192)  # the same class: the same block
    assert b.value == a.value
    s1 = rt.device_cache()
    assert s1["hits"] == s0["hits"] + 1 and s1["idle_bytes"] == s0["idle_bytes"]
    c = _alloc(L, ctx, 7000)  # class 7168: a new block
    assert c.value != b.value
    assert rt.device_cache()["misses"] == s1["misses"] + 1
    for p in (b, c):
        assert L.xh_free(ctx, p) == 0


def test_a_block_freed_behind_a_queued_launch_serves_the_next_request():
    rt = Runtime(0)
    L, ctx = lib(), rt.ctx()
    r = np.random.default_rng(2)
    n = 1 << 22
    x, x2 = r.standard_normal(n), r.standard_normal(n)
    dx, dy = _alloc(L, ctx, 8 * n), _alloc(L, ctx, 8 * n)
    assert L.xh_upload(ctx, dx, x.ctypes.data, 8 * n) == 0
    # y = x / max|x| is queued; x's block is freed and taken again before the launch has run
    assert L.xh_normalize_max_abs(ctx, dx, dy, n) == 0
    assert L.xh_free(ctx, dx) == 0
    dz = _alloc(L, ctx, 8 * n)
    assert dz.value == dx.value
    assert L.xh_upload(ctx, dz, x2.ctypes.data, 8 * n) == 0
    y, z = np.empty(n), np.empty(n)
    assert L.xh_download(ctx, y.ctypes.data, dy, 8 * n) == 0
    assert L.xh_download(ctx, z.ctypes.data, dz, 8 * n) == 0
    assert np.array_equal(y, x * (1.0 / np.abs(x).max()))
    assert np.array_equal(z, x2)
    for p in (dy, dz):
        assert L.xh_free(ctx, p) == 0


def test_zero_bound_keeps_no_idle_blocks():
    rt = Runtime(0)
    L, ctx = lib(), rt.ctx()
    a = _alloc(L, ctx, 1 << 20)
    assert L.xh_free(ctx, a) == 0
    assert rt.device_cache()["idle_bytes"] >= 1 << 20
    assert rt.device_cache(0)["idle_bytes"] == 0
    m0 = rt.device_cache()["misses"]
    b = _alloc(L, ctx, 1 << 20)
    assert L.xh_free(ctx, b) == 0
    s = rt.device_cache()
    assert s["idle_bytes"] == 0 and s["misses"] == m0 + 1
    # with a bound again, the same request is served by the cache
    rt.device_cache(1 << 30)
    c = _alloc(L, ctx, 1 << 20)
    assert L.xh_free(ctx, c) == 0
    d = _alloc(L, ctx, 1 << 20)
    assert d.value == c.value
    assert L.xh_free(ctx, d) == 0
