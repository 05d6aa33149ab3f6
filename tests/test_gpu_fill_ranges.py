"""Zero sectors of a product in one launch (xh_fill_ranges).  Needs an MI355X.

operator* sets the output to zero before the sector products (Tout.setZero(), TensorBase.cpp:256-258),
so a sector of the product that no pair of blocks writes stays zero.  The library fills those
sectors, merged into ranges of adjacent sectors, with one launch per product.  Checks:

* xh_fill_ranges writes exactly its ranges (ranges passed as arguments, up to 8, and from a table),
  leaves the rest untouched and refuses a range outside the buffer before launching;
* a product with a sector that no block pair writes gives zeros there, also when its output block
  comes from the context's cache of freed blocks holding NaNs, and matches the oracle.
"""
import ctypes as C

import numpy as np
import pytest

import xped_amd.tensor as xt
from oracle import tensor as oten
from xped_amd._lib import lib
from xped_amd.tensor import Runtime

from .helpers import both_tensor, oracle_structure, structure_of

pytestmark = pytest.mark.gpu


def _buffer(L, ctx, n, value):
    p = C.c_void_p()
    assert L.xh_alloc(ctx, 8 * n, C.byref(p)) == 0
    assert L.xh_fill(ctx, p, n, value) == 0
    return p


def _download(L, ctx, p, n):
    out = np.empty(n)
    assert L.xh_download(ctx, out.ctypes.data, p, 8 * n) == 0
    return out


@pytest.mark.parametrize("nranges", [0, 3, 8, 9, 300])
def test_fill_ranges_writes_exactly_its_ranges(nranges):
    rt = Runtime(0)
    L, ctx = lib(), rt.ctx()
    n = 200_000
    r = np.random.default_rng(nranges)
    cuts = np.sort(r.choice(np.arange(1, n), size=2 * nranges, replace=False))
    ranges = np.array([(cuts[2 * i], cuts[2 * i + 1] - cuts[2 * i]) for i in range(nranges)], dtype=np.int64).reshape(-1)
    if nranges >= 3:
        ranges[3] = 0  # an empty range
    p = _buffer(L, ctx, n, 1.5)
    assert L.xh_fill_ranges(ctx, p, n, ranges.ctypes.data, nranges, -2.0) == 0
    want = np.full(n, 1.5)
    for i in range(nranges):
        want[ranges[2 * i]:ranges[2 * i] + ranges[2 * i + 1]] = -2.0
    assert np.array_equal(_download(L, ctx, p, n), want)
    # a range past the end is refused and nothing is written
    bad = np.array([10, 5, n - 3, 4], dtype=np.int64)
    assert L.xh_fill_ranges(ctx, p, n, bad.ctypes.data, 2, 7.0) != 0
    assert np.array_equal(_download(L, ctx, p, n), want)
    assert L.xh_free(ctx, p) == 0


def test_unwritten_product_sector_is_zero_on_a_reused_block():
    rt = Runtime(0)
    L, ctx = lib(), rt.ctx()
    # A's domain carries charge 1, which its codomain (the contracted leg) lacks: the product's
    # sector 1 (4 x 6) has no block pair
    oA, A = both_tensor(rt, xt.U1, [[(0, 3), (1, 4)], [(0, 2)]], 1, 3)
    oB, B = both_tensor(rt, xt.U1, [[(0, 2)], [(0, 5), (1, 6)]], 1, 4)
    oC = oten.mul(oA, oB)
    C1 = A * B
    assert structure_of(C1) == oracle_structure(oC)
    np.testing.assert_allclose(C1.download(), oC.data, rtol=0, atol=1e-13)
    # freed blocks of the product's size class now hold NaNs; the next product's output reuses one
    blocks = [_buffer(L, ctx, C1.size, float("nan")) for _ in range(16)]
    for p in blocks:
        assert L.xh_free(ctx, p) == 0
    h0 = rt.device_cache()["hits"]
    C2 = A * B
    assert rt.device_cache()["hits"] > h0
    got = C2.download()
    assert not np.isnan(got).any()
    np.testing.assert_allclose(got, oC.data, rtol=0, atol=1e-13)
    sec = [s for s in structure_of(C2) if s[0] == (1,)]
    assert len(sec) == 1
    _, off, rows, cols = sec[0]
    assert np.array_equal(got[off:off + rows * cols], np.zeros(rows * cols))
