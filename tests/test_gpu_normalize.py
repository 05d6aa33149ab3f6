"""Device-side max-norm normalisation of the CTM moves (needs an MI355X).

`xh_normalize_max_abs` computes y = x / max|x| without reading the norm back; the result must
equal the host formula x * (1 / max|x|) bit for bit (the reference's `C / C.maxNorm()`,
CTM.cpp:1372).  A zero or NaN norm copies x and is counted until `xh_normalize_status` reports it.
"""
import ctypes as C

import numpy as np
import pytest

from xped_amd import _lib
from xped_amd._lib import lib

pytestmark = pytest.mark.gpu


def _run(L, ctx, x):
    n = x.size
    dx, dy = C.c_void_p(), C.c_void_p()
    assert L.xh_alloc(ctx, 8 * n, C.byref(dx)) == 0
    assert L.xh_alloc(ctx, 8 * n, C.byref(dy)) == 0
    L.xh_upload(ctx, dx, x.ctypes.data, 8 * n)
    assert L.xh_normalize_max_abs(ctx, dx, dy, n) == 0
    y = np.empty(n)
    L.xh_download(ctx, y.ctypes.data, dy, 8 * n)
    L.xh_free(ctx, dx)
    L.xh_free(ctx, dy)
    return y


def _status(L, ctx):
    count, first = C.c_int64(-1), C.c_double(-1.0)
    assert L.xh_normalize_status(ctx, C.byref(count), C.byref(first)) == 0
    return count.value, first.value


def test_normalize_matches_host_formula_and_counts_bad_norms(rt):
    L = lib()
    ctx = rt.ctx()
    r = np.random.default_rng(5)
    assert _status(L, ctx)[0] == 0
    for n in (1, 7, 256, 1000, 1 << 20, 3_000_001):
        x = r.standard_normal(n) * 10.0 ** r.uniform(-3, 3)
        y = _run(L, ctx, x)
        assert np.array_equal(y, x * (1.0 / np.abs(x).max())), n
    assert _status(L, ctx) == (0, 0.0)
    z = np.zeros(300)
    assert np.array_equal(_run(L, ctx, z), z)
    w = r.standard_normal(5000)
    w[17] = np.nan
    y = _run(L, ctx, w)
    assert np.array_equal(np.isnan(y), np.isnan(w)) and np.array_equal(y[~np.isnan(w)], w[~np.isnan(w)])
    count, first = _status(L, ctx)
    assert count == 2 and first == 0.0
    assert _status(L, ctx) == (0, 0.0)  # reset by the report
    x = r.standard_normal(999)
    assert np.array_equal(_run(L, ctx, x), x * (1.0 / np.abs(x).max()))
    assert _status(L, ctx) == (0, 0.0)
