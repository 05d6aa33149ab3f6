"""xh_gemm_noalloc: the grouped GEMM whose split-K partials live in the library's fixed per-device
scratch array instead of a per-call allocation.  Checked against xh_gemm (the same kernels, so bit
for bit) and against numpy, on two calls in a row, the second with a smaller input, and refused
with a clear error for a plan whose partials exceed the array."""
import ctypes as C

import numpy as np
import pytest

from xped_amd import _lib
from xped_amd._lib import lib

pytestmark = pytest.mark.gpu


def _items(shapes):
    items, ao, bo, co = [], 0, 0, 0
    for m, n, k in shapes:
        items.append((ao, bo, co, m, n, k))
        ao += m * k
        bo += k * n
        co += m * n
    return items, ao, bo, co


def _run_pair(rt, shapes, seed, precision=_lib.XH_FP64):
    L = lib()
    ctx = rt.ctx()
    items, ao, bo, co = _items(shapes)
    r = np.random.default_rng(seed)
    A, B = r.uniform(-1, 1, ao), r.uniform(-1, 1, bo)
    C0 = r.uniform(-1, 1, co)  # beta != 0: the previous output is read
    bufs = [C.c_void_p() for _ in range(4)]
    for d, n in zip(bufs, (ao, bo, co, co)):
        assert L.xh_alloc(ctx, 8 * max(n, 1), C.byref(d)) == 0
    dA, dB, dC1, dC2 = bufs
    L.xh_upload(ctx, dA, A.ctypes.data, 8 * ao)
    L.xh_upload(ctx, dB, B.ctypes.data, 8 * bo)
    L.xh_upload(ctx, dC1, C0.ctypes.data, 8 * co)
    L.xh_upload(ctx, dC2, C0.ctypes.data, 8 * co)
    arr = (_lib.xh_gemm_item * len(items))(*[_lib.xh_gemm_item(*it) for it in items])
    plan = C.c_void_p()
    assert L.xh_gemm_plan_create(ctx, arr, len(items), 0, 0, C.byref(plan)) == 0
    ws = C.c_int64(-1)
    assert L.xh_gemm_plan_workspace(plan, C.byref(ws)) == 0
    assert ws.value > 0, "the shapes must give a split-K plan"
    assert L.xh_gemm(ctx, plan, dA, dB, dC1, 1.5, -0.5, precision) == 0
    assert L.xh_gemm_noalloc(ctx, plan, dA, dB, dC2, 1.5, -0.5, precision) == 0, L.xh_last_error(ctx)
    out1, out2 = np.empty(co), np.empty(co)
    L.xh_download(ctx, out1.ctypes.data, dC1, 8 * co)
    L.xh_download(ctx, out2.ctypes.data, dC2, 8 * co)
    assert np.array_equal(out1, out2), "xh_gemm_noalloc differs from xh_gemm"
    for (a, b, c, m, n, k) in items:
        ref = 1.5 * A[a:a + m * k].reshape(k, m).T @ B[b:b + k * n].reshape(n, k).T - 0.5 * C0[c:c + m * n].reshape(n, m).T
        got = out2[c:c + m * n].reshape(n, m).T
        assert np.abs(got - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max()) * np.sqrt(k)
    L.xh_gemm_plan_destroy(ctx, plan)
    for d in bufs:
        L.xh_free(ctx, d)
    return ws.value


def test_noalloc_gemm_matches_xh_gemm_twice(rt):
    # long K and few output tiles: the plans split K into chunks whose partials use the scratch
    big = _run_pair(rt, [(130, 70, 2048), (40, 30, 1500), (5, 3, 777), (200, 150, 129)], seed=11)
    small = _run_pair(rt, [(33, 17, 900), (7, 9, 300)], seed=12)
    assert 0 < small < big


def test_noalloc_gemm_refuses_plans_beyond_the_scratch(rt):
    L = lib()
    ctx = rt.ctx()
    items = (_lib.xh_gemm_item * 1)(_lib.xh_gemm_item(0, 0, 0, 1000, 1000, 20000))
    plan = C.c_void_p()
    assert L.xh_gemm_plan_create(ctx, items, 1, 0, 0, C.byref(plan)) == 0
    ws = C.c_int64(0)
    L.xh_gemm_plan_workspace(plan, C.byref(ws))
    assert ws.value > (32 << 20)
    d = C.c_void_p()
    assert L.xh_alloc(ctx, 64, C.byref(d)) == 0
    # rejected before any launch, so the undersized buffers are never touched
    assert L.xh_gemm_noalloc(ctx, plan, d, d, d, 1.0, 0.0, _lib.XH_FP64) == _lib.XH_EINVAL
    assert b"scratch" in L.xh_last_error(ctx)
    L.xh_gemm_plan_destroy(ctx, plan)
    L.xh_free(ctx, d)
