"""HIP path vs the CPU oracle through the C ABI (needs an MI355X).

Tolerances: permute of abelian tensors is pure data movement (negation for fermionic signs) and
must be bit-exact; SU(2) permutes use recoupling coefficients and agree to 1e-13 absolute on
O(1) data; FP64 GEMM-based results agree to 1e-12 relative to the largest entry; FP32 mode to
1e-5 and BF16x3 mode to 5e-5 relative (stated per test).
"""
import ctypes as C
import itertools
import math
import random

import numpy as np
import pytest

import xped_amd.tensor as xt
from oracle import tensor as oten
from oracle.structure import inverse
from xped_amd import _lib, configs
from xped_amd._lib import lib

from .helpers import both_basis, both_tensor, oracle_structure, rand_entries, structure_of

pytestmark = pytest.mark.gpu

PCODES = [xt.U1, xt.FU1xU1, xt.FU1, xt.SU2]
IDS = ["U1", "FU1xU1", "FU1", "SU2"]


def test_set_random_matches_reference_generator(rt):
    r = random.Random(1)
    for dist in (0, 1):
        legs = [rand_entries(xt.U1, r, 4, 5) for _ in range(3)]
        o, t = both_tensor(rt, xt.U1, legs, 2, 77, dist)
        t.setRandom(77, dist)
        assert np.array_equal(t.download(), o.data)


@pytest.mark.parametrize("pcode", PCODES, ids=IDS)
def test_permute_all_permutations(rt, pcode):
    r = random.Random(2)
    for rank, corank in [(2, 2), (2, 1), (1, 2), (3, 1)]:
        legs = [rand_entries(pcode, r, 3, 4) for _ in range(rank + corank)]
        o, t = both_tensor(rt, pcode, legs, rank, 10 + rank)
        for p in itertools.permutations(range(rank + corank)):
            for shift in range(-corank, rank + 1):
                op = oten.permute(o, shift, p)
                tp = t.permute(shift, p)
                assert structure_of(tp) == oracle_structure(op)
                got = tp.download()
                if pcode == xt.SU2:
                    np.testing.assert_allclose(got, op.data, rtol=0, atol=1e-13)
                else:
                    assert np.array_equal(got, op.data), (p, shift)


@pytest.mark.parametrize("pcode", [xt.U1, xt.FU1xU1], ids=["U1", "FU1xU1"])
def test_permute_large_subblocks_transpose_path(rt, pcode):
    # degeneracies >= 16 route subblocks through the LDS transpose kernel
    r = random.Random(3)
    legs = [rand_entries(pcode, r, 3, 70, mindim=16) for _ in range(4)]
    o, t = both_tensor(rt, pcode, legs, 2, 5)
    for p, shift in [((1, 0, 2, 3), 0), ((2, 3, 0, 1), 0), ((3, 1, 2, 0), 1), ((1, 2, 3, 0), -1), ((0, 3, 2, 1), 2)]:
        op = oten.permute(o, shift, p)
        tp = t.permute(shift, p)
        assert structure_of(tp) == oracle_structure(op)
        assert np.array_equal(tp.download(), op.data), (p, shift)


def test_permute_timing_counts_launches_and_bytes(rt):
    # an abelian permute reads and writes every element once: 16 bytes per FP64 element
    r = random.Random(4)
    legs = [rand_entries(xt.U1, r, 3, 40, mindim=16) for _ in range(4)]
    o, t = both_tensor(rt, xt.U1, legs, 2, 6)
    rt.permute_timing(True, reset=True)
    outs = [t.permute(0, (2, 3, 0, 1)), t.permute(1, (3, 1, 2, 0))]
    ms, nbytes, n = rt.permute_timing(False, reset=True)
    assert n == 2 and ms > 0.0
    assert nbytes == 16.0 * sum(x.size for x in outs)
    assert rt.permute_timing(False)[2] == 0
    assert np.array_equal(outs[0].download(), oten.permute(o, 0, (2, 3, 0, 1)).data)


def test_permute_round_trip_su2(rt):
    r = random.Random(4)
    legs = [rand_entries(xt.SU2, r, 3, 5) for _ in range(4)]
    o, t = both_tensor(rt, xt.SU2, legs, 2, 6)
    d0 = t.download()
    for p in itertools.permutations(range(4)):
        for shift in (-2, -1, 0, 1, 2):
            back = t.permute(shift, p).permute(-shift, inverse(p))
            np.testing.assert_allclose(back.download(), d0, rtol=0, atol=1e-13)


def _gemm_ref(A, B, items, ta, tb):
    C = {}
    for it in items:
        m, n, k = it["m"], it["n"], it["k"]
        a = A[it["a"]:it["a"] + m * k].reshape((m, k) if ta else (k, m)).T if True else None
        a = A[it["a"]:it["a"] + m * k].reshape(m, k) if ta else A[it["a"]:it["a"] + m * k].reshape(k, m).T
        b = B[it["b"]:it["b"] + k * n].reshape(k, n) if tb else B[it["b"]:it["b"] + k * n].reshape(n, k).T
        C[it["c"]] = (a @ b, m, n)
    return C


@pytest.mark.parametrize("precision,tol", [(_lib.XH_FP64, 1e-12), (_lib.XH_FP32, 1e-5), (_lib.XH_BF16X3, 5e-5), (_lib.XH_OZAKI, 1e-12)],
                         ids=["fp64", "fp32", "bf16x3", "ozaki"])
@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_grouped_gemm_raw(rt, precision, tol, ta, tb):
    L = lib()
    ctx = rt.ctx()
    r = np.random.default_rng(5)
    # (40, 30, 1500), (130, 70, 2048), (5, 3, 777) have a long K and few output tiles: the plan
    # splits their K into chunks; (260, 300, 70) and (193, 450, 33) take the 128 x 128 FP64 tiles
    shapes = [(1, 1, 1), (17, 5, 3), (64, 64, 16), (65, 63, 17), (130, 7, 200), (8, 300, 33), (200, 150, 129), (3, 3, 0),
              (40, 30, 1500), (130, 70, 2048), (5, 3, 777), (260, 300, 70), (193, 450, 33)]
    items, ao, bo, co = [], 0, 0, 0
    for m, n, k in shapes:
        items.append({"a": ao, "b": bo, "c": co, "m": m, "n": n, "k": k})
        ao += m * k
        bo += k * n
        co += m * n
    A = r.uniform(-1, 1, ao)
    B = r.uniform(-1, 1, bo)
    dA, dB, dC = C.c_void_p(), C.c_void_p(), C.c_void_p()
    for d, n in ((dA, ao), (dB, bo), (dC, co)):
        assert L.xh_alloc(ctx, 8 * max(n, 1), C.byref(d)) == 0
    L.xh_upload(ctx, dA, A.ctypes.data, 8 * ao)
    L.xh_upload(ctx, dB, B.ctypes.data, 8 * bo)
    L.xh_fill(ctx, dC, co, 0.0)
    arr = (_lib.xh_gemm_item * len(items))(*[_lib.xh_gemm_item(it["a"], it["b"], it["c"], it["m"], it["n"], it["k"]) for it in items])
    plan = C.c_void_p()
    assert L.xh_gemm_plan_create(ctx, arr, len(items), ta, tb, C.byref(plan)) == 0
    assert L.xh_gemm(ctx, plan, dA, dB, dC, 1.0, 0.0, precision) == 0
    out = np.empty(co)
    L.xh_download(ctx, out.ctypes.data, dC, 8 * co)
    ref = _gemm_ref(A, B, items, ta, tb)
    for it in items:
        c, m, n = ref[it["c"]]
        got = out[it["c"]:it["c"] + m * n].reshape(n, m).T
        scale = max(1.0, np.abs(c).max() if c.size else 1.0) * math.sqrt(max(it["k"], 1))
        assert np.abs(got - c).max(initial=0.0) <= tol * scale, (it, np.abs(got - c).max())
    # alpha/beta accumulate
    assert L.xh_gemm(ctx, plan, dA, dB, dC, 2.0, -1.0, precision) == 0
    out2 = np.empty(co)
    L.xh_download(ctx, out2.ctypes.data, dC, 8 * co)
    for it in items:
        c, m, n = ref[it["c"]]
        got = out2[it["c"]:it["c"] + m * n].reshape(n, m).T
        scale = max(1.0, np.abs(c).max() if c.size else 1.0) * math.sqrt(max(it["k"], 1))
        assert np.abs(got - c).max(initial=0.0) <= 4 * tol * scale
    L.xh_gemm_plan_destroy(ctx, plan)
    for d in (dA, dB, dC):
        L.xh_free(ctx, d)


@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 1)])
def test_ozaki_gemm_error_bound_and_nonfinite_rows(rt, ta, tb):
    # XH_OZAKI against an extended-precision reference, on rows and columns scaled over 16 decades,
    # K beyond one int32 fold (4096) and split K: every entry within the classical FP64 GEMM bound
    # k u (|A| |B|)_ij (u = 2^-53), and the largest error relative to max|A_i.| max|B_.j| within 4x
    # that of the FP64 matrix cores on the same products; a NaN in A spoils exactly its row, an inf
    # in B exactly its column.  Items with a side or K below 64 (the second, fourth and fifth here)
    # run on the FP64 kernels in the same call.
    L = lib()
    ctx = rt.ctx()
    r = np.random.default_rng(9)
    shapes = [(70, 90, 5000), (33, 65, 129), (16, 16, 9000), (1, 200, 64), (130, 1, 63)]
    items, ao, bo, co = [], 0, 0, 0
    for m, n, k in shapes:
        items.append({"a": ao, "b": bo, "c": co, "m": m, "n": n, "k": k})
        ao += m * k
        bo += k * n
        co += m * n
    A = r.standard_normal(ao)
    B = r.standard_normal(bo)
    mats = []
    for it in items:
        m, n, k = it["m"], it["n"], it["k"]
        a = 10.0 ** r.uniform(-8, 8, (m, 1)) * r.standard_normal((m, k))
        b = r.standard_normal((k, n)) * 10.0 ** r.uniform(-8, 8, (1, n))
        A[it["a"]:it["a"] + m * k] = (a if ta else a.T).ravel()
        B[it["b"]:it["b"] + k * n] = (b if tb else b.T).ravel()
        mats.append((a, b))
    a0 = mats[0][0]
    a0[3, 17] = np.nan
    A[items[0]["a"]:items[0]["a"] + a0.size] = (a0 if ta else a0.T).ravel()
    b1 = mats[1][1]
    b1[5, 7] = np.inf
    B[items[1]["b"]:items[1]["b"] + b1.size] = (b1 if tb else b1.T).ravel()
    dA, dB, dC = C.c_void_p(), C.c_void_p(), C.c_void_p()
    for d, n in ((dA, ao), (dB, bo), (dC, co)):
        assert L.xh_alloc(ctx, 8 * max(n, 1), C.byref(d)) == 0
    L.xh_upload(ctx, dA, A.ctypes.data, 8 * ao)
    L.xh_upload(ctx, dB, B.ctypes.data, 8 * bo)
    arr = (_lib.xh_gemm_item * len(items))(*[_lib.xh_gemm_item(it["a"], it["b"], it["c"], it["m"], it["n"], it["k"]) for it in items])
    plan = C.c_void_p()
    assert L.xh_gemm_plan_create(ctx, arr, len(items), ta, tb, C.byref(plan)) == 0
    outs = {}
    for prec in (_lib.XH_OZAKI, _lib.XH_FP64):
        L.xh_fill(ctx, dC, co, 0.0)
        assert L.xh_gemm(ctx, plan, dA, dB, dC, 1.0, 0.0, prec) == 0
        out = np.empty(co)
        L.xh_download(ctx, out.ctypes.data, dC, 8 * co)
        outs[prec] = out
    L.xh_gemm_plan_destroy(ctx, plan)
    for d in (dA, dB, dC):
        L.xh_free(ctx, d)
    for idx, (it, (a, b)) in enumerate(zip(items, mats)):
        m, n, k = it["m"], it["n"], it["k"]
        got = outs[_lib.XH_OZAKI][it["c"]:it["c"] + m * n].reshape(n, m).T
        f64 = outs[_lib.XH_FP64][it["c"]:it["c"] + m * n].reshape(n, m).T
        bad = np.zeros((m, n), bool)
        if idx == 0:
            bad[3, :] = True
        if idx == 1:
            bad[:, 7] = True
        assert not np.any(np.isfinite(got[bad])), "a non-finite input must spoil its row / column"
        ok = ~bad
        with np.errstate(invalid="ignore"):
            ref = (a.astype(np.longdouble) @ b.astype(np.longdouble)).astype(np.float64)
            af = np.where(np.isfinite(a), np.abs(a), 0.0)
            bf = np.where(np.isfinite(b), np.abs(b), 0.0)
            amax = af.max(axis=1)[:, None]
            bmax = bf.max(axis=0)[None, :]
        bound = k * 2.0 ** -53 * (af @ bf) + 2.0 ** -53 * np.abs(ref)
        err = np.abs(got - ref)
        assert np.all(err[ok] <= bound[ok]), (it, float(np.max(err[ok] / bound[ok])))
        # no worse than the FP64 matrix cores on the same products (relative to |A_i.| |B_.j|)
        scale = amax * bmax
        assert np.max(err[ok] / scale[ok]) <= 4.0 * max(np.max(np.abs(f64 - ref)[ok] / scale[ok]), 2.0 ** -53)


@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_grouped_gemm_skinny_items(rt, ta, tb):
    # items with a short output side (<= 32) and K <= 32 run on the memory-bound skinny kernel,
    # mixed in one plan with items on the matrix-core tiles; FP64 to 1e-12
    L = lib()
    ctx = rt.ctx()
    r = np.random.default_rng(9)
    shapes = [(16, 5000, 9), (7000, 24, 13), (32, 300, 32), (31, 31, 31), (1, 2000, 1), (3000, 1, 5), (5, 7, 0), (200, 150, 40), (9, 9000, 8)]
    items, ao, bo, co = [], 0, 0, 0
    for m, n, k in shapes:
        items.append({"a": ao, "b": bo, "c": co, "m": m, "n": n, "k": k})
        ao += m * k
        bo += k * n
        co += m * n
    A, B, C0 = r.uniform(-1, 1, ao), r.uniform(-1, 1, bo), r.uniform(-1, 1, co)
    dA, dB, dC = C.c_void_p(), C.c_void_p(), C.c_void_p()
    for d, n in ((dA, ao), (dB, bo), (dC, co)):
        assert L.xh_alloc(ctx, 8 * max(n, 1), C.byref(d)) == 0
    L.xh_upload(ctx, dA, A.ctypes.data, 8 * ao)
    L.xh_upload(ctx, dB, B.ctypes.data, 8 * bo)
    L.xh_upload(ctx, dC, C0.ctypes.data, 8 * co)
    arr = (_lib.xh_gemm_item * len(items))(*[_lib.xh_gemm_item(it["a"], it["b"], it["c"], it["m"], it["n"], it["k"]) for it in items])
    plan = C.c_void_p()
    assert L.xh_gemm_plan_create(ctx, arr, len(items), ta, tb, C.byref(plan)) == 0
    assert L.xh_gemm(ctx, plan, dA, dB, dC, 1.5, -0.5, _lib.XH_FP64) == 0
    out = np.empty(co)
    L.xh_download(ctx, out.ctypes.data, dC, 8 * co)
    ref = _gemm_ref(A, B, items, ta, tb)
    for it in items:
        c, m, n = ref[it["c"]]
        want = 1.5 * c - 0.5 * C0[it["c"]:it["c"] + m * n].reshape(n, m).T
        got = out[it["c"]:it["c"] + m * n].reshape(n, m).T
        assert np.abs(got - want).max(initial=0.0) <= 1e-12 * max(1.0, np.abs(want).max(initial=0.0)) * math.sqrt(max(it["k"], 1)), it
    L.xh_gemm_plan_destroy(ctx, plan)
    for d in (dA, dB, dC):
        L.xh_free(ctx, d)


def test_gemm_timing_and_marker(rt):
    L = lib()
    ctx = rt.ctx()
    m, n, k = 96, 80, 1200
    r = np.random.default_rng(6)
    A, B = r.uniform(-1, 1, m * k), r.uniform(-1, 1, k * n)
    dA, dB, dC = C.c_void_p(), C.c_void_p(), C.c_void_p()
    for d, sz in ((dA, m * k), (dB, k * n), (dC, m * n)):
        assert L.xh_alloc(ctx, 8 * sz, C.byref(d)) == 0
    L.xh_upload(ctx, dA, A.ctypes.data, 8 * m * k)
    L.xh_upload(ctx, dB, B.ctypes.data, 8 * k * n)
    arr = (_lib.xh_gemm_item * 1)(_lib.xh_gemm_item(0, 0, 0, m, n, k))
    plan = C.c_void_p()
    assert L.xh_gemm_plan_create(ctx, arr, 1, 0, 0, C.byref(plan)) == 0
    ms, fl, nl = C.c_double(), C.c_double(), C.c_int64()
    assert L.xh_gemm_timing(ctx, 1, 1, C.byref(ms), C.byref(fl), C.byref(nl)) == 0
    assert L.xh_trace_marker(ctx, 1) == 0
    for _ in range(3):
        assert L.xh_gemm(ctx, plan, dA, dB, dC, 1.0, 0.0, _lib.XH_FP64) == 0
    assert L.xh_trace_marker(ctx, 2) == 0
    assert L.xh_gemm_timing(ctx, 0, 1, C.byref(ms), C.byref(fl), C.byref(nl)) == 0
    assert nl.value == 3 and fl.value == 3 * 2.0 * m * n * k and ms.value > 0.0
    assert L.xh_trace_marker(ctx, 0) != 0
    out = np.empty(m * n)
    L.xh_download(ctx, out.ctypes.data, dC, 8 * m * n)
    ref = A.reshape(k, m).T @ B.reshape(n, k).T
    np.testing.assert_allclose(out.reshape(n, m).T, ref, rtol=0, atol=1e-12 * np.abs(ref).max() * math.sqrt(k))
    L.xh_gemm_plan_destroy(ctx, plan)
    for d in (dA, dB, dC):
        L.xh_free(ctx, d)


def test_gemm_autotuner_picks_a_tiling_and_reuses_it(rt):
    # items with both output sides >= 192 have 128 x 128 tiles, so both tilings are candidates
    L = lib()
    ctx = rt.ctx()
    r = np.random.default_rng(9)
    items_mnk = [(320, 256, 200), (200, 192, 64), (96, 80, 50)]
    offs, a_tot, b_tot, c_tot = [], 0, 0, 0
    for m, n, k in items_mnk:
        offs.append((a_tot, b_tot, c_tot))
        a_tot += m * k
        b_tot += k * n
        c_tot += m * n
    A, B = r.uniform(-1, 1, a_tot), r.uniform(-1, 1, b_tot)
    dA, dB, dC = C.c_void_p(), C.c_void_p(), C.c_void_p()
    for d, sz in ((dA, a_tot), (dB, b_tot), (dC, c_tot)):
        assert L.xh_alloc(ctx, 8 * sz, C.byref(d)) == 0
    L.xh_upload(ctx, dA, A.ctypes.data, 8 * a_tot)
    L.xh_upload(ctx, dB, B.ctypes.data, 8 * b_tot)
    arr = (_lib.xh_gemm_item * 3)(*[_lib.xh_gemm_item(ao, bo, co, m, n, k) for (ao, bo, co), (m, n, k) in zip(offs, items_mnk)])
    plan = C.c_void_p()
    assert L.xh_gemm_plan_create(ctx, arr, 3, 0, 0, C.byref(plan)) == 0
    stats = [C.c_int64() for _ in range(4)]

    def run():
        assert L.xh_gemm(ctx, plan, dA, dB, dC, 1.0, 0.0, _lib.XH_FP64) == 0
        out = np.empty(c_tot)
        L.xh_download(ctx, out.ctypes.data, dC, 8 * c_tot)
        return out

    assert L.xh_set_gemm_autotune(ctx, 1) == 0
    try:
        tuned = run()
        assert L.xh_gemm_autotune_stats(ctx, *[C.byref(x) for x in stats]) == 0
        entries, runs, hits, chose64 = (x.value for x in stats)
        assert runs >= 1 and entries >= 1
        again = run()
        assert L.xh_gemm_autotune_stats(ctx, *[C.byref(x) for x in stats]) == 0
        assert stats[2].value == hits + 1  # the second call reuses the cached choice
        assert np.array_equal(again, tuned)
    finally:
        assert L.xh_set_gemm_autotune(ctx, 0) == 0
    # the tuned result is exactly the chosen tiling's result, and a correct product
    assert L.xh_set_gemm_tiles(ctx, 1 if chose64 else 0) == 0
    try:
        assert np.array_equal(run(), tuned)
    finally:
        assert L.xh_set_gemm_tiles(ctx, 0) == 0
    for (ao, bo, co), (m, n, k) in zip(offs, items_mnk):
        ref = A[ao:ao + m * k].reshape(k, m).T @ B[bo:bo + k * n].reshape(n, k).T
        got = tuned[co:co + m * n].reshape(n, m).T
        np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12 * np.abs(ref).max() * math.sqrt(k))
    L.xh_gemm_plan_destroy(ctx, plan)
    for d in (dA, dB, dC):
        L.xh_free(ctx, d)


@pytest.mark.parametrize("pcode", PCODES, ids=IDS)
def test_mul_and_contract(rt, pcode):
    r = random.Random(6)
    legs = [rand_entries(pcode, r, 3, 5) for _ in range(3)]
    legs2 = [rand_entries(pcode, r, 3, 5) for _ in range(2)]
    oA, A = both_tensor(rt, pcode, legs, 2, 1)
    oB, B = both_tensor(rt, pcode, [legs[2]] + legs2, 1, 2)
    oC = oten.mul(oA, oB)
    Cm = A * B
    assert structure_of(Cm) == oracle_structure(oC)
    np.testing.assert_allclose(Cm.download(), oC.data, rtol=0, atol=1e-12 * max(1, np.abs(oC.data).max(initial=0.0)))
    # contract over the shared leg with a reordered result
    oD = oten.contract(oA, [-2, -1, 1], oB, [1, -3, -4], 2)
    D = A.contract(B, [-2, -1, 1], [1, -3, -4], 2)
    assert structure_of(D) == oracle_structure(oD)
    np.testing.assert_allclose(D.download(), oD.data, rtol=0, atol=1e-12 * max(1, np.abs(oD.data).max(initial=0.0)))


def _c1_pair(rt):
    v = configs.c1_virtual()
    ph = configs.c1_physical()
    oA, A = both_tensor(rt, xt.U1, [v, ph, v], 2, 11)
    oB, B = both_tensor(rt, xt.U1, [v, ph, v], 2, 12)
    return oA, A, oB, B


def test_c1_two_site_contraction_and_truncation(rt):
    # configuration 1: contract over r, permute to (l,p|p',r'), tSVD truncated to chi = 64
    oA, A, oB, B = _c1_pair(rt)
    oC = oten.contract(oA, [-1, -2, 1], oB, [1, -3, -4], 2)
    Cm = A.contract(B, [-1, -2, 1], [1, -3, -4], 2)
    assert structure_of(Cm) == oracle_structure(oC)
    assert [s[2] for s in Cm.sectors()] == [1, 2, 8, 16, 25, 28, 25, 16, 8, 2, 1]
    np.testing.assert_allclose(Cm.download(), oC.data, rtol=0, atol=1e-12 * np.abs(oC.data).max())
    U, S, V, tw, ent = Cm.tSVD(64, 0.0, PRESERVE_MULTIPLETS=False)
    oU, oS, oV, otw, oent, kept = oten.tsvd(oC, 64, 0.0, False)
    assert [s[0] for s in S.sectors()] == list(kept.keys())
    assert [s[2] for s in S.sectors()] == [len(v) for v in kept.values()]
    smax = oS.data.max()
    np.testing.assert_allclose(S.download(), oS.data, rtol=0, atol=1e-12 * smax)
    assert abs(tw - otw) <= 1e-12 * smax ** 2
    assert abs(ent - oent) <= 1e-11 * max(1, abs(oent))
    # gauge-invariant check of the kept subspace: U S Vdag equals the oracle's
    rec = (U * S) * V
    orec = oten.mul(oten.mul(oU, oS), oV)
    np.testing.assert_allclose(rec.download(), orec.data, rtol=0, atol=1e-12 * smax)
    # orthonormal columns of U
    UtU = U.adjoint() * U
    ident = oten.mul(oten.adjoint(oU), oU)
    np.testing.assert_allclose(UtU.download(), ident.data, rtol=0, atol=1e-12)


@pytest.mark.parametrize("pcode", [xt.U1, xt.SU2, xt.FU1xU1], ids=["U1", "SU2", "FU1xU1"])
def test_full_svd_invariants(rt, pcode):
    r = random.Random(8)
    legs = [rand_entries(pcode, r, 4, 9) for _ in range(4)]
    o, t = both_tensor(rt, pcode, legs, 2, 3)
    U, S, V, tw, ent = t.tSVD(10 ** 6, 0.0, PRESERVE_MULTIPLETS=False)
    assert tw == 0.0
    rec = (U * S) * V
    np.testing.assert_allclose(rec.download(), o.data, rtol=0, atol=1e-12 * max(1, np.abs(o.data).max()) * 10)
    _, oS, _, _, _, kept = oten.tsvd(o, 10 ** 6, 0.0, False)
    np.testing.assert_allclose(np.sort(S.download()), np.sort(oS.data), rtol=0, atol=1e-12 * oS.data.max())


def test_tqr_and_teigh(rt):
    r = random.Random(9)
    legs = [rand_entries(xt.U1, r, 3, 6) for _ in range(3)]
    # tall domain so that the middle basis is the codomain
    o, t = both_tensor(rt, xt.U1, [legs[0], legs[1], legs[2]], 2, 4)
    Q, R = t.tQR()
    oQ, oR = oten.tqr(o)
    assert structure_of(Q) == oracle_structure(oQ) and structure_of(R) == oracle_structure(oR)
    np.testing.assert_allclose(Q.download(), oQ.data, rtol=0, atol=1e-12)
    np.testing.assert_allclose(R.download(), oR.data, rtol=0, atol=1e-12 * max(1, np.abs(oR.data).max()))
    # symmetric matrix for eigh: T^dagger T
    H = t.adjoint() * t
    oH = oten.mul(oten.adjoint(o), o)
    D, V = H.teigh()
    oD, oV = oten.teigh(oH)
    np.testing.assert_allclose(D.download(), oD.data, rtol=0, atol=1e-12 * max(1, np.abs(oD.data).max(initial=0.0)))
    rec = (V * D) * V.adjoint()
    np.testing.assert_allclose(rec.download(), oH.data, rtol=0, atol=1e-11 * max(1, np.abs(oH.data).max()))


@pytest.mark.parametrize("pcode", PCODES, ids=IDS)
def test_elementwise_ops(rt, pcode):
    r = random.Random(10)
    legs = [rand_entries(pcode, r, 3, 4) for _ in range(4)]
    o, t = both_tensor(rt, pcode, legs, 2, 8)
    for leg in range(4):
        assert np.array_equal(t.twist(leg).download(), oten.twist(o, leg).data)
    ad = t.adjoint()
    oad = oten.adjoint(o)
    assert structure_of(ad) == oracle_structure(oad)
    assert np.array_equal(ad.download(), oad.data)
    assert abs(t.squaredNorm() - oten.squared_norm(o)) <= 1e-12 * oten.squared_norm(o)
    assert t.maxNorm() == oten.max_norm(o)
    sq = t * t.adjoint()
    assert abs(sq.trace() - oten.trace(oten.mul(o, oad))) <= 1e-12 * abs(oten.trace(oten.mul(o, oad)))
    assert np.array_equal(t.scale(-2.5).download(), -2.5 * o.data)


def test_axpby_with_zero_b_does_not_read_the_output(rt):
    # scaling into a fresh buffer must not depend on its previous contents (0 * NaN = NaN)
    L = lib()
    ctx = rt.ctx()
    n = 1000
    x = np.random.default_rng(3).uniform(-1, 1, n)
    x[7] = -0.0
    dx, dy = C.c_void_p(), C.c_void_p()
    assert L.xh_alloc(ctx, 8 * n, C.byref(dx)) == 0 and L.xh_alloc(ctx, 8 * n, C.byref(dy)) == 0
    L.xh_upload(ctx, dx, x.ctypes.data, 8 * n)
    L.xh_fill(ctx, dy, n, float("nan"))
    assert L.xh_axpby(ctx, dy, dx, n, 2.5, 0.0) == 0
    y = np.empty(n)
    L.xh_download(ctx, y.ctypes.data, dy, 8 * n)
    L.xh_free(ctx, dx)
    L.xh_free(ctx, dy)
    assert np.array_equal(y, 2.5 * x) and np.signbit(y[7])


def test_cpp_facade_runs_the_hot_path_on_the_gpu(rt):
    """include/xped_amd.hpp driven from C++ (tests/cpp/adapter_test.cpp): setRandom, contract,
    permute, tSVD with SVspec and tQR on the device, against the oracle on the same seeds."""
    from oracle.structure import Qbasis
    from oracle.symmetry import U1, Sym

    from .cpp_adapter import build_and_run

    out = build_and_run("gpu")
    s = Sym(U1)
    virt, phys = Qbasis(s), Qbasis(s)
    for q in range(-2, 3):
        virt.push_back((q,), 3 - abs(q))
    phys.push_back((-1,), 1)
    phys.push_back((1,), 1)
    A = oten.Tensor(s, [virt, phys], [virt]).set_random(1)
    B = oten.Tensor(s, [virt, phys], [virt]).set_random(2)
    assert np.array_equal(np.array(out["A_data"]), A.data) and np.array_equal(np.array(out["B_data"]), B.data)
    Cc = oten.contract(A, [-1, -2, 1], B, [1, -3, -4], 2)
    assert out["C"] == [[q[0], o, r, c] for q, o, r, c in Cc.structure()]
    np.testing.assert_allclose(out["C_data"], Cc.data, rtol=0, atol=1e-12 * np.abs(Cc.data).max())
    P = oten.permute(Cc, -1, [0, 3, 1, 2])
    assert out["P"] == [[q[0], o, r, c] for q, o, r, c in P.structure()]
    np.testing.assert_allclose(out["P_data"], P.data, rtol=0, atol=1e-12 * np.abs(P.data).max())
    U, S, V, tw, ent, kept = oten.tsvd(Cc, 8, 0.0, False)
    assert out["S"] == [[q[0], o, r, c] for q, o, r, c in S.structure()]
    assert abs(out["truncWeight"] - tw) <= 1e-12 * max(1.0, tw) and abs(out["entropy"] - ent) <= 1e-12 * max(1.0, abs(ent))
    assert set(out["SVspec"]) == {str(q[0]) for q in kept}
    for q, v in kept.items():
        np.testing.assert_allclose(out["SVspec"][str(q[0])], v, rtol=0, atol=1e-12 * S.data.max())
    # C's blocks are rank-deficient (rank <= the dimension of the contracted sector), so the
    # trailing columns of Q are any orthonormal completion: compare the factorisation instead
    Q, R = oten.tqr(Cc)
    Q.data, R.data = np.array(out["Q_data"]), np.array(out["R_data"])
    for i, q in enumerate(Cc.sector):
        qb, rb = Q.block(Q.find(q)), R.block(R.find(q))
        np.testing.assert_allclose(qb @ rb, Cc.block(i), rtol=0, atol=1e-12 * np.abs(Cc.data).max())
        np.testing.assert_allclose(qb.T @ qb, np.eye(qb.shape[1]), rtol=0, atol=1e-12)
        assert np.abs(np.tril(rb, -1)).max(initial=0.0) == 0.0


@pytest.mark.parametrize("pcode", [xt.U1, xt.SU2, xt.FU1xU1], ids=["U1", "SU2", "FU1xU1"])
def test_storage_round_trip_with_data(rt, pcode, tmp_path):
    # xt_tensor_save / load: data and structure bit for bit, and the loaded tensor computes as the original
    r = random.Random(12)
    legs = [rand_entries(pcode, r, 3, 4) for _ in range(4)]
    o, t = both_tensor(rt, pcode, legs, 2, 33)
    u = t.permute(1, (2, 0, 3, 1))
    f = tmp_path / "u.xt"
    u.save(f)
    v = xt.Tensor.load(rt, f)
    assert structure_of(v) == structure_of(u)
    assert np.array_equal(v.download(), u.download())
    assert np.array_equal(v.permute(-1, (1, 3, 0, 2)).download(), u.permute(-1, (1, 3, 0, 2)).download())
    w = xt.Tensor.deserialize(rt, t.serialize())
    assert np.array_equal(w.download(), o.data)
