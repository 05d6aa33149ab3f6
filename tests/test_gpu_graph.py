"""HIP-graph capture of grouped launches (needs an MI355X).

A contraction's grouped GEMM is captured once (`mul_into` on existing tensors) and replayed.
Between replays the test does eager work and synchronises. Every replay must equal the eager
product bit for bit: the same kernels run on the same buffers. An operand changed between replays
must be read by the next replay. Work that cannot be captured (a split-K plan, which allocates its
partials; a plan not built yet) must fail the capture cleanly and leave the stream usable.
"""
import numpy as np
import pytest

import xped_amd.tensor as xt
from xped_amd._lib import XpedError

pytestmark = pytest.mark.gpu


def _basis(dims):
    b = xt.Qbasis(xt.U1)
    for q, d in enumerate(dims):
        b.push_back((q,), d)
    return b


def _pair(rt, m, k, n, seed):
    A = xt.Tensor(rt, [_basis(m)], [_basis(k)])
    B = xt.Tensor(rt, [_basis(k)], [_basis(n)])
    A.setRandom(seed, 0)
    B.setRandom(seed + 1, 0)
    return A, B


def test_captured_grouped_gemm_replays_match_eager(rt):
    # inner dimensions <= 96 cannot be split along K (chunks of >= 128), so no workspace is needed
    A, B = _pair(rt, (1536, 512, 96), (96, 64, 40), (1536, 512, 80), 11)
    ref = A.mul(B).download()  # builds the plan
    out = A.mul(B).zeros_like()
    rt.synchronize()
    rt.graph_begin()
    A.mul_into(B, out)
    g = rt.graph_end()
    rng = np.random.default_rng(3)
    for _ in range(3):
        out.upload(rng.uniform(-1, 1, out.size))  # overwritten by the replay
        g.launch()
        rt.synchronize()
        assert np.array_equal(out.download(), ref)
        assert np.array_equal(A.mul(B).download(), ref)  # eager work between replays
    # the replay reads the operands' buffers as they are at launch time
    A.setRandom(77, 0)
    ref2 = A.mul(B).download()
    g.launch()
    rt.synchronize()
    assert np.array_equal(out.download(), ref2)
    g.close()


def test_split_k_plan_is_refused_under_capture(rt):
    A, B = _pair(rt, (64,), (4096,), (64,), 21)
    ref = A.mul(B).download()  # split-K plan, built eagerly
    out = A.mul(B).zeros_like()
    rt.synchronize()
    rt.graph_begin()
    with pytest.raises(XpedError):
        A.mul_into(B, out)
    with pytest.raises(XpedError):
        rt.graph_end()
    assert np.array_equal(A.mul(B).download(), ref)


def test_plan_creation_is_refused_under_capture(rt):
    A, B = _pair(rt, (128, 64), (32, 16), (96, 48), 31)
    out = xt.Tensor(rt, [_basis((128, 64))], [_basis((96, 48))])
    rt.synchronize()
    rt.graph_begin()
    with pytest.raises(XpedError):
        A.mul_into(B, out)  # the product's plan does not exist yet
    with pytest.raises(XpedError):
        rt.graph_end()
    want = A.download().reshape(-1)  # the stream works eagerly again
    assert want.size == A.size
    np.testing.assert_allclose(A.mul_into(B, out).download(), A.mul(B).download(), rtol=0, atol=0)


def test_graph_keeps_its_plans_after_cache_eviction():
    # an own runtime with an 8-plan bound, so that the captured product's plan is dropped from the
    # cache while the graph is alive; the graph must keep the plan's device tables until it is
    # destroyed, however much memory later plans take from the pool
    rt2 = xt.Runtime(0)
    rt2.set_max_plans(8)
    A, B = _pair(rt2, (1536, 512, 96), (96, 64, 40), (1536, 512, 80), 41)
    ref = A.mul(B).download()
    out = A.mul(B).zeros_like()
    rt2.synchronize()
    rt2.graph_begin()
    A.mul_into(B, out)
    g = rt2.graph_end()
    evicted0 = rt2.cache_sizes()[2]
    keep = []
    for i in range(24):  # new plans of other shapes push the captured one out of the cache
        X, Y = _pair(rt2, (40 + i, 3), (17, 2), (30 + i, 5), 100 + i)
        keep.append(X.mul(Y))
    rt2.synchronize()
    assert rt2.cache_sizes()[2] > evicted0
    assert rt2.cache_sizes()[1] <= 8
    rng = np.random.default_rng(5)
    for _ in range(3):
        out.upload(rng.uniform(-1, 1, out.size))
        g.launch()
        rt2.synchronize()
        assert np.array_equal(out.download(), ref)
    g.close()
    # the plan was freed with the graph; the product plans anew and still agrees
    assert np.array_equal(A.mul(B).download(), ref)
