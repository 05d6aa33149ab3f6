"""The learned rank predictor: hand-written HIP forward/backward/Adam kernels against the float64
oracle, training on the GPU from the exact FP64 CTMRG spectra, and what the trained network gains
over the rank guess without it (needs an MI355X).

Tolerances: FP32 kernels against the float64 oracle on the same FP32 inputs: forward 1e-5 and
gradients 1e-4 relative to the largest entry; one Adam step 1e-6 absolute.
"""
import numpy as np
import pytest

import xped_amd.tensor as xt
from oracle import mlp as om
from xped_amd import configs, rank_predictor as rp

from .helpers import make_ipeps

pytestmark = pytest.mark.gpu


def _batch(n, seed):
    r = np.random.default_rng(seed)
    return r.standard_normal((n, rp.NFEAT)).astype(np.float32), r.standard_normal(n).astype(np.float32)


def test_forward_and_gradient_match_oracle(rt):
    net = rp.RankNet(rt, seed=11)
    X, T = _batch(777, 1)
    p = net.params().astype(np.float64)
    y = net.forward(X)
    yo, _ = om.forward(p, X.astype(np.float64), rp.NFEAT, rp.WIDTH)
    np.testing.assert_allclose(y, yo, rtol=0, atol=1e-5 * np.abs(yo).max())
    loss, g = net.grad(X, T)
    lo, go = om.loss_and_grad(p, X.astype(np.float64), T.astype(np.float64), rp.NFEAT, rp.WIDTH)
    assert abs(loss - lo) <= 1e-5 * lo
    np.testing.assert_allclose(g, go, rtol=0, atol=1e-4 * np.abs(go).max())


def test_adam_step_matches_oracle(rt):
    net = rp.RankNet(rt, seed=12)
    X, T = _batch(300, 2)
    p0 = net.params().astype(np.float64)
    # the update rule is judged on the kernels' own gradient (the first Adam step is about
    # lr * sign(g), so gradients at rounding level would decide the sign otherwise)
    _, g = net.grad(X, T)
    net.train(X, T, steps=1, lr=1e-2)
    p1, _, _ = om.adam_step(p0, g.astype(np.float64), np.zeros_like(p0), np.zeros_like(p0), 1, 1e-2)
    np.testing.assert_allclose(net.params(), p1, rtol=0, atol=1e-6)


def test_training_on_gpu_and_gain_over_proportional_ranks(rt):
    # train on the spectra of the first recorded CTM step, judge on the second one
    recs = rp.records_from_npz(rp.SPECTRA)
    train = [r for r in recs if r["key"][0] == 0]
    held = [r for r in recs if r["key"][0] == 1]
    X, T = rp.dataset(train, 16)
    net = rp.RankNet(rt, seed=3)
    losses = net.train(X, T, steps=1500, lr=3e-3)
    assert losses[-1] < 0.05 * losses[0]
    chi = 256
    learned, proportional, exact = [], [], []
    for r in held:
        tops = [sv[:rp.TOP] for sv in r["sv"]]
        dims = [min(int(a), int(b)) for a, b in zip(r["rows"], r["cols"])]
        learned.append(rp.sketch_cost_and_safety(r, net.ranks(tops, dims, chi)))
        proportional.append(rp.sketch_cost_and_safety(r, rp.proportional_ranks(r, chi)))
        exact.append(rp.sketch_cost_and_safety(r, r["kept"]))
    n_learned = sum(s for _, s in learned)
    n_prop = sum(s for _, s in proportional)
    print(f"safe sketches: learned {n_learned}/{len(held)}, proportional {n_prop}/{len(held)}; "
          f"cost learned/exact {np.mean([c for c, _ in learned]) / np.mean([c for c, _ in exact]):.2f}")
    # the gain: with the predictor every held-out decomposition passes the safety check (no full
    # SVD fallback) at a sketch cost close to that of the exact kept counts; sharing chi out in
    # proportion to the sector sizes needs the full SVD almost everywhere
    assert n_learned >= len(held) - 1
    assert n_learned > n_prop
    assert np.mean([c for c, _ in learned]) <= 1.6 * np.mean([c for c, _ in exact])


def test_ctm_with_learned_ranks_matches_full_svd_step(rt):
    # the committed weights drive the randomized SVDs from the very first step (no full-SVD step
    # to take ranks from); unsafe sketches fall back to the full SVD, so the step must match
    params, width = rp.load_weights()
    _, A = make_ipeps(rt, configs.c4_small_virtual(), configs.c4_physical(), 7)
    chi = 24
    full = xt.CTM(rt, A, chi).init()
    fast = xt.CTM(rt, A, chi).init()
    full.grow_all()
    fast.grow_all()
    fast.set_options(svd="randomized", oversample=16, power_iters=6).set_predictor(params, width)
    Ef = full.step()
    Er = fast.step()
    n, nfb, _ = fast.rsvd_stats()
    assert n == 16
    assert abs(Er - Ef) <= 1e-10 * abs(Ef), (Er, Ef, nfb)
