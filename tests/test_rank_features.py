"""Input features of the learned rank predictor (xt_rank_features), restated from their definition.

Per sector, with decades scaled by 1/4 (s_1 >= s_2 >= ... its leading sketched singular values,
s_glob the largest over all sectors, c = 10^log_cut * s_glob the candidate cut):
f[0] = log10(s_1 / s_glob), f[j] = log10(s_{j+1} / s_1) for j = 1..15 (-16 where the sector has
fewer values), f[16] = log10(min(m, n)), f[17] = log_cut, f[18] = log10(c / s_last) clamped to
[-16, 16].  The library computes a sector's cut-independent features once and its two cut features
per cut; every cut must give what the definition gives for it.
"""
import numpy as np

from xped_amd import rank_predictor as rp


def _lg(x):
    return np.log10(x) if x > 0 else -300.0


def _restated(top, dim, s_glob, log_cut):
    t = list(top[: rp.TOP])
    n = len(t)
    s1 = t[0] if n else 0.0
    f = np.zeros(rp.NFEAT)
    f[0] = 0.25 * max(-16.0, _lg(s1) - _lg(s_glob))
    for j in range(1, rp.TOP):
        f[j] = 0.25 * (max(-16.0, _lg(t[j]) - _lg(s1)) if j < n else -16.0)
    f[16] = 0.25 * np.log10(max(1, dim))
    f[17] = 0.25 * log_cut
    slast = t[min(n, rp.TOP) - 1] if n else 0.0
    f[18] = 0.25 * max(-16.0, min(16.0, log_cut + _lg(s_glob) - _lg(slast)))
    return f.astype(np.float32)


def test_features_follow_their_definition_for_every_cut():
    r = np.random.default_rng(4)
    for trial in range(40):
        n = int(r.integers(0, 20))
        top = np.sort(10.0 ** r.uniform(-14, 1, n))[::-1]
        if trial % 7 == 3 and n > 2:
            top[-1] = 0.0  # an exactly zero value
        s_glob = max(float(top[0]) if n else 0.0, 10.0 ** r.uniform(-1, 1))
        dim = int(r.integers(1, 3000))
        for log_cut in (-14.0, -7.25, -0.5, 0.0, float(r.uniform(-14, 0))):
            got = rp.features(top, dim, s_glob, log_cut)
            want = _restated(top, dim, s_glob, log_cut)
            np.testing.assert_allclose(got, want, rtol=0, atol=1e-6)
            assert got[17] == np.float32(0.25 * log_cut)


def test_a_slip_in_the_last_feature_is_caught():
    # the restatement is not a copy of the library: with s_last swapped for s_1, f[18] differs
    top = np.array([3.0, 1.0, 1e-3])
    got = rp.features(top, 10, 3.0, -2.0)
    wrong = 0.25 * max(-16.0, min(16.0, -2.0 + np.log10(3.0) - np.log10(3.0)))
    assert abs(got[18] - wrong) > 0.1
