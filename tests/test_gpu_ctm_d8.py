"""The benchmark's CTMRG workload (U(1) iPEPS, 2x2 cell, D = 8, chi = 256) against the oracle's
FP64 run, committed as tests/golden/ctm_d8_chi256.json (written by make_ctm_fixture.py in the build
container: about two hours of CPU).  Needs an MI355X.

From FROM_A initialisation the first grow_all cuts chi from 64^2-dimensional corner products to
256.  The double layer's ket <-> bra symmetry makes +q/-q sector pairs of singular values exactly
degenerate, and at the cut last-bit rounding decides which partner of such a pair is kept.  The
run is therefore only piecewise continuous in its input: perturbing the site tensors by 1e-16
relative moves the HIP run from one branch to the other (energy after the first step moving by
4.2e-9 relative onto the oracle's value, spectra by 1.5e-7 sigma_max), while on one branch two runs
agree to 1e-14.  The test follows the decompositions in order: until one of them keeps the other
partner of a pair, its kept values must equal the oracle's as multisets to 1e-12 sigma_max and the
kept counts must agree; the decomposition that branches may differ from the oracle's counts only by
one +q/-q swap (values still at 1e-12); after it, values are compared at 1e-6 sigma_max.  The energy
is compared at 1e-10 relative on the oracle's branch and 1e-8 otherwise, the second step at 1e-8,
and the fast (randomized, learned-rank) second step against the HIP FP64 one at 1e-10."""
import json
from pathlib import Path

import numpy as np
import pytest

import xped_amd.tensor as xt
from xped_amd import configs, rank_predictor
from xped_amd.ipeps import random_ipeps

pytestmark = pytest.mark.gpu

FIX = json.loads((Path(__file__).parent / "golden" / "ctm_d8_chi256.json").read_text())


def _sites(rt):
    return [a.scale(10.0) for a in random_ipeps(rt, configs.c4_virtual(), configs.c4_physical(), 1)]


@pytest.mark.timeout(600)
def test_d8_chi256_steps_match_the_oracle_fixture(rt):
    A = _sites(rt)
    c = xt.CTM(rt, A, 256).set_options(record_spectra=True).init()
    E0 = c.step()
    sp = c.spectra()
    assert len(sp) == len(FIX["first_step_spectra"]) == 16
    branched = False
    for g, o in zip(sp, FIX["first_step_spectra"]):
        assert (g["site"], g["dir"]) == (o["site"], o["dir"])
        vals = sorted(x for s in g["sectors"] for x in np.sort(s["sv"])[::-1][: s["kept"]])
        assert len(vals) == len(o["sv"])
        np.testing.assert_allclose(vals, o["sv"], rtol=0, atol=(1e-6 if branched else 1e-12) * max(vals))
        kept = {str(s["q"]): s["kept"] for s in g["sectors"] if s["kept"] > 0}
        okept = {k: v for k, v in o["kept"].items() if v > 0}
        if kept != okept and not branched:
            diff = {q: kept.get(q, 0) - okept.get(q, 0) for q in set(kept) | set(okept)}
            moved = sorted((d, int(q)) for q, d in diff.items() if d)
            assert len(moved) == 2 and moved[0][0] == -1 and moved[1][0] == 1 and moved[0][1] == -moved[1][1], moved
            branched = True
    e0 = FIX["energies"][0]
    assert abs(E0 - e0) <= (1e-8 if branched else 1e-10) * abs(e0), (E0, e0, branched)
    # second step: HIP FP64 and the fast path (randomized SVD, learned ranks) from the same state
    fast = xt.CTM(rt, A, 256).init()
    fast.step()
    params, width = rank_predictor.load_weights()
    fast.set_options(precision="fp64", svd="randomized", oversample=16, power_iters=6)
    fast.set_rsvd(oversample=16, rank_factor=2.0, power_iters=6, sketch_precision="fp64")
    fast.set_predictor(params, width)
    E1 = c.step()
    E1f = fast.step()
    e1 = FIX["energies"][1]
    assert abs(E1 - e1) <= 1e-8 * abs(e1), (E1, e1)
    assert abs(E1f - E1) <= 1e-10 * abs(E1), (E1f, E1)
    assert fast.rsvd_stats()[1] == 0  # no fallback to the full SVD


@pytest.mark.timeout(900)
def test_chi384_fast_step_matches_fp64_step(rt):
    # C5's chi = 384 step: after a common full-SVD warm-up, the fast step (randomized SVD with
    # the learned ranks, trained at chi = 256) against the HIP FP64 step from the same state
    A = _sites(rt)
    runs = []
    for mode in ("fp64", "fast"):
        c = xt.CTM(rt, A, 384).init()
        c.grow_all()
        if mode == "fast":
            params, width = rank_predictor.load_weights()
            c.set_options(precision="fp64", svd="randomized", oversample=16, power_iters=6)
            c.set_rsvd(oversample=16, rank_factor=2.0, power_iters=6, sketch_precision="fp64")
            c.set_predictor(params, width)
        runs.append((c.step(), c.rsvd_stats()))
    (Ef, _), (Er, (n, nfb, _)) = runs
    assert n > 0 and nfb == 0
    assert abs(Er - Ef) <= 1e-10 * abs(Ef), (Er, Ef)
