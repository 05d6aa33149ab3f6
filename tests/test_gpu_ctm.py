"""CTMRG step on the GPU (C++ driver over the HIP tensor layer) vs the oracle's FP64 CTMRG.

Energies and reduced density matrices are gauge invariant, so they are compared directly:
<= 1e-10 relative for the energy (the bar for the FP64 path), 1e-9 absolute for rho entries
(O(1) normalised); sector structure of every environment tensor and the kept count of every
projector SVD are compared exactly."""
import numpy as np
import pytest

import xped_amd.tensor as xt
from oracle import ctm as octm
from xped_amd import configs

from .helpers import make_ipeps, make_neel, oracle_structure, structure_of

pytestmark = pytest.mark.gpu


def test_neel_product_state_energy(rt):
    oA, pA = make_neel(rt)
    c = xt.CTM(rt, pA, 4).init()
    E = c.step()
    assert abs(E - (-0.5)) < 1e-14


def test_random_ipeps_steps_match_oracle(rt):
    # D = 2 and chi = 400: chi is not reached in two steps, so only the eps = 1e-10 rule truncates and
    # the kept states are unambiguous; energies, structures and density matrices must agree
    oA, pA = make_ipeps(rt, [((0,), 1), ((1,), 1)], configs.c4_physical(), 7)
    chi = 400
    oc = octm.CTM(oA, chi)
    oc.init_from_A()
    c = xt.CTM(rt, pA, chi).init()
    for it in range(2):
        Eo = oc.step()
        E = c.step()
        assert abs(E - Eo) <= 1e-10 * abs(Eo), (it, E, Eo)
        for name, otab in (("C1", oc.C1), ("T1", oc.T1), ("C3", oc.C3), ("T4", oc.T4), ("T2", oc.T2)):
            for site in range(4):
                assert structure_of(c.tensor(name, site)) == oracle_structure(otab[site]), (it, name, site)
        for site in range(4):
            for name, otab in (("rho_h", oc.rho_h), ("rho_v", oc.rho_v)):
                r = c.tensor(name, site)
                assert structure_of(r) == oracle_structure(otab[site])
                np.testing.assert_allclose(r.download(), otab[site].data, rtol=0, atol=1e-9)


def test_truncated_projectors_keep_the_same_singular_values(rt):
    # D = 3, chi = 9 truncates inside the exactly degenerate +q/-q pairs that the double-layer
    # environment produces; which partner survives depends on last-bit rounding (as it does in the
    # reference's LAPACK path), so the kept singular VALUES are compared as a multiset up to and
    # including the first call where the per-sector choice differs
    oA, pA = make_ipeps(rt, configs.c4_small_virtual(), configs.c4_physical(), 7)
    chi = 9
    oc = octm.CTM(oA, chi)
    oc.init_from_A()
    c = xt.CTM(rt, pA, chi).set_options(record_spectra=True).init()
    oc.grow_all()
    c.grow_all()
    sp = c.spectra()
    assert len(sp) == len(oc.spectra) == 16
    for g, o in zip(sp, oc.spectra):
        assert (g["site"], g["dir"]) == o[:2]
        kept_g = {s["q"]: s["kept"] for s in g["sectors"] if s["kept"] > 0}
        vals_g = sorted(x for s in g["sectors"] for x in s["sv"][: s["kept"]])
        assert sum(kept_g.values()) == len(o[3])
        smax = max(vals_g)
        np.testing.assert_allclose(vals_g, o[3], rtol=0, atol=1e-12 * smax)
        if kept_g != {q[0]: n for q, n in o[2].items()}:
            break


def test_randomized_svd_step_matches_full_step(rt):
    # the second step with randomized projector SVDs (target ranks = kept counts of the first,
    # full step; unsafe sketches fall back to the full SVD) against the same step with full SVDs
    _, A = make_ipeps(rt, configs.c4_small_virtual(), configs.c4_physical(), 7)
    chi = 24
    full = xt.CTM(rt, A, chi).init()
    rand = xt.CTM(rt, A, chi).init()
    full.grow_all()
    rand.grow_all()
    rand.set_options(svd="randomized", oversample=8, power_iters=2)
    Ef = full.step()
    Er = rand.step()
    n, nfb, ratio = rand.rsvd_stats()
    assert n == 16
    assert abs(Er - Ef) <= 1e-10 * abs(Ef), (Er, Ef, nfb, ratio)


def test_batched_move_contractions_give_the_same_steps(rt):
    # the independent contractions of a move in lockstep, each stage's products in one grouped GEMM
    # launch (mul_batch: offsets from the lowest data addresses of several tensors), against one
    # contraction at a time.  The split-K chunking of a shared launch differs, so results differ in
    # the last bits; energies must agree to 1e-10 while every decomposition keeps the same counts,
    # and a first difference in kept counts may only be a +q/-q partner swap at the cut
    _, A = make_ipeps(rt, configs.c4_small_virtual(), configs.c4_physical(), 7)
    chi = 24
    runs = []
    for batch in (False, True):
        c = xt.CTM(rt, A, chi).set_options(batch_contractions=batch, record_spectra=True).init()
        E = [c.step() for _ in range(3)]
        kept = [{s["q"]: s["kept"] for s in r["sectors"] if s["kept"] > 0} for r in c.spectra()]
        runs.append((E, kept))
    (E0, k0), (E1, k1) = runs
    assert len(k0) == len(k1) == 48
    first = next((i for i, (a, b) in enumerate(zip(k0, k1)) if a != b), None)
    if first is not None:
        diff = {q: k1[first].get(q, 0) - k0[first].get(q, 0) for q in set(k0[first]) | set(k1[first])}
        moved = sorted((d, q) for q, d in diff.items() if d)
        assert len(moved) == 2 and moved[0][0] == -1 and moved[1][0] == 1 and moved[0][1] == -moved[1][1], moved
    for step, (a, b) in enumerate(zip(E0, E1)):
        same = first is None or first >= 16 * (step + 1)
        assert abs(a - b) <= (1e-10 if same else 1e-6) * abs(a), (step, E0, E1, first)


def test_bounded_plan_cache_gives_the_same_steps():
    # a runtime whose plan caches hold at most 8 plans drops and rebuilds plans all through the
    # steps (device tables freed in stream order while launches may still be queued); energies must
    # equal those of a runtime with the default bound bit for bit
    energies = []
    evicted = []
    for bound in (None, 8):
        rt = xt.Runtime(0)
        if bound:
            rt.set_max_plans(bound)
        oA, pA = make_ipeps(rt, [((0,), 1), ((1,), 1)], configs.c4_physical(), 7)
        c = xt.CTM(rt, pA, 16).init()
        energies.append([c.step() for _ in range(3)])
        evicted.append(rt.cache_sizes()[2])
        assert rt.cache_sizes()[0] <= (bound or 2048)
    assert energies[0] == energies[1]
    assert evicted[0] == 0 and evicted[1] > 0


def _hip_bond_energies(c, oc):
    """Bond energies of the HIP CTM's density matrices (the oracle's gate on HIP data; the
    oracle CTM run alongside supplies the structures, which must match)."""
    from oracle import ctm as octm_
    from oracle import tensor as ot

    out = []
    for name, otab in (("rho_h", oc.rho_h), ("rho_v", oc.rho_v)):
        for site in range(4):
            r = c.tensor(name, site)
            o = otab[site].copy()
            assert structure_of(r) == oracle_structure(o)
            o.data = r.download()
            out.append(ot.trace(octm_.ct(octm_.tw(o, [0, 1]), [1, 2, 3, 4], oc._gate(o, 1.0, 1.0), [3, 4, 1, 2], 0)))
    return out


@pytest.mark.parametrize("orientation", ["h", "v"])
@pytest.mark.parametrize("alpha,beta", [(-1.0, 1.0), (0.8, 0.3)], ids=["singlet", "magnetised"])
def test_dimer_states_keep_exact_bond_energies(rt, orientation, alpha, beta):
    # closed forms (helpers.dimer_bond_energies) through three HIP CTMRG steps
    from .helpers import dimer_bond_energies, make_dimer

    oA, pA = make_dimer(rt, orientation, alpha, beta)
    oc = octm.CTM(oA, 16)
    oc.init_from_A()
    c = xt.CTM(rt, pA, 16).init()
    want = np.array(dimer_bond_energies(orientation, alpha, beta))
    for _ in range(3):
        oc.step()
        E = c.step()
        assert abs(E - want.sum() / 4) <= 1e-12
        assert np.abs(np.array(_hip_bond_energies(c, oc)) - want).max() <= 1e-12


def test_truncating_d4_steps_match_oracle(rt):
    # D = 4 over charges 0, 1, 2 and chi = 30: every projector truncates.  The double layer is
    # symmetric under exchanging ket and bra, which maps the fused charge q to -q, so +q/-q sectors
    # carry exactly degenerate singular values and a cut through such a pair keeps either partner
    # (last-bit rounding decides, in the reference as here).  Compared: energies to 1e-10 relative,
    # the kept singular values of every decomposition as a multiset to 1e-12 sigma_max, and the
    # density matrices to 1e-9 in every step whose per-sector kept counts all agree.
    virtual = [((0,), 2), ((1,), 1), ((2,), 1)]
    oA, pA = make_ipeps(rt, virtual, configs.c4_physical(), 11)
    chi = 30
    oc = octm.CTM(oA, chi)
    oc.init_from_A()
    c = xt.CTM(rt, pA, chi).set_options(record_spectra=True).init()
    same_counts = 0
    for it in range(2):
        oc.spectra = []
        c.clear_spectra()
        Eo = oc.step()
        E = c.step()
        assert abs(E - Eo) <= 1e-10 * abs(Eo), (it, E, Eo)
        sp = c.spectra()
        assert len(sp) == len(oc.spectra) == 16
        all_same = True
        for g, o in zip(sp, oc.spectra):
            assert (g["site"], g["dir"]) == o[:2]
            kept_g = {s["q"]: s["kept"] for s in g["sectors"] if s["kept"] > 0}
            vals_g = sorted(x for s in g["sectors"] for x in np.sort(s["sv"])[::-1][: s["kept"]])
            assert len(vals_g) == len(o[3]) == chi
            np.testing.assert_allclose(vals_g, o[3], rtol=0, atol=1e-12 * max(vals_g))
            all_same &= kept_g == {q[0]: n for q, n in o[2].items() if n > 0}
        if all_same:
            same_counts += 1
            for site in range(4):
                for name, otab in (("rho_h", oc.rho_h), ("rho_v", oc.rho_v)):
                    r = c.tensor(name, site)
                    assert structure_of(r) == oracle_structure(otab[site])
                    np.testing.assert_allclose(r.download(), otab[site].data, rtol=0, atol=1e-9)
