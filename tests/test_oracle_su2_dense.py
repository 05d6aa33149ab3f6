"""Pins of the oracle's SU(2) recoupling and fermionic swap signs by dense expansion (no GPU).

The reference's own checks (tests/tensor_tests.hpp:1-200, tests/tensor.cpp:75-384) expand a
symmetric tensor into a dense one with Clebsch-Gordan tensors (Tensor::plainTensor,
Tensor.cpp:1133-1310; FusionTree::asTensor, FusionTree.cpp:154-307) and require
plain(T.permute<shift, p>()) == shuffle_p(plain(T)) for every permutation and shift, and closed-form
spin matrices for chains of permutes and products.  The Clebsch-Gordan tensors come from the exact
3j symbols that the committed sympy fixtures pin, and the dense shuffle is numpy's transpose, so
these checks are independent of the recoupling formulas they test: a sign or index slip in
coeff_swap / coeff_recouple / coeff_turn / coeff_FS fails them (test_planted_mistakes_fail)."""
import itertools
import math

import numpy as np
import pytest

from oracle import symmetry as osym
from oracle import tensor as oten
from oracle.symmetry import FU1, SU2, U1, Sym

from .su2_cases import SU2_BASES, OracleOps, graded_transpose, obasis, sorted_legs, spin_chain_checks

TOL = 1e-13


def _su2_tensor(entries, rank, seed):
    s = Sym(SU2)
    legs = [obasis(s, e) for e in entries]
    return oten.Tensor(s, legs[:rank], legs[rank:]).set_random(seed)


def _worst_permute_error(t, shifts, perms):
    P = oten.plain(t)
    worst = 0.0
    for shift in shifts:
        for p in perms:
            worst = max(worst, np.abs(oten.plain(oten.permute(t, shift, p)) - np.transpose(P, p)).max())
    return worst


def test_cg_tensors_are_unitary_and_couple_spin_halves():
    s = Sym(SU2)
    # 1/2 x 1/2 -> 0 is the singlet (|up down> - |down up>)/sqrt(2), -> 1 the triplet
    c0 = oten.cgc(s, (2,), (2,), (1,))[:, :, 0]
    np.testing.assert_allclose(c0, np.array([[0, 1], [-1, 0]]) / math.sqrt(2), atol=1e-15)
    c1 = oten.cgc(s, (2,), (2,), (3,))
    np.testing.assert_allclose(c1[:, :, 0], np.array([[1, 0], [0, 0]]), atol=1e-15)
    np.testing.assert_allclose(c1[:, :, 1], np.array([[0, 1], [1, 0]]) / math.sqrt(2), atol=1e-15)
    # the vertex tensors of all outputs of a q1 x q2 product form an orthogonal matrix
    for q1, q2 in [(2, 3), (3, 3), (4, 2), (5, 4)]:
        cols = [oten.cgc(s, (q1,), (q2,), q).reshape(q1 * q2, q[0]) for q in s.combine((q1,), (q2,))]
        U = np.concatenate(cols, axis=1)
        np.testing.assert_allclose(U.T @ U, np.eye(q1 * q2), atol=1e-14)


@pytest.mark.parametrize("b,c", [(0, 1), (1, 2), (2, 0)])
def test_su2_permute_matches_dense_shuffle(b, c):
    # tests/tensor_tests.hpp:1-48: Tensor<2,2>({B, C}, {B, C}), all 24 permutations, shifts -2..2
    t = _su2_tensor([SU2_BASES[b], SU2_BASES[c], SU2_BASES[b], SU2_BASES[c]], 2, 5 + b)
    assert _worst_permute_error(t, range(-2, 3), list(itertools.permutations(range(4)))) <= TOL


@pytest.mark.parametrize("rank", [4, 0])
def test_su2_permute_within_domain_or_codomain(rank):
    # tests/tensor_tests.hpp:115-200 (Tensor<4,0> and Tensor<0,4>, shift 0)
    t = _su2_tensor([SU2_BASES[0]] * 4, rank, 9)
    assert _worst_permute_error(t, [0], list(itertools.permutations(range(4)))) <= TOL


def test_su2_rank3_all_shifts():
    t = _su2_tensor([SU2_BASES[0], SU2_BASES[2], SU2_BASES[1]], 2, 13)
    assert _worst_permute_error(t, range(-1, 3), list(itertools.permutations(range(3)))) <= TOL


@pytest.mark.parametrize("two_s1", range(1, 11))
def test_su2_spin_matrix_closed_forms(two_s1):
    # tests/tensor.cpp:230-384: spin matrices, S^2, S/sqrt(2), rank-2 products and S1.S2
    assert spin_chain_checks(OracleOps(), two_s1) <= 1e-12


def test_planted_mistakes_fail():
    """Each plausible slip in the recoupling coefficients breaks dense equivalence."""
    t = _su2_tensor([SU2_BASES[0], SU2_BASES[1], SU2_BASES[0], SU2_BASES[1]], 2, 3)
    perms = list(itertools.permutations(range(4)))
    assert _worst_permute_error(t, range(-2, 3), perms) <= TOL
    S = osym.Sym
    saved = (S._swap1, S._recouple1, S._turn1, S.coeff_FS)
    slips = {
        "swap sign": ("_swap1", lambda self, k, ql, qr, qf: -saved[0](self, k, ql, qr, qf) if k == SU2 and ql != qr else saved[0](self, k, ql, qr, qf)),
        "recouple phase": ("_recouple1", lambda self, k, q1, q2, q3, Q, Q12, Q23: saved[1](self, k, q1, q2, q3, Q, Q12, Q23) * (-1.0 if Q12 != Q23 else 1.0)),
        "turn factor": ("_turn1", lambda self, k, ql, qr, qf: saved[2](self, k, ql, qr, qf) * (math.sqrt(ql / qf) if k == SU2 else 1.0)),
        "FS sign": ("coeff_FS", lambda self, q: 1.0),
    }
    try:
        for name, (attr, f) in slips.items():
            setattr(S, attr, f)
            assert _worst_permute_error(t, range(-2, 3), perms) > 1e-3, name
            setattr(S, attr, saved[["_swap1", "_recouple1", "_turn1", "coeff_FS"].index(attr)])
    finally:
        S._swap1, S._recouple1, S._turn1, S.coeff_FS = saved


@pytest.mark.parametrize("kinds", [(FU1, U1), (FU1,)], ids=["FU1xU1", "FU1"])
def test_fermionic_split_permutes_are_graded_shuffles(kinds):
    """Shift-0 permutations that keep domain and codomain legs apart: the dense tensor is shuffled
    with the sign (-1)^(parity_a parity_b) for every pair of legs the permutation reorders (the
    coeff_swap of U1.hpp:174-182 composed along Permutation::decompose).  Bends of fermionic legs
    have no dense counterpart in the reference's tests and are not covered here."""
    s = Sym(*kinds)
    ent = [((1, 0), 2), ((0, 0), 1), ((-1, 0), 2), ((1, 1), 1), ((-1, -1), 1)] if len(kinds) == 2 else [((1,), 2), ((0,), 1), ((-1,), 2), ((2,), 1)]
    legs = [obasis(s, ent) for _ in range(4)]
    for rank in (4, 2, 0):
        t = oten.Tensor(s, legs[:rank], legs[rank:]).set_random(21 + rank)
        P, L = oten.plain(t), sorted_legs(t)
        for p in itertools.permutations(range(4)):
            if rank == 2 and (set(p[:2]) != {0, 1}):
                continue
            got = oten.plain(oten.permute(t, 0, list(p)))
            assert np.array_equal(got, graded_transpose(s, L, P, list(p))), p
        if rank == 4:
            # a dropped sign would make the plain shuffle agree instead
            got = oten.plain(oten.permute(t, 0, [1, 0, 2, 3]))
            assert not np.array_equal(got, np.transpose(P, [1, 0, 2, 3]))


def test_c2_reduced_contraction_orders_agree_in_the_oracle():
    """The identity the workload-size C2 GPU test relies on, on the oracle: absorbing a site into
    T in two contraction orders gives the same tensor."""
    from xped_amd import configs
    from xped_amd import su2_layer as L

    from .c2_ops import OracleOps

    s = Sym(SU2)
    ops = OracleOps(s)
    v, p, env = configs.c2_small_virtual(), configs.c2_physical(), configs.c2_env(9)
    A = oten.Tensor(s, [obasis(s, v), obasis(s, v)], [obasis(s, v), obasis(s, v), obasis(s, p)]).set_random(7, 1)
    T = oten.Tensor(s, [obasis(s, env)], [obasis(s, env), obasis(s, v), obasis(s, v)]).set_random(8, 1)
    Ad = L.adag_of(ops, A)
    X, Y = L.absorb(ops, T, A, Ad), L.absorb_other_order(ops, T, A, Ad)
    assert X.sector == Y.sector
    assert np.abs(X.data - Y.data).max() <= 1e-12 * np.abs(X.data).max()
    # the double layer's dense expansion equals the dense ket-bra contraction joined by the dense
    # expansions of the fusion isometries (spin SU(2) is bosonic, so the twists are identities)
    a = L.contract_aadag(ops, A, Ad)
    l, t, r, b = A.unc_dom[0], A.unc_dom[1], A.unc_cod[0], A.unc_cod[1]
    lA, tA, rA, bA = Ad.unc_dom[0], Ad.unc_dom[1], Ad.unc_cod[0], Ad.unc_cod[1]
    F = [oten.plain(ops.identity([ops.fuse(l, lA)], [l, lA])), oten.plain(ops.identity([ops.fuse(t, tA)], [t, tA])),
         oten.plain(ops.identity([r, rA], [ops.fuse(r, rA)])), oten.plain(ops.identity([b, bA], [ops.fuse(b, bA)]))]
    want = np.einsum("ltrbs,LTsRB,XlL,YtT,rRZ,bBW->XYZW", oten.plain(A), oten.plain(Ad), *F)
    np.testing.assert_allclose(oten.plain(a), want, rtol=0, atol=1e-13 * np.abs(want).max())
