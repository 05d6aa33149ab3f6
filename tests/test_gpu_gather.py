"""The permute of contract()'s left operand read inside the product (the gather GEMM).  Needs an MI355X.

contract(a, a1, b, a2) permutes a and b, multiplies, and permutes the product (Tensor.hpp:328-351).
When every sector block of the product is a skinny FP64 item with the permuted a on its long side
(k <= 32, n <= 32, n < m) and the permute of a is single-source with coefficients +-1 (U(1),
fermionic U(1), U(1)xU(1)), the product reads a through the permute's tasks instead of
materialising the permuted copy (xh_gemm_gather).  Checks:

* fused and materialising paths give the same result bit for bit (the same skinny-kernel sums in
  the same order), and the fused path ran (its counter moved);
* the result matches the oracle's contract at 1e-12 of the largest entry;
* products that do not qualify (a wide right operand, SU(2) recoupling sums) take the materialising
  path and still match the oracle;
* CTMRG moves with and without the fusion give the same environment and energy bit for bit.

Each check runs with the gather alone (mode 1, the default) and with the product's own permute
folded into its stores as well (mode 2, xh_gemm_gather_scatter).
"""
import random

import numpy as np
import pytest

import xped_amd.tensor as xt
from oracle import tensor as oten
from oracle.symmetry import Sym
from xped_amd import configs

from .helpers import KINDS, both_basis, both_tensor, make_ipeps, oracle_structure, structure_of

pytestmark = pytest.mark.gpu

# the last contraction of a CTM corner: a rank-4 seven-leg tensor with three small legs (1, 3, 6)
# contracted against a five-leg right operand with two small open legs
A1 = [-1, 1, -4, 2, -5, -2, 3]
A2 = [1, 2, 3, -6, -3]
P1 = [0, 2, 4, 5, 1, 3, 6]  # get_permutations(A1, 4, A2, 3, 3).p1, shift 0


def _q(pcode, v):
    return (v, (v + 1) % 3 - 1) if pcode == xt.FU1xU1 else v


def _legs(pcode, big, small):
    b = [(_q(pcode, q), d) for q, d in big]
    s = [(_q(pcode, q), d) for q, d in small]
    return [b, s, b, s, b, b, s]


def _pair(rt, pcode, big, small, open_b, seed):
    oA, A = both_tensor(rt, pcode, _legs(pcode, big, small), 4, seed)
    oPA = oten.permute(oA, 0, P1)
    PA = A.permute(0, P1)
    # the right operand's domain is the permuted left operand's codomain
    pairs = [both_basis(pcode, [(_q(pcode, q), d) for q, d in open_b]) for _ in range(2)]
    oB = oten.Tensor(Sym(*KINDS[pcode]), oPA.unc_cod, [p[0] for p in pairs]).set_random(seed + 1, 0)
    B = xt.Tensor(rt, PA.uncoupledCodomain(), [p[1] for p in pairs])
    assert structure_of(B) == oracle_structure(oB)
    B.upload(oB.data)
    return oA, A, oB, B


def _fused_and_plain(rt, A, B, mode=1):
    n0 = rt.fuse_permute(mode)
    Cf = A.contract(B, A1, A2, 3)
    n1 = rt.fuse_permute(False)
    Cp = A.contract(B, A1, A2, 3)
    n2 = rt.fuse_permute(True)
    assert n2 == n1
    return Cf, Cp, n1 - n0


@pytest.mark.parametrize("mode", [1, 2], ids=["gather", "gather+scatter"])
@pytest.mark.parametrize("pcode", [xt.U1, xt.FU1, xt.FU1xU1], ids=["U1", "FU1", "FU1xU1"])
def test_gather_gemm_is_the_permute_then_the_product(rt, pcode, mode):
    # mode 2 also writes the product straight into its permuted layout (the result permute of
    # this contraction, (0, 3, 5 | 1, 2, 4), is not the identity)
    big = [(-2, 3), (0, 5), (2, 4)]
    small = [(0, 1), (1, 2)]
    oA, A, oB, B = _pair(rt, pcode, big, small, [(-1, 2), (1, 1)], 7)
    Cf, Cp, nf = _fused_and_plain(rt, A, B, mode)
    assert nf == 1, "the product did not run fused"
    assert structure_of(Cf) == structure_of(Cp)
    assert np.array_equal(Cf.download(), Cp.download())
    oC = oten.contract(oA, A1, oB, A2, 3)
    assert structure_of(Cf) == oracle_structure(oC)
    np.testing.assert_allclose(Cf.download(), oC.data, rtol=0, atol=1e-12 * np.abs(oC.data).max())
    # a second call reuses the cached plans
    rt.fuse_permute(mode)
    assert np.array_equal(A.contract(B, A1, A2, 3).download(), Cf.download())
    rt.fuse_permute(True)


@pytest.mark.parametrize("mode", [1, 2], ids=["gather", "gather+scatter"])
def test_gather_gemm_rows_across_many_domain_trees(rt, mode):
    # many small domain trees: the plan's row blocks meet up to four trees each, and blocks end
    # early where a fifth would begin (the first test's trees are longer than a block instead)
    r = random.Random(3)
    big = [(q, r.randint(1, 3)) for q in range(-3, 4)]
    small = [(-1, 2), (0, 1), (1, 1)]
    oA, A, oB, B = _pair(rt, xt.U1, big, small, [(-1, 1), (0, 2), (1, 1)], 11)
    Cf, Cp, nf = _fused_and_plain(rt, A, B, mode)
    assert nf == 1
    assert np.array_equal(Cf.download(), Cp.download())
    oC = oten.contract(oA, A1, oB, A2, 3)
    np.testing.assert_allclose(Cf.download(), oC.data, rtol=0, atol=1e-12 * np.abs(oC.data).max())


def test_products_that_do_not_qualify_are_materialised(rt):
    # open right-operand legs of dimension 40: n > 32, not a skinny product
    big = [(-1, 3), (1, 4)]
    small = [(0, 1), (1, 2)]
    oA, A, oB, B = _pair(rt, xt.U1, big, small, [(-1, 20), (1, 20)], 5)
    Cf, Cp, nf = _fused_and_plain(rt, A, B)
    assert nf == 0
    assert np.array_equal(Cf.download(), Cp.download())
    oC = oten.contract(oA, A1, oB, A2, 3)
    np.testing.assert_allclose(Cf.download(), oC.data, rtol=0, atol=1e-12 * np.abs(oC.data).max())
    # SU(2): the permute sums recoupled sources, so the product never reads through it
    big2 = [(1, 2), (2, 2), (3, 1)]
    small2 = [(1, 1), (2, 1)]
    oA, A, oB, B = _pair(rt, xt.SU2, big2, small2, [(1, 1), (2, 1)], 9)
    Cf, Cp, nf = _fused_and_plain(rt, A, B)
    assert nf == 0
    oC = oten.contract(oA, A1, oB, A2, 3)
    np.testing.assert_allclose(Cf.download(), oC.data, rtol=0, atol=1e-12 * np.abs(oC.data).max())


@pytest.mark.parametrize("mode", [1, 2], ids=["gather", "gather+scatter"])
def test_ctm_moves_with_and_without_the_fusion_agree_bit_for_bit(mode):
    def run(fuse):
        rt = xt.Runtime(0)
        rt.fuse_permute(fuse)
        _, A = make_ipeps(rt, configs.c4_small_virtual(), configs.c4_physical(), 5)
        c = xt.CTM(rt, A, 12).init()
        for _ in range(3):
            c.grow_all()
        c.computeRDM()
        env = [c.tensor(w, s).download() for w in ("C1", "C2", "C3", "C4", "T1", "T2", "T3", "T4") for s in range(4)]
        return c.energy(), env, rt.fuse_permute(None)

    Ef, envf, nf = run(mode)
    Ep, envp, np_ = run(False)
    assert nf > 0 and np_ == 0
    assert Ef == Ep
    assert len(envf) == len(envp)
    for x, y in zip(envf, envp):
        assert np.array_equal(x, y)
