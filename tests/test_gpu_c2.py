"""C2, the SU(2) double layer (needs an MI355X): contractAAdag (PEPSContractions.cpp:106-129) and
the absorption of a site into an environment tensor T(chi | chi, D, D') on the HIP path.

* Reduced size (virtual spins {0, 1/2}, chi = 9): against the oracle, structures bit-exact and
  data to 1e-13 relative to the largest entry (SU(2) recoupling sums).
* Workload size (virtual {0: 2, 1/2: 1, 1: 1}, chi = 100): the oracle's Python recoupling takes
  hours on these rank-8 intermediates, so the HIP result is checked by contracting the same
  network in a second order (ket and bra first, over the physical leg): different permutes,
  recoupling coefficients and GEMM plans must give the same tensor to 1e-12 relative; the
  oracle itself satisfies this identity (tests/test_oracle_su2_dense.py)."""
import numpy as np
import pytest

import xped_amd.tensor as xt
from oracle import tensor as oten
from oracle.structure import Qbasis
from oracle.symmetry import SU2, Sym
from xped_amd import configs
from xped_amd import su2_layer as L

from .c2_ops import OracleOps
from .helpers import both_basis, oracle_structure, structure_of

pytestmark = pytest.mark.gpu


def _both(rt, legs, rank, seed):
    pairs = [both_basis(xt.SU2, e) for e in legs]
    o = oten.Tensor(Sym(SU2), [p[0] for p in pairs[:rank]], [p[0] for p in pairs[rank:]]).set_random(seed, 1)
    t = xt.Tensor(rt, [p[1] for p in pairs[:rank]], [p[1] for p in pairs[rank:]])
    t.upload(o.data)
    return o, t


def test_c2_reduced_matches_oracle(rt):
    v, p, env = configs.c2_small_virtual(), configs.c2_physical(), configs.c2_env(9)
    oA, A = _both(rt, [v, v, v, v, p], 2, 7)
    oT, T = _both(rt, [env, env, v, v], 1, 8)
    oo, ho = OracleOps(Sym(SU2)), L.HipOps(rt)
    oAd, Ad = L.adag_of(oo, oA), L.adag_of(ho, A)
    for f in (lambda ops, a, ad: L.contract_aadag(ops, a, ad),
              lambda ops, a, ad: L.absorb(ops, oT if ops is oo else T, a, ad)):
        want, got = f(oo, oA, oAd), f(ho, A, Ad)
        assert structure_of(got) == oracle_structure(want)
        np.testing.assert_allclose(got.download(), want.data, rtol=0, atol=1e-13 * np.abs(want.data).max())


def test_c2_workload_two_contraction_orders_agree(rt):
    v, p, env = configs.c2_virtual(), configs.c2_physical(), configs.c2_env(100)
    A = xt.Tensor(rt, [L_basis(v), L_basis(v)], [L_basis(v), L_basis(v), L_basis(p)]).setRandom(7, 1)
    T = xt.Tensor(rt, [L_basis(env)], [L_basis(env), L_basis(v), L_basis(v)]).setRandom(8, 1)
    ops = L.HipOps(rt)
    Ad = L.adag_of(ops, A)
    a = L.contract_aadag(ops, A, Ad)
    assert a.size > 0
    X = L.absorb(ops, T, A, Ad)
    Y = L.absorb_other_order(ops, T, A, Ad)
    assert structure_of(X) == structure_of(Y)
    x, y = X.download(), Y.download()
    assert np.abs(x - y).max() <= 1e-12 * np.abs(x).max()
    # the environment leg really has full dimension 100 and the site legs 7
    assert T.leg(0).fullDim() == 100 and A.leg(0).fullDim() == 7


def L_basis(entries):
    b = xt.Qbasis(xt.SU2)
    for q, d in entries:
        b.push_back(q, d)
    return b
