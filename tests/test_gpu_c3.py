"""C3 at its stated leg dimension (needs an MI355X): the U(1)xU(1) fermionic grouped contraction
A(l1, l2 | l3, l4) . B(l3, l4 | l5, l6) with every leg of dimension 512 (sector degeneracies from
the seeded power law), 313 matched sector blocks up to 3591 x 3172, 1.15 TFLOP.  The product is
one grouped launch per tiling, chosen by the autotuner, compared with the oracle's per-block FP64
products at 1e-12 relative, and replayed from a captured HIP graph bit for bit."""
import numpy as np
import pytest

import xped_amd.tensor as xt
from oracle import tensor as oten
from oracle.structure import Qbasis as OQ
from oracle.symmetry import FU1, U1, Sym
from xped_amd import configs
from xped_amd._lib import lib

from .helpers import oracle_structure, structure_of

pytestmark = pytest.mark.gpu


@pytest.mark.timeout(600)
def test_c3_leg512_matches_oracle_autotuned_and_replayed(rt):
    s = Sym(FU1, U1)
    ol, pl = [], []
    for i in range(6):
        e = configs.c3_leg(i, 512)
        o, p = OQ(s, 1), xt.Qbasis(xt.FU1xU1)
        for q, d in e:
            o.push_back(q, d)
            p.push_back(q, d)
        ol.append(o)
        pl.append(p)
    oA = oten.Tensor(s, ol[0:2], ol[2:4]).set_random(1)
    oB = oten.Tensor(s, ol[2:4], ol[4:6]).set_random(2)
    A = xt.Tensor(rt, pl[0:2], pl[2:4]).setRandom(1)
    B = xt.Tensor(rt, pl[2:4], pl[4:6]).setRandom(2)
    L = lib()
    L.xh_set_gemm_autotune(rt.ctx(), 1)
    try:
        C = A.mul(B)  # first call of the plan: both tilings timed, the faster one run last
        oC = oten.mul(oA, oB)
        assert structure_of(C) == oracle_structure(oC)
        got = C.download()
        assert np.abs(got - oC.data).max() <= 1e-12 * np.abs(oC.data).max()
        # replay: the product captured into a graph (no split-K at this size) equals the eager one
        out = C.zeros_like()
        rt.synchronize()
        rt.graph_begin()
        A.mul_into(B, out)
        g = rt.graph_end()
        for _ in range(2):
            g.launch()
            rt.synchronize()
            assert np.array_equal(out.download(), got)
        g.close()
    finally:
        L.xh_set_gemm_autotune(rt.ctx(), 0)
