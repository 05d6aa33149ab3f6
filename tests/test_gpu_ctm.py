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


def test_random_ipeps_step_matches_oracle(rt):
    oA, pA = make_ipeps(rt, configs.c4_small_virtual(), configs.c4_physical(), 7)
    chi = 9
    oc = octm.CTM(oA, chi)
    oc.init_from_A()
    c = xt.CTM(rt, pA, chi).init()
    for it in range(2):
        Eo = oc.step()
        E = c.step()
        assert abs(E - Eo) <= 1e-10 * abs(Eo), (it, E, Eo)
        for name, otab in (("C1", oc.C1), ("T1", oc.T1), ("C3", oc.C3), ("T4", oc.T4)):
            for site in range(4):
                assert structure_of(c.tensor(name, site)) == oracle_structure(otab[site]), (it, name, site)
        for site in range(4):
            for name, otab in (("rho_h", oc.rho_h), ("rho_v", oc.rho_v)):
                r = c.tensor(name, site)
                assert structure_of(r) == oracle_structure(otab[site])
                np.testing.assert_allclose(r.download(), otab[site].data, rtol=0, atol=1e-9)
