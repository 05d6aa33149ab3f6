"""Pins of the oracle's CTMRG (oracle/ctm.py) beyond the D = 1 product state (no GPU).

* Dense networks: at FROM_A initialisation every environment tensor is one site's double layer
  with its outward legs traced, so each reduced density matrix is a finite traced-boundary patch.
  A symmetry-free dense contraction written from the lattice geometry (tests/ctm_dense.py) gives
  the 8 bond energies of random D = 2 and D = 3 states; they must equal the oracle's to 1e-12.
  This pins the init contractions, computeRDM and the Heisenberg gate.
* Closed forms through the moves: columnar dimer product states (D = 2 on the dimer bonds) keep
  their exact bond energies through CTMRG steps (finite correlations), in both orientations, for
  the singlet and for an unequal superposition whose magnetised sites make the non-dimer bonds
  nonzero (helpers.dimer_bond_energies).  A wrong site, leg or projector in a move changes the
  pattern or breaks the structure.
* Planted mistakes (one index-list or site slip each, in a copy of the oracle's source) fail at
  least one of these checks."""
import types
from pathlib import Path

import numpy as np
import pytest

from oracle import ctm as octm
from oracle import tensor as ot
from xped_amd import configs

from .ctm_dense import DenseCTM
from .helpers import dimer_bond_energies, make_dimer, oracle_bond_energies, oracle_ipeps


def _dense_vs_oracle(mod, virtual, seed):
    oA = oracle_ipeps(virtual, configs.c4_physical(), seed)
    oc = mod.CTM(oA, 400)
    oc.init_from_A()
    oc.compute_rdm()
    d = DenseCTM([ot.plain(a) for a in oA]).init()
    got = [ot.trace(mod.ct(mod.tw(r, [0, 1]), [1, 2, 3, 4], oc._gate(r, 1.0, 1.0), [3, 4, 1, 2], 0)) for r in oc.rho_h + oc.rho_v]
    return np.abs(np.array(got) - np.array(d.bond_energies())).max(), got


@pytest.mark.parametrize("virtual,seed", [([((0,), 1), ((1,), 1)], 7), (configs.c4_small_virtual(), 3)], ids=["D2", "D3"])
def test_init_density_matrices_match_dense_patch(virtual, seed):
    err, e = _dense_vs_oracle(octm, virtual, seed)
    assert err <= 1e-12
    assert max(abs(x) for x in e) > 1e-3  # a nontrivial state


def _dimer_error(mod, orientation, alpha, beta, steps):
    oA, _ = make_dimer(None, orientation, alpha, beta)
    oc = mod.CTM(oA, 16)
    oc.init_from_A()
    want = np.array(dimer_bond_energies(orientation, alpha, beta))
    worst = 0.0
    for _ in range(steps):
        oc.step()
        got = [ot.trace(mod.ct(mod.tw(r, [0, 1]), [1, 2, 3, 4], oc._gate(r, 1.0, 1.0), [3, 4, 1, 2], 0)) for r in oc.rho_h + oc.rho_v]
        worst = max(worst, float(np.abs(np.array(got) - want).max()))
    return worst


@pytest.mark.parametrize("orientation", ["h", "v"])
@pytest.mark.parametrize("alpha,beta", [(-1.0, 1.0), (0.8, 0.3)], ids=["singlet", "magnetised"])
def test_dimer_states_keep_exact_bond_energies(orientation, alpha, beta):
    assert _dimer_error(octm, orientation, alpha, beta, 3) <= 1e-12


def test_singlet_dimer_energy_is_minus_three_eighths():
    oA, _ = make_dimer(None, "h")
    oc = octm.CTM(oA, 16)
    oc.init_from_A()
    assert abs(oc.step() - (-0.375)) <= 1e-14
    assert oracle_bond_energies(oc)[:2] == pytest.approx([-0.75, -0.75], abs=1e-14)


PLANTS = {
    "rho_h joins the wrong ket/bra legs": ("ct(left_half, [1, 2, -3, 3, -1, 4], right_half, [2, -4, 1, 3, 4, -2], 2)",
                                          "ct(left_half, [1, 2, -1, 3, -3, 4], right_half, [2, -4, 1, 3, 4, -2], 2)"),
    "init C1 legs transposed": ("[1, 2, -2], fuse_dd, [1, 2, -1], 0)", "[1, 2, -1], fuse_dd, [1, 2, -2], 0)"),
    "top move reads T1 of the wrong column": ("self.T1[idx(x, y - 1)], [1, -4, -5, -6], 3)\n                P1T1A",
                                               "self.T1[idx(x + 1, y - 1)], [1, -4, -5, -6], 3)\n                P1T1A"),
    "rho_v reads T4 of the wrong row": ("self.T4[idx(x - 1, y + 1)], [-1, 1, -2, -3], C4T3", "self.T4[idx(x - 1, y)], [-1, 1, -2, -3], C4T3"),
    "left move pairs T4 with the wrong projector": ("tw(P2[idx(x, y - 1)], [0, 1, 2]), [1, 2, 3, -1], 1))\n            for y in range(LY):\n                self.C1",
                                                    "tw(P2[idx(x, y)], [0, 1, 2]), [1, 2, 3, -1], 1))\n            for y in range(LY):\n                self.C1"),
}


@pytest.mark.parametrize("name", list(PLANTS))
def test_planted_mistakes_fail(name):
    src = Path(octm.__file__).read_text()
    a, b = PLANTS[name]
    assert src.count(a) >= 1
    mod = types.ModuleType("oracle.ctm_planted")
    mod.__package__ = "oracle"
    exec(compile(src.replace(a, b, 1), "ctm_planted", "exec"), mod.__dict__)
    caught = False
    for check in (lambda: _dimer_error(mod, "h", 0.8, 0.3, 2) > 1e-6, lambda: _dimer_error(mod, "v", 0.8, 0.3, 2) > 1e-6,
                  lambda: _dense_vs_oracle(mod, [((0,), 1), ((1,), 1)], 7)[0] > 1e-6):
        try:
            caught = caught or check()
        except Exception:  # a structure mismatch raises inside the planted oracle
            caught = True
    assert caught
