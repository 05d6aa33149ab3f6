"""The HIP permute and product of SU(2) and fermionic tensors against dense expansions (needs an
MI355X).  Same checks as tests/test_oracle_su2_dense.py, on data that the HIP kernels computed:
the structure comes out bit-exact (compared with the oracle's), the data is expanded with exact
Clebsch-Gordan tensors and compared with numpy's dense shuffle (tests/tensor_tests.hpp:1-200) and
with the closed-form spin matrices of tests/tensor.cpp:230-384.  Tolerance 1e-13 (absolute, O(1)
entries) for permutes, 1e-12 for the product chains."""
import itertools

import numpy as np
import pytest

import xped_amd.tensor as xt
from oracle import tensor as oten
from oracle.symmetry import FU1, SU2, U1, Sym

from .helpers import both_basis, oracle_structure, structure_of
from .su2_cases import SU2_BASES, OracleOps, graded_transpose, sorted_legs, spin_chain_checks

pytestmark = pytest.mark.gpu


def _both(rt, pcode, entries, rank, seed):
    pairs = [both_basis(pcode, e) for e in entries]
    o = oten.Tensor(pairs[0][0].sym, [p[0] for p in pairs[:rank]], [p[0] for p in pairs[rank:]]).set_random(seed)
    t = xt.Tensor(rt, [p[1] for p in pairs[:rank]], [p[1] for p in pairs[rank:]])
    t.upload(o.data)
    return o, t


def _expanded(o_like, t):
    """Dense expansion of the HIP tensor t, whose structure must equal the oracle tensor o_like's."""
    assert structure_of(t) == oracle_structure(o_like)
    o = o_like.copy()
    o.data = t.download()
    return oten.plain(o)


@pytest.mark.parametrize("b,c", [(0, 1), (1, 2)])
def test_hip_su2_permute_matches_dense_shuffle(rt, b, c):
    o, t = _both(rt, xt.SU2, [SU2_BASES[b], SU2_BASES[c], SU2_BASES[b], SU2_BASES[c]], 2, 5 + b)
    P = oten.plain(o)
    worst = 0.0
    for shift in range(-2, 3):
        for p in itertools.permutations(range(4)):
            got = _expanded(oten.permute(o, shift, p), t.permute(shift, p))
            worst = max(worst, np.abs(got - np.transpose(P, p)).max())
    assert worst <= 1e-13


def test_hip_su2_permute_within_domain(rt):
    o, t = _both(rt, xt.SU2, [SU2_BASES[0]] * 4, 4, 9)
    P = oten.plain(o)
    for p in itertools.permutations(range(4)):
        got = _expanded(oten.permute(o, 0, p), t.permute(0, p))
        assert np.abs(got - np.transpose(P, p)).max() <= 1e-13, p


class GpuOps(OracleOps):
    """The spin-matrix chains on the HIP path; the oracle runs alongside for the structures only
    (the data compared with the closed forms is the HIP result)."""

    def __init__(self, rt):
        self.rt = rt

    def tensor(self, dom, cod, value):
        o = OracleOps.tensor(self, dom, cod, value)
        pb = [both_basis(xt.SU2, e)[1] for e in list(dom) + list(cod)]
        t = xt.Tensor(self.rt, pb[:len(dom)], pb[len(dom):])
        t.upload(o.data)
        return (o, t)

    def adjoint(self, x):
        return (oten.adjoint(x[0]), x[1].adjoint())

    def permute(self, x, shift, p):
        return (oten.permute(x[0], shift, p), x[1].permute(shift, list(p)))

    def mul(self, a, b):
        return (oten.mul(a[0], b[0]), a[1] * b[1])

    def plain(self, x):
        return _expanded(x[0], x[1])

    def blocks(self, x):
        o = x[0].copy()
        assert structure_of(x[1]) == oracle_structure(o)
        o.data = x[1].download()
        return [(q, o.block(i)) for i, q in enumerate(o.sector)]


@pytest.mark.parametrize("two_s1", [1, 2, 3, 6, 10])
def test_hip_su2_spin_matrix_closed_forms(rt, two_s1):
    assert spin_chain_checks(GpuOps(rt), two_s1, two_s2_values=[1, 2, 5, 10]) <= 1e-12


def test_hip_fermionic_split_permutes_are_graded_shuffles(rt):
    s = Sym(FU1, U1)
    ent = [((1, 0), 2), ((0, 0), 1), ((-1, 0), 2), ((1, 1), 1), ((-1, -1), 1)]
    for rank in (4, 2):
        o, t = _both(rt, xt.FU1xU1, [ent] * 4, rank, 21 + rank)
        P, L = oten.plain(o), sorted_legs(o)
        for p in itertools.permutations(range(4)):
            if rank == 2 and set(p[:2]) != {0, 1}:
                continue
            got = _expanded(oten.permute(o, 0, list(p)), t.permute(0, list(p)))
            assert np.array_equal(got, graded_transpose(s, L, P, list(p))), p
