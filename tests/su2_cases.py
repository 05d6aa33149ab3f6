"""Shared cases of the SU(2) dense-expansion checks (the reference's tests/tensor.cpp:75-384 and
tests/tensor_tests.hpp:1-200): bases, the closed-form spin matrices and the operator chains of the
spin-matrix test, written once for the oracle and for the HIP path."""
import math

import numpy as np

from oracle import tensor as oten
from oracle.structure import Qbasis as OQbasis
from oracle.symmetry import FU1, SU2, U1, Sym

SU2_BASES = [
    [((1,), 2), ((2,), 1), ((3,), 1)],
    [((2,), 2), ((1,), 1), ((3,), 1)],
    [((1,), 1), ((2,), 2), ((4,), 1)],
]


def pauli_vec(two_s):
    """-sqrt(1/2) S+, Sz, sqrt(1/2) S- of spin S = two_s / 2 in the basis m = S, S-1, .. (tests/tensor.cpp:251-277)."""
    S = 0.5 * two_s
    n = two_s + 1
    sp, sz, sm = np.zeros((n, n)), np.zeros((n, n)), np.zeros((n, n))
    for i in range(1, n):
        m = S - i
        sp[i - 1, i] = math.sqrt(S * (S + 1) - m * (m + 1))
    for i in range(n):
        sz[i, i] = S - i
    for i in range(n - 1):
        m = S - i
        sm[i + 1, i] = math.sqrt(S * (S + 1) - m * (m - 1))
    return [-math.sqrt(0.5) * sp, sz, math.sqrt(0.5) * sm]


def rank2_products(P):
    """Closed forms of the K = 2 coupling of two spin operators (tests/tensor.cpp:347-355)."""
    return [-math.sqrt(0.6) * (P[0] @ P[0]), -math.sqrt(0.3) * (P[0] @ P[1] + P[1] @ P[0]),
            -math.sqrt(0.1) * (P[0] @ P[2] + P[2] @ P[0] + 2 * P[1] @ P[1]), -math.sqrt(0.3) * (P[2] @ P[1] + P[1] @ P[2]),
            -math.sqrt(0.6) * (P[2] @ P[2])]


def graded_transpose(sym, legs, dense, p):
    """Leg permutation p of a dense fermionic tensor with the graded sign: every pair of legs whose
    order p inverts contributes (-1)^(parity_a parity_b), parity = fermion number mod 2 of the
    entry's sector on that leg.  legs: the sorted oracle bases of the dense legs."""
    n = len(legs)
    par = []
    for b in legs:
        v = np.zeros(b.dim(), dtype=np.int64)
        for q, _, dim, _ in b.data:
            o = b.full_outer_num(q)
            f = sum(x for k, x in zip(sym.kinds, q) if k == FU1)
            v[o:o + dim] = f % 2
        par.append(v)
    sign = np.ones(dense.shape)
    for a in range(n):
        for b in range(a + 1, n):
            # output positions of legs a and b; inverted when a lands after b
            if p.index(a) > p.index(b):
                pa = par[a].reshape([-1 if i == a else 1 for i in range(n)])
                pb = par[b].reshape([-1 if i == b else 1 for i in range(n)])
                sign = sign * np.where((pa * pb) % 2 == 1, -1.0, 1.0)
    return np.transpose(sign * dense, p)


def sorted_legs(t):
    out = []
    for b in t.unc_dom + t.unc_cod:
        c = b.copy()
        c.sort()
        out.append(c)
    return out


def obasis(sym, entries):
    b = OQbasis(sym, 1)
    for q, d in entries:
        b.push_back(q, d)
    return b


class OracleOps:
    """The operations of the spin-matrix chains on the oracle."""

    sym = Sym(SU2)

    def tensor(self, dom, cod, value):
        t = oten.Tensor(self.sym, [obasis(self.sym, e) for e in dom], [obasis(self.sym, e) for e in cod])
        t.data[:] = value
        return t

    adjoint = staticmethod(oten.adjoint)
    permute = staticmethod(oten.permute)
    mul = staticmethod(oten.mul)

    def plain(self, t):
        return oten.plain(t)


def spin_chain_checks(ops, two_s1, two_s2_values=range(1, 11), tol=1e-12):
    """tests/tensor.cpp:244-382 for one S1: the spin matrices of the reduced operator, the K = 0, 1, 2
    couplings of two of them (S^2, S/sqrt(2), rank-2 products) and the K = 0 coupling S1.S2 with a
    second spin, whose 1 x 1 blocks equal (Stot(Stot+1) - S1(S1+1) - S2(S2+1)) / 2.  Returns the
    largest deviation from the closed forms."""
    S1 = 0.5 * two_s1
    n = two_s1 + 1
    P = pauli_vec(two_s1)
    worst = 0.0
    s1 = ops.tensor([[((n,), 1)], [((3,), 1)]], [[((n,), 1)]], math.sqrt(S1 * (S1 + 1)))
    tp = ops.plain(ops.adjoint(s1))
    for k in range(3):
        worst = max(worst, np.abs(tp[:, :, k] - P[k]).max())
    sa = ops.permute(ops.adjoint(s1), -1, [0, 1, 2])
    chains = {}
    for K, q in ((0, 1), (1, 3), (2, 5)):
        couple = ops.tensor([[((3,), 1)], [((q,), 1)]], [[((3,), 1)]], 1.0)
        left = ops.mul(sa, ops.permute(couple, +1, [0, 2, 1]))
        chains[K] = left
        prod = ops.mul(ops.permute(left, 0, [0, 3, 1, 2]), s1)
        if K == 0:
            c = ops.plain(prod)
            for j in range(c.shape[1]):
                worst = max(worst, np.abs(c[:, j, :] - S1 * (S1 + 1) * np.eye(n)).max())
        else:
            c = ops.plain(ops.adjoint(prod))
            want = [math.sqrt(0.5) * x for x in P] if K == 1 else rank2_products(P)
            for k in range(c.shape[2]):
                worst = max(worst, np.abs(c[:, :, k] - want[k]).max())
    for two_s2 in two_s2_values:
        S2 = 0.5 * two_s2
        m = two_s2 + 1
        s2 = ops.tensor([[((m,), 1)], [((3,), 1)]], [[((m,), 1)]], math.sqrt(S2 * (S2 + 1)))
        outer = ops.permute(ops.mul(ops.permute(chains[0], -1, [0, 1, 3, 2]), ops.permute(s2, +1, [1, 2, 0])), 0, [0, 4, 2, 1, 3])
        for q, blk in ops.blocks(outer):
            assert blk.shape == (1, 1)
            Stot = 0.5 * (q[0] - 1)
            worst = max(worst, abs(blk[0, 0] - 0.5 * (Stot * (Stot + 1) - S1 * (S1 + 1) - S2 * (S2 + 1))))
    return worst


def _oracle_blocks(t):
    return [(q, t.block(i)) for i, q in enumerate(t.sector)]


OracleOps.blocks = staticmethod(_oracle_blocks)
