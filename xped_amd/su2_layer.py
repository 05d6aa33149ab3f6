"""The SU(2) double-layer workload (C2): the reference's contractAAdag
(src/PEPS/PEPSContractions.cpp:106-129) and the absorption of a site into a CTM edge tensor
T(chi | chi, D_ket, D_bra) (the T4 * A * A^dagger chain of the CTM moves, src/PEPS/CTM.cpp:924-1183),
written once over any tensor backend that has contract / permute / adjoint / twist and an
Identity constructor: the HIP path (xped_amd.tensor) and the oracle (for the parity tests) share
the index lists.  The contraction index lists are the reference's (an interface kept alike)."""
from __future__ import annotations


def adag_of(ops, A):
    """iPEPS::updateAdags: A^dagger permuted to (l, t, p | r, b)."""
    return ops.permute(ops.adjoint(A), 0, [3, 4, 2, 0, 1])


def contract_aadag(ops, A, Adag):
    """contractAAdag (PEPSContractions.cpp:106-129): the double layer (ll, tt | rr, dd)."""
    l, t, r, b = ops.leg(A, 0), ops.leg(A, 1), ops.leg(A, 2), ops.leg(A, 3)
    lA, tA, rA, bA = ops.leg(Adag, 0), ops.leg(Adag, 1), ops.leg(Adag, 3), ops.leg(Adag, 4)
    fuse_ll = ops.identity([ops.fuse(l, lA)], [l, lA])
    fuse_tt = ops.identity([ops.fuse(t, tA)], [t, tA])
    fuse_rr = ops.identity([r, rA], [ops.fuse(r, rA)])
    fuse_dd = ops.identity([b, bA], [ops.fuse(b, bA)])
    tw = ops.twist
    X = ops.contract(A, [-1, -2, -3, -4, 1], tw(tw(Adag, 3), 4), [-5, -6, 1, -7, -8], 8)
    X = ops.contract(X, [1, -1, -2, -3, 2, -4, -5, -6], tw(tw(fuse_ll, 1), 2), [-7, 1, 2], 6)
    X = ops.contract(X, [1, -1, -2, 2, -3, -4, -5], tw(tw(fuse_tt, 1), 2), [-6, 1, 2], 5)
    X = ops.contract(X, [1, -1, 2, -2, -3, -4], fuse_rr, [1, 2, -5], 4)
    return ops.contract(X, [1, 2, -1, -2, -3], fuse_dd, [1, 2, -4], 2)


def absorb(ops, T, A, Adag):
    """T(chi_u | chi_d, D_ket, D_bra) absorbing the site (A, Adag) over its ket and bra legs and the
    physical leg: (chi_u, chi_d, t, r | b, t', r', b')."""
    TA = ops.contract(T, [-1, -2, 1, -3], A, [1, -4, -5, -6, -7], 3)
    return ops.contract(TA, [-1, -2, 1, -3, -4, -5, 2], Adag, [1, -6, 2, -7, -8], 4)


def absorb_other_order(ops, T, A, Adag):
    """The same network as absorb, contracted in another order (ket and bra joined first over the
    physical leg), permuted to absorb's leg order: a full-size consistency check."""
    AA = ops.contract(A, [-1, -2, -3, -4, 1], Adag, [-5, -6, 1, -7, -8], 4)  # (l, t, r, b | l', t', r', b')
    return ops.contract(T, [-1, -2, 1, 2], AA, [1, -3, -4, -5, 2, -6, -7, -8], 4)


class HipOps:
    """xped_amd.tensor operations for the shared index lists."""

    def __init__(self, rt):
        self.rt = rt

    def contract(self, a, a1, b, a2, rank):
        return a.contract(b, a1, a2, rank)

    def permute(self, t, shift, p):
        return t.permute(shift, p)

    def adjoint(self, t):
        return t.adjoint()

    def twist(self, t, leg):
        return t.twist(leg)

    def leg(self, t, i):
        return t.leg(i)

    def fuse(self, a, b):
        return a.combine(b).forgetHistory()

    def identity(self, dom, cod):
        from .tensor import Tensor

        return Tensor.identity(self.rt, dom, cod)
