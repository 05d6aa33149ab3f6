"""FP64 CPU restatement of one CTMRG step on a 2x2 unit cell (oracle).

Follows src/PEPS/CTM.cpp for TRank = 2, ALL_OUT_LEGS = false, CORNER projection:
init FROM_A (:137-240), grow_all (:386-420) = left/right/top/bottom moves (:924-1183),
get_projectors (:1185-1337) with decompose (PEPSContractions.cpp:16-43), renormalize_* (:1339-1641),
contractCorner (:1643-1760), computeRDM_h/_v (:528-874), and the nearest-neighbour Heisenberg energy
of avg (LinearAlgebra.cpp:198-260, CTMSolver.cpp:60-62, Heisenberg.hpp:47-53).
Tensors follow the oracle's block-sparse layout (oracle/tensor.py).
"""
from __future__ import annotations

import numpy as np

from . import tensor as ot

LX, LY = 2, 2
LEFT, TOP, RIGHT, BOTTOM, PHYS = range(5)


def idx(x, y):
    return (x % LX) * LY + (y % LY)


def tw(t, legs):
    for l in legs:
        t = ot.twist(t, l)
    return t


def ct(a, a1, b, a2, rank):
    return ot.contract(a, list(a1), b, list(a2), rank)


def identity(sym, dom, cod):
    t = ot.Tensor(sym, dom, cod)
    for i in range(len(t.sector)):
        b = np.zeros((t.rows[i], t.cols[i]))
        k = min(t.rows[i], t.cols[i])
        b[range(k), range(k)] = 1.0
        t.set_block(i, b)
    return t


def scaled(t, a):
    out = t.copy()
    out.data = a * t.data
    return out


def normalized(t):
    return scaled(t, 1.0 / ot.max_norm(t))


def diag_map(t, f):
    """diag_sqrt / diag_inv (TensorBase.cpp:191-201): the result block is the diagonal matrix of the
    mapped diagonal (PlainInterface_Eigen_impl.hpp:141-144, M.diagonal().unaryExpr(f).asDiagonal())."""
    out = t.copy()
    for i in range(len(t.sector)):
        d = np.diag(t.block(i)).copy()
        out.set_block(i, np.diag(f(d)))
    return out


class CTM:
    def __init__(self, A, chi):
        self.sym = A[0].sym
        self.A = list(A)
        # iPEPS::updateAdags: A^dagger permuted to (left, top, phys | right, bottom)
        self.Adag = [ot.permute(ot.adjoint(a), 0, [3, 4, 2, 0, 1]) for a in self.A]
        self.chi = chi
        n = LX * LY
        self.C1, self.C2, self.C3, self.C4 = [None] * n, [None] * n, [None] * n, [None] * n
        self.T1, self.T2, self.T3, self.T4 = [None] * n, [None] * n, [None] * n, [None] * n
        self.rho_h, self.rho_v = [None] * n, [None] * n
        self.spectra = []

    def a(self, x, y):
        return self.A[idx(x, y)]

    def ad(self, x, y):
        return self.Adag[idx(x, y)]

    def ket(self, x, y, leg):
        a = self.a(x, y)
        return {LEFT: a.unc_dom[0], TOP: a.unc_dom[1], RIGHT: a.unc_cod[0], BOTTOM: a.unc_cod[1], PHYS: a.unc_cod[2]}[leg]

    def bra(self, x, y, leg):
        a = self.ad(x, y)
        return {LEFT: a.unc_dom[0], TOP: a.unc_dom[1], RIGHT: a.unc_cod[0], BOTTOM: a.unc_cod[1], PHYS: a.unc_dom[2]}[leg]

    # -------------------------------------------------------------------- init FROM_A
    def init_from_A(self):
        s = self.sym
        for x in range(LX):
            for y in range(LY):
                p = idx(x, y)
                a, ad = self.a(x, y), self.ad(x, y)

                def fused(leg):
                    return self.ket(x, y, leg).combine(self.bra(x, y, leg)).forget_history()

                fuse_ll = identity(s, [fused(LEFT)], [self.ket(x, y, LEFT), self.bra(x, y, LEFT)])
                fuse_uu = identity(s, [fused(TOP)], [self.ket(x, y, TOP), self.bra(x, y, TOP)])
                fuse_rr = identity(s, [self.ket(x, y, RIGHT), self.bra(x, y, RIGHT)], [fused(RIGHT)])
                fuse_dd = identity(s, [self.ket(x, y, BOTTOM), self.bra(x, y, BOTTOM)], [fused(BOTTOM)])
                self.C1[p] = ct(ct(ct(a, [1, 2, -1, -2, 3], tw(ad, [3, 4]), [1, 2, 3, -3, -4], 2), [1, -1, 2, -2], fuse_rr, [1, 2, -3], 2),
                                [1, 2, -2], fuse_dd, [1, 2, -1], 0)
                self.C2[p] = ct(ct(ct(a, [-1, 1, 2, -2, 3], tw(ad, [4]), [-3, 1, 3, 2, -4], 2), [1, -1, 2, -2], tw(fuse_ll, [1, 2]), [-3, 1, 2], 2),
                                [1, 2, -1], fuse_dd, [1, 2, -2], 1)
                self.C3[p] = ct(ct(ct(a, [-1, -2, 1, 2, 3], ad, [-3, -4, 3, 1, 2], 2), [1, -1, 2, -2], tw(fuse_ll, [1, 2]), [-3, 1, 2], 2),
                                [1, 2, -2], tw(fuse_uu, [1, 2]), [-1, 1, 2], 2)
                self.C4[p] = ct(ct(ct(a, [1, -1, -2, 2, 3], tw(ad, [3]), [1, -3, 3, -4, 2], 2), [1, -1, 2, -2], tw(fuse_uu, [1, 2]), [-3, 1, 2], 2),
                                [1, 2, -1], fuse_rr, [1, 2, -2], 1)
                self.T1[p] = ct(ct(ct(a, [-1, 1, -2, -3, 2], tw(ad, [3, 4]), [-4, 1, 2, -5, -6], 3), [1, -1, -2, 2, -3, -4], tw(fuse_ll, [1, 2]),
                                   [-5, 1, 2], 4), [1, -3, 2, -4, -1], fuse_rr, [1, 2, -2], 1)
                self.T2[p] = ct(ct(ct(a, [-1, -2, 1, -3, 2], tw(ad, [4]), [-4, -5, 2, 1, -6], 3), [-1, 1, -2, -3, 2, -4], tw(fuse_uu, [1, 2]),
                                   [-5, 1, 2], 4), [-1, 1, -2, 2, -3], fuse_dd, [1, 2, -4], 3)
                self.T3[p] = ct(ct(ct(a, [-1, -2, -3, 1, 2], tw(ad, [3]), [-4, -5, 2, -6, 1], 3), [1, -1, -2, 2, -3, -4], tw(fuse_ll, [1, 2]),
                                   [-5, 1, 2], 4), [-1, 1, -2, 2, -3], fuse_rr, [1, 2, -4], 3)
                self.T4[p] = ct(ct(ct(a, [1, -1, -2, -3, 2], tw(ad, [3, 4]), [1, -4, 2, -5, -6], 3), [1, -1, -2, 2, -3, -4], tw(fuse_uu, [1, 2]),
                                   [-5, 1, 2], 4), [-3, 1, -4, 2, -1], fuse_dd, [1, 2, -2], 1)

    # -------------------------------------------------------------------- corners and projectors
    def corner(self, x, y, c):
        if c == 0:  # upper left
            C1T1 = ct(self.C1[idx(x - 1, y - 1)], [-1, 1], self.T1[idx(x, y - 1)], [1, -2, -3, -4], 1)
            T4C1T1 = ct(self.T4[idx(x - 1, y)], [1, -1, -2, -3], tw(C1T1, [0]), [1, -4, -5, -6], 3)
            T4C1T1A = ct(T4C1T1, [-1, 1, -2, -3, 2, -4], self.a(x, y), [1, 2, -5, -6, -7], 4)
            return ct(T4C1T1A, [-1, 1, -4, 2, -5, -2, 3], tw(self.ad(x, y), [3, 4]), [1, 2, 3, -6, -3], 3)
        if c == 1:  # lower left
            C4T3 = ct(self.C4[idx(x - 1, y + 1)], [-1, 1], self.T3[idx(x, y + 1)], [-2, -3, 1, -4], 1)
            T4C4T3 = ct(self.T4[idx(x - 1, y)], [-1, 1, -2, -3], C4T3, [1, -4, -5, -6], 3)
            T4C4T3A = ct(T4C4T3, [-1, 1, -2, 2, -3, -4], tw(self.a(x, y), [3]), [1, -5, -6, 2, -7], 4)
            return ct(T4C4T3A, [-4, 1, 2, -1, -5, -2, 3], tw(self.ad(x, y), [3]), [1, -6, 3, -3, 2], 3)
        if c == 2:  # upper right
            T1C2 = ct(self.T1[idx(x, y - 1)], [-1, 1, -2, -3], self.C2[idx(x + 1, y - 1)], [1, -4], 3)
            T1C2T2 = ct(T1C2, [-1, -2, -3, 1], self.T2[idx(x + 1, y)], [-4, -5, 1, -6], 3)
            T1C2T2A = ct(T1C2T2, [-1, 1, -2, 2, -3, -4], tw(self.a(x, y), [2]), [-5, 1, 2, -6, -7], 4)
            return ct(T1C2T2A, [-1, 1, 2, -4, -2, -5, 3], tw(self.ad(x, y), [4]), [-3, 1, 3, 2, -6], 3)
        C3T3 = ct(tw(self.C3[idx(x + 1, y + 1)], [1]), [-1, 1], self.T3[idx(x, y + 1)], [-2, -3, -4, 1], 1)
        T2C3T3 = ct(self.T2[idx(x + 1, y)], [-1, -2, -3, 1], C3T3, [1, -4, -5, -6], 3)
        T2C3T3A = ct(T2C3T3, [1, -1, -2, 2, -3, -4], tw(self.a(x, y), [2, 3]), [-5, -6, 1, 2, -7], 4)
        return ct(T2C3T3A, [1, -1, 2, -4, -5, -2, 3], self.ad(x, y), [-6, -3, 3, 1, 2], 3)

    def decompose(self, Ta, Tb, site, direction):
        prod = ot.mul(Ta, Tb)
        U, S, V, tw_, ent, kept = ot.tsvd(prod, self.chi, 1.0e-10, False)
        self.spectra.append((site, direction, {q: len(v) for q, v in kept.items()}, sorted(x for v in kept.values() for x in v)))
        isqrtS = diag_map(S, lambda d: 1.0 / np.sqrt(d))
        P1 = ot.mul(ot.mul(isqrtS, ot.adjoint(U)), Ta)
        P2 = ot.mul(ot.mul(Tb, ot.adjoint(V)), isqrtS)
        return P1, P2

    def projectors(self, x, y, d):
        site = idx(x, y)
        if d == 0:
            return self.decompose(tw(self.corner(x, y + 1, 1), [3, 4, 5]), self.corner(x, y, 0), site, d)
        if d == 1:
            return self.decompose(self.corner(x, y, 2), self.corner(x, y + 1, 3), site, d)
        if d == 2:
            return self.decompose(self.corner(x, y, 0), self.corner(x + 1, y, 2), site, d)
        return self.decompose(self.corner(x + 1, y, 3), tw(self.corner(x, y, 1), [0, 1, 2]), site, d)

    # -------------------------------------------------------------------- moves
    def left_move(self):
        for x in range(LX):
            P1, P2 = {}, {}
            for y in range(LY):
                P1[idx(x, y)], P2[idx(x, y)] = self.projectors(x, y, 0)
            C1n, T4n, C4n = {}, {}, {}
            for y in range(LY):
                C1T1 = ct(self.C1[idx(x - 1, y - 1)], [-1, 1], self.T1[idx(x, y - 1)], [1, -4, -2, -3], 3)
                C1n[idx(x, y - 1)] = normalized(ct(tw(P1[idx(x, y - 1)], [1, 2, 3]), [-1, 1, 2, 3], C1T1, [1, 2, 3, -2], 0))
                C4T3 = ct(self.C4[idx(x - 1, y + 1)], [-2, 1], self.T3[idx(x, y + 1)], [-3, -4, 1, -1], 1)
                C4n[idx(x, y + 1)] = normalized(ct(C4T3, [-2, 1, 2, 3], tw(P2[idx(x, y)], [0, 1, 2]), [1, 2, 3, -1], 1))
                P1T4 = ct(tw(P1[idx(x, y)], [1, 2, 3]), [-1, 1, -2, -3], self.T4[idx(x - 1, y)], [-4, 1, -5, -6], 3)
                P1T4A = ct(P1T4, [-1, 2, -2, -3, 1, -4], self.a(x, y), [1, -5, -6, 2, -7], 4)
                P1T4AAdag = ct(P1T4A, [-1, 2, -4, 1, -5, -2, 3], tw(self.ad(x, y), [3, 4]), [1, -6, 3, -3, 2], 3)
                T4n[idx(x, y)] = normalized(ct(P1T4AAdag, [-2, -3, -4, 1, 2, 3], tw(P2[idx(x, y - 1)], [0, 1, 2]), [1, 2, 3, -1], 1))
            for y in range(LY):
                self.C1[idx(x, y)], self.T4[idx(x, y)], self.C4[idx(x, y)] = C1n[idx(x, y)], T4n[idx(x, y)], C4n[idx(x, y)]

    def right_move(self):
        for xs in range(LX, 0, -1):
            x = xs % LX
            P1, P2 = {}, {}
            for y in range(LY):
                P1[idx(x, y)], P2[idx(x, y)] = self.projectors(x, y, 1)
            C2n, T2n, C3n = {}, {}, {}
            for y in range(LY):
                T1C2 = ct(self.T1[idx(x, y - 1)], [-1, 1, -3, -4], self.C2[idx(x + 1, y - 1)], [1, -2], 1)
                C2n[idx(x, y - 1)] = normalized(ct(T1C2, [-1, 1, 2, 3], P2[idx(x, y - 1)], [1, 2, 3, -2], 1))
                C3T3 = ct(tw(self.C3[idx(x + 1, y + 1)], [1]), [-1, 1], self.T3[idx(x, y + 1)], [-2, -3, -4, 1], 3)
                C3n[idx(x, y + 1)] = normalized(ct(P1[idx(x, y)], [-1, 1, 2, 3], C3T3, [1, 2, 3, -2], 2))
                P1T2 = ct(P1[idx(x, y - 1)], [-1, 1, -2, -3], self.T2[idx(x + 1, y)], [-4, -5, 1, -6], 3)
                P1T2A = ct(P1T2, [-1, 1, -2, 2, -3, -4], tw(self.a(x, y), [2]), [-5, 1, 2, -6, -7], 4)
                P1T2AAdag = ct(P1T2A, [-1, 1, 2, -4, -2, -5, 3], tw(self.ad(x, y), [4]), [-3, 1, 3, 2, -6], 3)
                T2n[idx(x, y)] = normalized(ct(P1T2AAdag, [-3, -1, -2, 1, 2, 3], P2[idx(x, y)], [1, 2, 3, -4], 3))
            for y in range(LY):
                self.C2[idx(x, y)], self.T2[idx(x, y)], self.C3[idx(x, y)] = C2n[idx(x, y)], T2n[idx(x, y)], C3n[idx(x, y)]

    def top_move(self):
        for y in range(LY):
            P1, P2 = {}, {}
            for x in range(LX):
                P1[idx(x, y)], P2[idx(x, y)] = self.projectors(x, y, 2)
            C1n, T1n, C2n = {}, {}, {}
            for x in range(LX):
                T4C1 = ct(self.T4[idx(x - 1, y)], [1, -1, -3, -4], tw(self.C1[idx(x - 1, y - 1)], [0]), [1, -2], 1)
                C1n[idx(x - 1, y)] = normalized(ct(T4C1, [-1, 1, 2, 3], P2[idx(x - 1, y)], [1, 2, 3, -2], 0))
                C2T2 = ct(self.C2[idx(x + 1, y - 1)], [-1, 1], self.T2[idx(x + 1, y)], [-2, -3, 1, -4], 3)
                C2n[idx(x + 1, y)] = normalized(ct(P1[idx(x, y)], [-1, 1, 2, 3], C2T2, [1, 2, 3, -2], 1))
                P1T1 = ct(P1[idx(x - 1, y)], [-1, 1, -2, -3], self.T1[idx(x, y - 1)], [1, -4, -5, -6], 3)
                P1T1A = ct(P1T1, [-1, 1, -2, -3, 2, -4], self.a(x, y), [1, 2, -5, -6, -7], 4)
                P1T1AAdag = ct(P1T1A, [-1, 1, -4, 2, -5, -2, 3], tw(self.ad(x, y), [3, 4]), [1, 2, 3, -6, -3], 3)
                T1n[idx(x, y)] = normalized(ct(P1T1AAdag, [-1, -3, -4, 1, 2, 3], P2[idx(x, y)], [1, 2, 3, -2], 1))
            for x in range(LX):
                self.C1[idx(x, y)], self.T1[idx(x, y)], self.C2[idx(x, y)] = C1n[idx(x, y)], T1n[idx(x, y)], C2n[idx(x, y)]

    def bottom_move(self):
        for ys in range(LY, 0, -1):
            y = ys % LY
            P1, P2 = {}, {}
            for x in range(LX):
                P1[idx(x, y)], P2[idx(x, y)] = self.projectors(x, y, 3)
            C4n, T3n, C3n = {}, {}, {}
            for x in range(LX):
                T4C4 = ct(self.T4[idx(x - 1, y)], [-4, 1, -2, -3], self.C4[idx(x - 1, y + 1)], [1, -1], 3)
                C4n[idx(x - 1, y)] = normalized(ct(tw(P1[idx(x - 1, y)], [1, 2, 3]), [-2, 1, 2, 3], T4C4, [1, 2, 3, -1], 1))
                T2C3 = ct(self.T2[idx(x + 1, y)], [-3, -4, -1, 1], self.C3[idx(x + 1, y + 1)], [1, -2], 1)
                C3n[idx(x + 1, y)] = normalized(ct(T2C3, [-1, 1, 2, 3], tw(P2[idx(x, y)], [0, 1, 2]), [1, 2, 3, -2], 2))
                P1T3 = ct(tw(P1[idx(x, y)], [1, 2, 3]), [-1, 1, -2, -3], tw(self.T3[idx(x, y + 1)], [0]), [-4, -5, -6, 1], 3)
                P1T3A = ct(P1T3, [-1, 1, -2, 2, -3, -4], self.a(x, y), [-5, -6, 1, 2, -7], 4)
                P1T3AAdag = ct(P1T3A, [-1, 1, 2, -4, -5, -2, 3], tw(self.ad(x, y), [3]), [-6, -3, 3, 1, 2], 3)
                T3n[idx(x, y)] = normalized(ct(P1T3AAdag, [-4, -1, -2, 1, 2, 3], tw(P2[idx(x - 1, y)], [0, 1, 2]), [1, 2, 3, -3], 3))
            for x in range(LX):
                self.C4[idx(x, y)], self.T3[idx(x, y)], self.C3[idx(x, y)] = C4n[idx(x, y)], T3n[idx(x, y)], C3n[idx(x, y)]

    def grow_all(self):
        self.left_move()
        self.right_move()
        self.top_move()
        self.bottom_move()

    # -------------------------------------------------------------------- observables
    def _normalize_rho(self, rho):
        Id2 = identity(self.sym, rho.unc_cod, rho.unc_dom)
        norm = ot.trace(ct(rho, [1, 2, 3, 4], tw(Id2, [0, 1]), [3, 4, 1, 2], 0))
        return scaled(rho, 1.0 / norm)

    def compute_rdm(self):
        for x in range(LX):
            for y in range(LY):
                C1T1 = ct(self.C1[idx(x - 1, y - 1)], [-1, 1], self.T1[idx(x, y - 1)], [1, -2, -3, -4], 1)
                T4C1T1 = ct(self.T4[idx(x - 1, y)], [1, -1, -2, -3], tw(C1T1, [0]), [1, -4, -5, -6], 3)
                T4C1T1A = ct(T4C1T1, [-1, 1, -2, -3, 2, -4], self.a(x, y), [1, 2, -5, -6, -7], 4)
                C4T3 = ct(self.C4[idx(x - 1, y + 1)], [-1, 1], self.T3[idx(x, y + 1)], [-2, -3, 1, -4], 1)
                C4T3Ad = ct(C4T3, [-1, -2, 1, -3], tw(self.ad(x, y), [3]), [-4, -5, -6, -7, 1], 3)
                left_half = ct(T4C1T1A, [1, 2, -1, 3, -2, 4, -3], C4T3Ad, [1, 4, -4, 2, 3, -5, -6], 3)
                T1C2 = ct(self.T1[idx(x + 1, y - 1)], [-1, 1, -2, -3], self.C2[idx(x + 2, y - 1)], [1, -4], 3)
                T1C2T2 = ct(T1C2, [-1, -2, -3, 1], self.T2[idx(x + 2, y)], [-4, -5, 1, -6], 3)
                AT1C2T2 = ct(tw(self.a(x + 1, y), [1]), [-1, 1, 2, -2, -3], T1C2T2, [-4, 1, -5, 2, -6, -7], 3)
                T3C3 = ct(self.T3[idx(x + 1, y + 1)], [-1, -2, -3, 1], self.C3[idx(x + 2, y + 1)], [-4, 1], 3)
                T3C3Ad = ct(T3C3, [-1, 1, -2, -3], self.ad(x + 1, y), [-4, -5, -6, -7, 1], 3)
                right_half = ct(AT1C2T2, [-1, 1, -2, -3, 2, 3, 4], T3C3Ad, [1, -4, 4, -5, 2, -6, 3], 3)
                self.rho_h[idx(x, y)] = self._normalize_rho(ct(left_half, [1, 2, -3, 3, -1, 4], right_half, [2, -4, 1, 3, 4, -2], 2))
        for x in range(LX):
            for y in range(LY):
                C1T1 = ct(self.C1[idx(x - 1, y - 1)], [-1, 1], self.T1[idx(x, y - 1)], [1, -2, -3, -4], 1)
                T4C1T1 = ct(self.T4[idx(x - 1, y)], [1, -1, -2, -3], tw(C1T1, [0]), [1, -4, -5, -6], 3)
                T4C1T1A = ct(T4C1T1, [-1, 1, -2, -3, 2, -4], self.a(x, y), [1, 2, -5, -6, -7], 4)
                C2T2 = ct(self.C2[idx(x + 1, y - 1)], [-1, 1], self.T2[idx(x + 1, y)], [-2, -3, 1, -4], 1)
                AdC2T2 = ct(tw(self.ad(x, y), [3, 4]), [-1, -2, -3, 1, -4], C2T2, [-5, -6, 1, -7], 4)
                upper_half = ct(T4C1T1A, [-1, 1, 2, 3, 4, -2, -3], AdC2T2, [1, 3, -4, -5, 2, 4, -6], 3)
                C4T3 = ct(self.C4[idx(x - 1, y + 2)], [-1, 1], self.T3[idx(x, y + 2)], [-2, -3, 1, -4], 1)
                T4C4T3 = ct(self.T4[idx(x - 1, y + 1)], [-1, 1, -2, -3], C4T3, [1, -4, -5, -6], 3)
                AT4C4T3 = ct(tw(self.a(x, y + 1), [0]), [1, -1, -2, 2, -3], T4C4T3, [-4, 1, -5, 2, -6, -7], 3)
                T2C3 = ct(self.T2[idx(x + 1, y + 1)], [-1, -2, -3, 1], self.C3[idx(x + 1, y + 2)], [1, -4], 3)
                T2C3Ad = ct(T2C3, [-1, 1, -2, -3], self.ad(x, y + 1), [-4, -5, -6, 1, -7], 3)
                lower_half = ct(AT4C4T3, [-1, 1, -2, -3, 2, 3, 4], T2C3Ad, [1, -4, 4, 2, -5, -6, 3], 3)
                self.rho_v[idx(x, y)] = self._normalize_rho(ct(upper_half, [1, 2, -3, -1, 3, 4], lower_half, [2, -4, 1, 4, 3, -2], 2))

    def _gate(self, rho, Jxy, Jz):
        g = ot.Tensor(self.sym, rho.unc_cod, rho.unc_dom)
        for i, q in enumerate(g.sector):
            b = np.zeros((g.rows[i], g.cols[i]))
            r0 = 0
            for td in g.dom.tree(q):
                c0 = 0
                for tc in g.cod.tree(q):
                    if td.dim == 1 and tc.dim == 1:
                        def sz(basis, lab):
                            return 0.5 * (-lab[0] if basis.conj_flag else lab[0])
                        c1, c2 = sz(g.unc_dom[0], td.unc[0]), sz(g.unc_dom[1], td.unc[1])
                        a1, a2 = sz(g.unc_cod[0], tc.unc[0]), sz(g.unc_cod[1], tc.unc[1])
                        v = 0.0
                        if c1 == a1 and c2 == a2:
                            v = Jz * a1 * a2
                        if c1 == a1 + 1 and c2 == a2 - 1:
                            v = 0.5 * Jxy
                        if c1 == a1 - 1 and c2 == a2 + 1:
                            v = 0.5 * Jxy
                        b[r0, c0] = v
                    c0 += tc.dim
                r0 += td.dim
            g.set_block(i, b)
        return g

    def energy(self, Jxy=1.0, Jz=1.0):
        e = 0.0
        for rho in self.rho_h + self.rho_v:
            e += ot.trace(ct(tw(rho, [0, 1]), [1, 2, 3, 4], self._gate(rho, Jxy, Jz), [3, 4, 1, 2], 0))
        return e / (LX * LY)

    def step(self):
        self.grow_all()
        self.compute_rdm()
        return self.energy()
