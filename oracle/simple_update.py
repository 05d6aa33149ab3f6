"""FP64 CPU restatement of the simple update on a 2x2 unit cell (oracle; test infrastructure only).

Follows the nearest-neighbour branch of TimePropagator (src/PEPS/TimePropagator.cpp): t_step
(:78-105), t_step_h (:113-190), t_step_v (:192-226), initU (:453-487); SimpleUpdate::renormalize
(src/PEPS/SimpleUpdate.cpp:5-14); applyWeights (src/PEPS/PEPSContractions.cpp:93-104);
iPEPS::initWeightTensors / updateAtensors (src/PEPS/iPEPS.cpp:575-602).  The gates are
TensorBase::mexp (TensorBase.cpp:180-183): Eigen's matrix exponential per block, here
scipy.linalg.expm (Pade with scaling and squaring, the method Eigen's MatrixFunctions use).

Sites and bonds are indexed by idx(x, y) = (x mod 2) * 2 + (y mod 2); whs[idx(x, y)] is the weight of
the bond (x, y)-(x+1, y), wvs[idx(x, y)] that of the bond (x, y-1)-(x, y).

Pinned (tests/test_oracle_su_pins.py) by dense invariants that do not restate this code: an
untruncated bond update reproduces the gate applied to the dense two-site block, the new weights
are the singular values of that block across the bond, and a truncated update is its best rank-D
approximation; the gate against a closed-form exponential of the Heisenberg bond.
"""
from __future__ import annotations

import sys

import numpy as np
import scipy.linalg

from . import tensor as ot
from .ctm import ct, diag_map, identity, scaled

L = 2
ALL = sys.maxsize


def idx(x, y):
    return (x % L) * L + (y % L)


def mexp(t: ot.Tensor, factor: float) -> ot.Tensor:
    out = t.copy()
    for i in range(len(t.sector)):
        out.set_block(i, scipy.linalg.expm(factor * t.block(i)))
    return out


def sqrt_w(w):
    return diag_map(w, np.sqrt)


def inv_w(w):
    return diag_map(w, lambda d: 1.0 / d)


def identity_on(w):
    return identity(w.sym, w.unc_dom, w.unc_cod)


def apply_weights(A, wL, wT, wR, wB):
    t = ct(A, [1, -2, -3, -4, -5], ot.twist(wL, 1), [-1, 1], 2)
    t = ct(t, [-1, 1, -3, -4, -5], ot.twist(wT, 1), [-2, 1], 2)
    t = ct(t, [-1, -2, 1, -4, -5], wR, [1, -3], 2)
    return ct(t, [-1, -2, -3, 1, -5], wB, [1, -4], 2)


class WeightedPEPS:
    def __init__(self, A, D):
        self.D = D
        self.Gs = [a.copy() for a in A]
        self.whs = [identity(a.sym, [a.unc_cod[0]], [a.unc_cod[0]]) for a in A]
        self.wvs = [identity(a.sym, [a.unc_dom[1]], [a.unc_dom[1]]) for a in A]

    def A(self):
        out = [None] * (L * L)
        for x in range(L):
            for y in range(L):
                out[idx(x, y)] = apply_weights(self.Gs[idx(x, y)], sqrt_w(self.whs[idx(x - 1, y)]), sqrt_w(self.wvs[idx(x, y)]),
                                               sqrt_w(self.whs[idx(x, y)]), sqrt_w(self.wvs[idx(x, y + 1)]))
        return out


def _normalized_weight(S):
    m = ot.max_norm(S)
    assert m > 0.0
    return scaled(S, 1.0 / m)


class TimePropagator:
    def __init__(self, Hh, Hv, dt):
        self.dt = dt
        self.Uh = [mexp(h, -dt) for h in Hh]
        self.Uv = [mexp(h, -dt) for h in Hv]
        self.spectrum_h = [None] * (L * L)
        self.spectrum_v = [None] * (L * L)

    def t_step(self, psi: WeightedPEPS):
        n = L * L
        for i in range(n):
            self.t_step_h(psi, i % L, i // L)
        for i in range(n):
            self.t_step_v(psi, i % L, i // L)
        for i in reversed(range(n)):
            self.t_step_v(psi, i % L, i // L)
        for i in reversed(range(n)):
            self.t_step_h(psi, i % L, i // L)

    def t_step_h(self, psi: WeightedPEPS, x, y):
        G, wh, wv = psi.Gs, psi.whs, psi.wvs
        left_full = apply_weights(G[idx(x, y)], wh[idx(x - 1, y)], wv[idx(x, y)], sqrt_w(wh[idx(x, y)]), wv[idx(x, y + 1)])
        right_full = apply_weights(G[idx(x + 1, y)], sqrt_w(wh[idx(x, y)]), wv[idx(x + 1, y)], wh[idx(x + 1, y)], wv[idx(x + 1, y + 1)])
        left, tmp_l, bond_l = ot.tsvd(ot.permute(left_full, -1, [0, 1, 3, 2, 4]), ALL, 0.0)[:3]
        bond_l = ot.mul(tmp_l, bond_l)
        bond_r, tmp_r, right = ot.tsvd(ot.permute(right_full, 0, [0, 4, 1, 2, 3]), ALL, 0.0)[:3]
        bond_r = ot.mul(bond_r, tmp_r)
        bond = ct(bond_l, [-1, 1, -2], bond_r, [1, -3, -4], 2)
        U = ot.twist(ot.twist(self.Uh[idx(x, y)], 0), 1)
        enlarged = ct(U, [1, 2, -1, -3], bond, [-2, 1, 2, -4], 2)
        bond_lp, weight, bond_rp = ot.tsvd(enlarged, psi.D, 1e-14)[:3]
        self.spectrum_h[idx(x, y)] = weight
        wh[idx(x, y)] = _normalized_weight(weight)
        g = ct(left, [-1, -2, -4, 1], bond_lp, [-5, 1, -3], 2)
        G[idx(x, y)] = apply_weights(g, inv_w(wh[idx(x - 1, y)]), inv_w(wv[idx(x, y)]), identity_on(wh[idx(x, y)]), inv_w(wv[idx(x, y + 1)]))
        g = ct(bond_rp, [-1, -5, 1], right, [1, -2, -3, -4], 2)
        G[idx(x + 1, y)] = apply_weights(g, identity_on(wh[idx(x, y)]), inv_w(wv[idx(x + 1, y)]), inv_w(wh[idx(x + 1, y)]), inv_w(wv[idx(x + 1, y + 1)]))

    def t_step_v(self, psi: WeightedPEPS, x, y):
        G, wh, wv = psi.Gs, psi.whs, psi.wvs
        top_full = apply_weights(G[idx(x, y - 1)], wh[idx(x - 1, y - 1)], wv[idx(x, y - 1)], wh[idx(x, y - 1)], sqrt_w(wv[idx(x, y)]))
        bottom_full = apply_weights(G[idx(x, y)], wh[idx(x - 1, y)], sqrt_w(wv[idx(x, y)]), wh[idx(x, y)], wv[idx(x, y + 1)])
        top, tmp_t, bond_t = ot.tsvd(ot.permute(top_full, -1, [0, 1, 2, 3, 4]), ALL, 0.0)[:3]
        bond_t = ot.mul(tmp_t, bond_t)
        bond_b, tmp_b, bottom = ot.tsvd(ot.permute(bottom_full, 0, [1, 4, 0, 2, 3]), ALL, 0.0)[:3]
        bond_b = ot.mul(bond_b, tmp_b)
        bond = ct(bond_t, [-1, 1, -2], bond_b, [1, -3, -4], 2)
        U = ot.twist(ot.twist(self.Uv[idx(x, y - 1)], 0), 1)
        enlarged = ct(U, [1, 2, -1, -3], bond, [-2, 1, 2, -4], 2)
        bond_tp, weight, bond_bp = ot.tsvd(enlarged, psi.D, 1e-14)[:3]
        self.spectrum_v[idx(x, y)] = weight
        wv[idx(x, y)] = _normalized_weight(weight)
        g = ct(top, [-1, -2, -3, 1], bond_tp, [-5, 1, -4], 2)
        G[idx(x, y - 1)] = apply_weights(g, inv_w(wh[idx(x - 1, y - 1)]), inv_w(wv[idx(x, y - 1)]), inv_w(wh[idx(x, y - 1)]), identity_on(wv[idx(x, y)]))
        g = ct(bond_bp, [-2, -5, 1], bottom, [1, -1, -3, -4], 2)
        G[idx(x, y)] = apply_weights(g, inv_w(wh[idx(x - 1, y)]), identity_on(wv[idx(x, y)]), inv_w(wh[idx(x, y)]), inv_w(wv[idx(x, y + 1)]))


def heisenberg_bond(sym, phys, Jxy=1.0, Jz=1.0):
    """Jz Sz.Sz + Jxy/2 (S+S- + S-S+) (Heisenberg.hpp:50-53) as Tensor<2,2>(p, p | p, p) on a U(1)
    leg with labels 2 Sz."""
    H = ot.Tensor(sym, [phys, phys], [phys, phys])
    for i, q in enumerate(H.sector):
        b = np.zeros((H.rows[i], H.cols[i]))
        r0 = 0
        for td in H.dom.tree(q):
            c0 = 0
            for tc in H.cod.tree(q):
                if td.dim == 1 and tc.dim == 1:
                    def sz(basis, lab):
                        return 0.5 * (-lab[0] if basis.conj_flag else lab[0])
                    c1, c2 = sz(H.unc_dom[0], td.unc[0]), sz(H.unc_dom[1], td.unc[1])
                    a1, a2 = sz(H.unc_cod[0], tc.unc[0]), sz(H.unc_cod[1], tc.unc[1])
                    v = 0.0
                    if c1 == a1 and c2 == a2:
                        v = Jz * a1 * a2
                    if (c1 == a1 + 1 and c2 == a2 - 1) or (c1 == a1 - 1 and c2 == a2 + 1):
                        v = 0.5 * Jxy
                    b[r0, c0] = v
                c0 += tc.dim
            r0 += td.dim
        H.set_block(i, b)
    return H
