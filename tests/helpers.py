"""Shared builders: identical legs for the product (xped_amd) and the oracle."""
import random

import numpy as np

import xped_amd.tensor as xt
from oracle import tensor as oten
from oracle.structure import Qbasis as OQbasis
from oracle.symmetry import FU1, SU2, U1, Sym

KINDS = {xt.U1: (U1,), xt.SU2: (SU2,), xt.FU1xU1: (FU1, U1), xt.FU1: (FU1,)}


def both_basis(pcode, entries):
    o = OQbasis(Sym(*KINDS[pcode]), 1)
    p = xt.Qbasis(pcode)
    for q, d in entries:
        q = tuple(q) if isinstance(q, (tuple, list)) else (q,)
        o.push_back(q, d)
        p.push_back(q, d)
    return o, p


def rand_entries(pcode, r, nsec, maxdim, mindim=1):
    out = []
    for _ in range(nsec):
        if pcode == xt.SU2:
            q = (r.randint(1, 4),)
        elif pcode[1]:
            q = (r.randint(-3, 3), r.randint(-2, 2))
        else:
            q = (r.randint(-4, 4),)
        out.append((q, r.randint(mindim, maxdim)))
    return out


def both_tensor(rt, pcode, legs_entries, rank, seed, dist=0):
    """legs_entries: list of entry lists; returns (oracle tensor, product tensor) with the same data."""
    pairs = [both_basis(pcode, e) for e in legs_entries]
    ob = [p[0] for p in pairs]
    pb = [p[1] for p in pairs]
    o = oten.Tensor(Sym(*KINDS[pcode]), ob[:rank], ob[rank:]).set_random(seed, dist)
    t = xt.Tensor(rt, pb[:rank], pb[rank:])
    t.upload(o.data)
    return o, t


def structure_of(t):
    return [(s[0], s[1], s[2], s[3]) for s in t.sectors()]


def oracle_structure(o):
    return [(tuple(q), off, r, c) for q, off, r, c in o.structure()]


def make_ipeps(rt, virtual, physical, seed):
    """Four 2x2-cell site tensors A(left, top | right, bottom, phys) with N(0,1)/sqrt(size) entries
    drawn from one mt19937(seed) in site order (0,0), (0,1), (1,0), (1,1)."""
    from oracle import rng
    from xped_amd import configs

    eng = rng.MT19937(seed)
    nd = rng.Normal()
    oA, pA = [], []
    for legs in configs.ipeps_site_legs(virtual, physical):
        pairs = [both_basis(xt.U1, e) for e in legs]
        o = oten.Tensor(Sym(U1), [p[0] for p in pairs[:2]], [p[0] for p in pairs[2:]])
        o.data = nd.sample(eng, o.size) / np.sqrt(o.size)
        t = xt.Tensor(rt, [p[1] for p in pairs[:2]], [p[1] for p in pairs[2:]])
        t.upload(o.data)
        oA.append(o)
        pA.append(t)
    return oA, pA


def make_neel(rt):
    """D = 1 Neel product state on the 2x2 cell (spins +1, -1, -1, +1 in units of 2Sz); virtual
    charges chosen so that every site conserves U(1) charge.  Energy per site of the
    nearest-neighbour Heisenberg model: exactly -1/2."""
    spin = {(0, 0): 1, (1, 0): -1, (0, 1): -1, (1, 1): 1}
    h = {(0, 0): 0, (1, 0): 1, (0, 1): 0, (1, 1): -1}
    oA, pA = [], []
    for x in range(2):
        for y in range(2):
            legs = [[((h[((x - 1) % 2, y)],), 1)], [((0,), 1)], [((h[(x, y)],), 1)], [((0,), 1)], [((-1,), 1), ((1,), 1)]]
            pairs = [both_basis(xt.U1, e) for e in legs]
            o = oten.Tensor(Sym(U1), [p[0] for p in pairs[:2]], [p[0] for p in pairs[2:]])
            for i, q in enumerate(o.sector):
                blk = o.block(i).copy()
                c0 = 0
                for tc in o.cod.tree(q):
                    if tc.unc[2] == (spin[(x, y)],):
                        blk[0, c0] = 1.0
                    c0 += tc.dim
                o.set_block(i, blk)
            t = xt.Tensor(rt, [p[1] for p in pairs[:2]], [p[1] for p in pairs[2:]])
            t.upload(o.data)
            oA.append(o)
            pA.append(t)
    return oA, pA


def make_dimer(rt, orientation="h", alpha=-1.0, beta=1.0):
    """Columnar dimer product state on the 2x2 cell, as a U(1) iPEPS with D = 2 on the dimer
    bonds and D = 1 elsewhere.  orientation "h": dimers on the bonds (0, y)-(1, y); "v": on
    (x, 0)-(x, 1).  The first site of a dimer sends its spin through the dimer bond (virtual
    charge -2Sz), the second site receives it with amplitude alpha (2Sz = -1) or beta (+1), so each
    dimer is alpha |up down> + beta |down up> (the singlet for alpha = -beta).  Exact bond energies
    in dimer_bond_energies.  rt None: oracle tensors only."""
    spin = [((-1,), 1), ((1,), 1)]
    dimer = [((-1,), 1), ((1,), 1)]
    none = [((0,), 1)]
    oA, pA = [], []
    for x in range(2):
        for y in range(2):
            first = (x == 0) if orientation == "h" else (y == 0)
            if orientation == "h":
                left, right = (none, dimer) if first else (dimer, none)
                top = bottom = none
            else:
                top, bottom = (none, dimer) if first else (dimer, none)
                left = right = none
            legs = [left, top, right, bottom, spin]
            pairs = [both_basis(xt.U1, e) for e in legs] if rt is not None else None
            obs = [p[0] for p in pairs] if pairs else [_obasis(e) for e in legs]
            o = oten.Tensor(Sym(U1), obs[:2], obs[2:])
            for i, q in enumerate(o.sector):
                blk = o.block(i).copy()
                r0 = 0
                for td in o.dom.tree(q):
                    c0 = 0
                    for tc in o.cod.tree(q):
                        l, t = td.unc[0][0], td.unc[1][0]
                        r, b, s = tc.unc[0][0], tc.unc[1][0], tc.unc[2][0]
                        vin, vout = (l, r) if orientation == "h" else (t, b)
                        if first and vin == 0 and vout == -s:
                            blk[r0, c0] = 1.0
                        if not first and vout == 0 and vin == s:
                            blk[r0, c0] = alpha if s == -1 else beta
                        c0 += tc.dim
                    r0 += td.dim
                o.set_block(i, blk)
            oA.append(o)
            if rt is not None:
                t = xt.Tensor(rt, [p[1] for p in pairs[:2]], [p[1] for p in pairs[2:]])
                t.upload(o.data)
                pA.append(t)
    return oA, pA


def dimer_bond_energies(orientation="h", alpha=-1.0, beta=1.0):
    """<S_i.S_j> of the 4 horizontal then 4 vertical bonds (index x * 2 + y of the bond's first
    site) of make_dimer's state: -1/4 + alpha beta / N on a dimer (N = alpha^2 + beta^2); between
    dimers, whose total 2Sz is 0 so that <S+> = <S-> = 0, the product m_i m_j of the sites'
    magnetisations, m = (alpha^2 - beta^2) / (2N) on the first site of a dimer and -m on the
    second."""
    N = alpha * alpha + beta * beta
    ed = -0.25 + alpha * beta / N
    m = 0.5 * (alpha * alpha - beta * beta) / N
    if orientation == "h":  # h bonds: x = 0 dimer, x = 1 second-first; v bonds: same x
        return [ed, ed, -m * m, -m * m] + [m * m] * 4
    return [m * m] * 4 + [ed, -m * m, ed, -m * m]


def _obasis(entries):
    b = OQbasis(Sym(U1), 1)
    for q, d in entries:
        b.push_back(q, d)
    return b


def oracle_bond_energies(oc, Jxy=1.0, Jz=1.0):
    """<H_bond> of the 4 horizontal then 4 vertical bonds from the oracle CTM's density matrices
    (the terms that CTM.energy sums)."""
    from oracle import ctm as octm

    return [oten.trace(octm.ct(octm.tw(rho, [0, 1]), [1, 2, 3, 4], oc._gate(rho, Jxy, Jz), [3, 4, 1, 2], 0)) for rho in oc.rho_h + oc.rho_v]


def oracle_ipeps(virtual, physical, seed):
    """Oracle-only site tensors of make_ipeps (same data)."""
    from oracle import rng
    from xped_amd import configs

    eng = rng.MT19937(seed)
    nd = rng.Normal()
    out = []
    for legs in configs.ipeps_site_legs(virtual, physical):
        obs = [_obasis(e) for e in legs]
        o = oten.Tensor(Sym(U1), obs[:2], obs[2:])
        o.data = nd.sample(eng, o.size) / np.sqrt(o.size)
        out.append(o)
    return out
