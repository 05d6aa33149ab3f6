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
