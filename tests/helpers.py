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
