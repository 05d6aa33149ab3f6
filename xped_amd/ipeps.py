"""Synthetic iPEPS site tensors for the CTMRG workloads.

The reference's iPEPS::setRandom (src/PEPS/iPEPS.cpp:343-364) fills the unit cell's site
tensors from one seeded engine with uniform entries.  The CTMRG workloads here use standard
normal entries scaled by 1/sqrt(plain size), each site drawn from std::mt19937(seed + site) in
storage order, column-major per block (the C++ engine of Tensor::setRandom, Tensor.cpp:32-43).
"""
from __future__ import annotations

import math
from typing import List

from . import configs
from .tensor import U1, Qbasis, Runtime, Tensor


def basis(entries, pcode=U1) -> Qbasis:
    b = Qbasis(pcode)
    for q, d in entries:
        b.push_back(tuple(q) if isinstance(q, (tuple, list)) else (q,), d)
    return b


def random_ipeps(rt: Runtime, virtual, physical, seed: int) -> List[Tensor]:
    """Four 2x2-cell site tensors A(left, top | right, bottom, phys), sites (0,0), (0,1), (1,0), (1,1)."""
    out = []
    for site, legs in enumerate(configs.ipeps_site_legs(virtual, physical)):
        b = [basis(e) for e in legs]
        t = Tensor(rt, b[:2], b[2:])
        t.setRandom(seed + site, 1)
        out.append(t.scale(1.0 / math.sqrt(t.plain_size())))
    return out
