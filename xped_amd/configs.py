"""Synthetic workloads of the benchmark and the parity tests (leg bases and seeds).

Each builder returns plain (q, degeneracy) lists so that the product tensors and the oracle's
tensors can be made from identical legs.

C1  U(1) MPS two-site: A(l,p|r) B(r,p'|r'), virtual charges -4..4 with
    deg = max(1, round(64 exp(-q^2/4) / Z)) (sums to 66 as written), physical {-1:1, +1:1}.
C3  U(1)xU(1) fermionic grouped contraction: three virtual legs of total dimension 512 over
    sectors (N, 2Sz), |N|, |2Sz| <= 6, N = 2Sz mod 2, degeneracies from a seeded power law.
C2  SU(2) double layer: site tensor A(l, t | r, b, p) with virtual multiplets {spin 0: 2,
    spin 1/2: 1, spin 1: 1} (q = 2j + 1 labels {1: 2, 2: 1, 3: 1}; full dimension 7, as written,
    where the workload text says D = 6), physical spin 1/2, N(0, 1) reduced entries; the
    environment leg of chi = 100 (full dimension) holds spins 0..2 with the fixed multiplicities
    {0: 10, 1/2: 15, 1: 10, 3/2: 5, 2: 2} (the workload leaves them open).
C4  U(1) iPEPS: D = 8 over charges -3..3 ({+-3:1, +-2:1, +-1:1, 0:2}), physical spin-1/2.
"""
from __future__ import annotations

import math
import random
from typing import List, Tuple

Entries = List[Tuple[tuple, int]]


def c1_virtual(chi: int = 64) -> Entries:
    qs = list(range(-4, 5))
    w = [math.exp(-q * q / 4.0) for q in qs]
    Z = sum(w)
    return [((q,), max(1, int(round(chi * x / Z)))) for q, x in zip(qs, w)]


def c1_physical() -> Entries:
    return [((-1,), 1), ((1,), 1)]


def c3_leg(seed: int, total: int = 512, exponent: float = 1.5) -> Entries:
    r = random.Random(seed)
    sectors = [(n, sz) for n in range(-6, 7) for sz in range(-6, 7) if (n - sz) % 2 == 0]
    w = [(1.0 + r.random() * 9.0) ** (-exponent) for _ in sectors]
    Z = sum(w)
    if total < len(sectors):
        raise ValueError(f"a leg of dimension {total} cannot hold {len(sectors)} sectors of degeneracy >= 1")
    degs = [max(1, int(round(total * x / Z))) for x in w]
    # renormalise to the exact total: add to / take from the largest sectors (never below 1)
    diff = total - sum(degs)
    order = sorted(range(len(degs)), key=lambda i: -degs[i])
    k = 0
    guard = 0
    while diff != 0:
        i = order[k % len(order)]
        step = 1 if diff > 0 else -1
        if degs[i] + step >= 1:
            degs[i] += step
            diff -= step
        k += 1
        guard += 1
        if guard > 100 * total:
            raise RuntimeError("degeneracy renormalisation did not terminate")
    return [((n, sz), d) for (n, sz), d in zip(sectors, degs)]


def c2_virtual() -> Entries:
    return [((1,), 2), ((2,), 1), ((3,), 1)]


def c2_physical() -> Entries:
    return [((2,), 1)]


def c2_env(chi: int = 100) -> Entries:
    """SU(2) environment leg; chi = 100 is the workload's (full dimension 100), chi = 9 a reduced
    one for parity tests against the oracle."""
    table = {100: [((1,), 10), ((2,), 15), ((3,), 10), ((4,), 5), ((5,), 2)], 9: [((1,), 2), ((2,), 2), ((3,), 1)]}
    if chi not in table:
        raise ValueError("c2_env: chi must be 100 (the workload) or 9 (reduced)")
    return table[chi]


def c2_small_virtual() -> Entries:
    """A reduced C2 virtual leg (spins 0 and 1/2, full dimension 3) for oracle parity tests."""
    return [((1,), 1), ((2,), 1)]


def c4_virtual() -> Entries:
    return [((-3,), 1), ((-2,), 1), ((-1,), 1), ((0,), 2), ((1,), 1), ((2,), 1), ((3,), 1)]


def c4_physical() -> Entries:
    return [((-1,), 1), ((1,), 1)]


def c4_small_virtual() -> Entries:
    """D = 3 over charges -1..1 (a reduced C4 shape for parity tests)."""
    return [((-1,), 1), ((0,), 1), ((1,), 1)]


def ipeps_site_legs(virtual: Entries, physical: Entries):
    """Leg entries of the four Tensor<2,3>(left, top | right, bottom, phys) site tensors of a 2x2
    cell (order (0,0), (0,1), (1,0), (1,1)) with every virtual bond carrying `virtual`."""
    return [[virtual, virtual, virtual, virtual, physical] for _ in range(4)]
