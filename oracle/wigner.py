"""Exact Wigner 3j/6j/9j symbols (oracle).

The reference calls GSL (include/Xped/Symmetry/SU2Wrappers.hpp:66-82) with doubled arguments
2j = q - 1.  GSL is not available here, so the oracle evaluates the Racah single-sum formulas
exactly in rational arithmetic (Edmonds, Angular Momentum in Quantum Mechanics, eqs. 3.7.3,
6.3.7, 6.4.3) and rounds the final value once.  The committed sympy fixtures pin these.
"""
from __future__ import annotations

from decimal import Decimal, getcontext
from fractions import Fraction
from functools import lru_cache
from math import factorial

getcontext().prec = 60


def _f(n: int) -> int:
    return factorial(n)


def _delta2(ta: int, tb: int, tc: int):
    s1, s2, s3, s4 = ta + tb - tc, ta - tb + tc, -ta + tb + tc, ta + tb + tc
    if s1 < 0 or s2 < 0 or s3 < 0 or s4 % 2:
        return None
    return Fraction(_f(s1 // 2) * _f(s2 // 2) * _f(s3 // 2), _f(s4 // 2 + 1))


def _to_float(s: Fraction, p: Fraction) -> float:
    """s * sqrt(p) rounded once to binary64."""
    if s == 0 or p == 0:
        return 0.0
    ds = Decimal(s.numerator) / Decimal(s.denominator)
    dp = (Decimal(p.numerator) / Decimal(p.denominator)).sqrt()
    return float(ds * dp)


@lru_cache(maxsize=None)
def wigner_3j_exact(tj1, tj2, tj3, tm1, tm2, tm3):
    """(sum, radicand) with 3j = sum * sqrt(radicand); doubled arguments."""
    if min(tj1, tj2, tj3) < 0 or tm1 + tm2 + tm3 != 0:
        return Fraction(0), Fraction(0)
    if abs(tm1) > tj1 or abs(tm2) > tj2 or abs(tm3) > tj3:
        return Fraction(0), Fraction(0)
    if (tj1 + tm1) % 2 or (tj2 + tm2) % 2 or (tj3 + tm3) % 2:
        return Fraction(0), Fraction(0)
    d = _delta2(tj1, tj2, tj3)
    if d is None:
        return Fraction(0), Fraction(0)
    a1 = (tj3 - tj2 + tm1) // 2
    a2 = (tj3 - tj1 - tm2) // 2
    a3 = (tj1 + tj2 - tj3) // 2
    a4 = (tj1 - tm1) // 2
    a5 = (tj2 + tm2) // 2
    kmin = max(0, -a1, -a2)
    kmax = min(a3, a4, a5)
    s = Fraction(0)
    for k in range(kmin, kmax + 1):
        term = Fraction(1, _f(k) * _f(a1 + k) * _f(a2 + k) * _f(a3 - k) * _f(a4 - k) * _f(a5 - k))
        s += -term if k % 2 else term
    rad = d * _f((tj1 + tm1) // 2) * _f((tj1 - tm1) // 2) * _f((tj2 + tm2) // 2) * _f((tj2 - tm2) // 2) \
        * _f((tj3 + tm3) // 2) * _f((tj3 - tm3) // 2)
    if ((tj1 - tj2 - tm3) // 2) % 2:
        s = -s
    return s, rad


def wigner_3j(tj1, tj2, tj3, tm1, tm2, tm3) -> float:
    return _to_float(*wigner_3j_exact(tj1, tj2, tj3, tm1, tm2, tm3))


@lru_cache(maxsize=None)
def wigner_6j_exact(tj1, tj2, tj3, tj4, tj5, tj6):
    if min(tj1, tj2, tj3, tj4, tj5, tj6) < 0:
        return Fraction(0), Fraction(0)
    ds = [_delta2(tj1, tj2, tj3), _delta2(tj1, tj5, tj6), _delta2(tj4, tj2, tj6), _delta2(tj4, tj5, tj3)]
    if any(d is None for d in ds):
        return Fraction(0), Fraction(0)
    a = [(tj1 + tj2 + tj3) // 2, (tj1 + tj5 + tj6) // 2, (tj4 + tj2 + tj6) // 2, (tj4 + tj5 + tj3) // 2]
    b = [(tj1 + tj2 + tj4 + tj5) // 2, (tj2 + tj3 + tj5 + tj6) // 2, (tj3 + tj1 + tj6 + tj4) // 2]
    s = Fraction(0)
    for t in range(max(a), min(b) + 1):
        den = 1
        for x in a:
            den *= _f(t - x)
        for y in b:
            den *= _f(y - t)
        term = Fraction(_f(t + 1), den)
        s += -term if t % 2 else term
    rad = ds[0] * ds[1] * ds[2] * ds[3]
    return s, rad


def wigner_6j(tj1, tj2, tj3, tj4, tj5, tj6) -> float:
    return _to_float(*wigner_6j_exact(tj1, tj2, tj3, tj4, tj5, tj6))


def wigner_9j(tj1, tj2, tj3, tj4, tj5, tj6, tj7, tj8, tj9) -> float:
    """Sum over x of (2x+1)(-1)^{2x} {j1 j4 j7; j8 j9 x}{j2 j5 j8; j4 x j6}{j3 j6 j9; x j1 j2},
    accumulated in Decimal and rounded once."""
    xmin = max(abs(tj1 - tj9), abs(tj4 - tj8), abs(tj2 - tj6))
    xmax = min(tj1 + tj9, tj4 + tj8, tj2 + tj6)
    acc = Decimal(0)
    for tx in range(xmin, xmax + 1, 2):
        vals = []
        for args in ((tj1, tj4, tj7, tj8, tj9, tx), (tj2, tj5, tj8, tj4, tx, tj6), (tj3, tj6, tj9, tx, tj1, tj2)):
            s, r = wigner_6j_exact(*args)
            if s == 0:
                vals = None
                break
            vals.append((s, r))
        if vals is None:
            continue
        s = vals[0][0] * vals[1][0] * vals[2][0]
        r = vals[0][1] * vals[1][1] * vals[2][1]
        term = (Decimal(s.numerator) / Decimal(s.denominator)) * (Decimal(r.numerator) / Decimal(r.denominator)).sqrt()
        term *= (tx + 1)
        acc += -term if tx % 2 else term
    return float(acc)
