"""Symmetry algebra (oracle): U(1), fermionic U(1), SU(2) and their products.

Follows include/Xped/Symmetry/U1.hpp:95-226, SU2.hpp:108-177, src/Symmetry/SU2.cpp:18-119,
SymBase.hpp:19-102 and S1xS2.hpp:105-308 (products of per-component coefficients; component 0
is the outer loop of basis_combine).  Quantum numbers are tuples of ints.
"""
from __future__ import annotations

import math

from .wigner import wigner_6j

U1, FU1, SU2 = 1, 2, 3


def phase(x: int) -> float:
    """functions.hpp:79-84 -- (-1)^x with C++ remainder semantics."""
    return -1.0 if math.fmod(x, 2) != 0 else 1.0


class Sym:
    def __init__(self, *kinds: int):
        self.kinds = tuple(k for k in kinds if k)
        self.nq = len(self.kinds)

    def __eq__(self, other):
        return isinstance(other, Sym) and self.kinds == other.kinds

    def __hash__(self):
        return hash(self.kinds)

    @property
    def non_abelian(self) -> bool:
        return SU2 in self.kinds

    @property
    def fermionic(self) -> bool:
        return FU1 in self.kinds

    def vacuum(self):
        return tuple(1 if k == SU2 else 0 for k in self.kinds)

    def conj(self, q):
        return tuple(x if k == SU2 else -x for k, x in zip(self.kinds, q))

    def degeneracy(self, q) -> int:
        d = 1
        for k, x in zip(self.kinds, q):
            if k == SU2:
                d *= x
        return d

    def compare(self, a, b) -> bool:
        # SymBase::compare(q1, q2) = q1 < q2 (lexicographic qarray comparison)
        return tuple(a) < tuple(b)

    def _comb1(self, k, a, b):
        if k == SU2:  # SU2.cpp:18-27
            return list(range(abs(a - b) + 1, abs(a + b) - 1 + 1, 2))
        return [a + b]  # U1.hpp:189-194

    def combine(self, a, b):
        if self.nq == 0:
            return [()]
        out = [(x,) for x in self._comb1(self.kinds[0], a[0], b[0])]
        for c in range(1, self.nq):
            out = [o + (y,) for o in out for y in self._comb1(self.kinds[c], a[c], b[c])]
        return out

    @staticmethod
    def _tri1(k, a, b, c):
        if k == SU2:  # SU2.cpp:114-119
            return (c - 1 >= abs(a - b)) and (c - 1 <= a + b - 2)
        return a + b == c

    def triangle(self, a, b, c) -> bool:
        return all(self._tri1(k, x, y, z) for k, x, y, z in zip(self.kinds, a, b, c))

    def _swap1(self, k, ql, qr, qf):
        if k == SU2:  # SU2.hpp:169-177
            return phase((ql + qr - qf - 1) // 2) if self._tri1(k, ql, qr, qf) else 0.0
        sign = 1.0
        if k == FU1 and ql % 2 != 0 and qr % 2 != 0:  # U1.hpp:174-182
            sign = -1.0
        return sign if self._tri1(k, ql, qr, qf) else 0.0

    def coeff_swap(self, ql, qr, qf) -> float:
        out = 1.0
        for c, k in enumerate(self.kinds):
            out *= self._swap1(k, ql[c], qr[c], qf[c])
        return out

    def _sixj1(self, k, q1, q2, q3, q4, q5, q6):
        if k == SU2:
            return wigner_6j(q1 - 1, q2 - 1, q3 - 1, q4 - 1, q5 - 1, q6 - 1)
        t = self._tri1
        return 1.0 if (t(k, q1, q2, q3) and t(k, q1, q6, q5) and t(k, q2, q4, q6) and t(k, q3, q4, q5)) else 0.0

    def _recouple1(self, k, q1, q2, q3, Q, Q12, Q23):
        if k == SU2:  # SU2.hpp:164-167 (the phase argument is truncated to int)
            return math.sqrt(Q12 * Q23) * phase(int((q1 + q2 + q3 + Q) / 2.0)) * self._sixj1(k, q1, q2, Q12, q3, Q, Q23)
        return self._sixj1(k, q1, q2, Q12, q3, Q, Q23)  # U1.hpp:163-166

    def coeff_recouple(self, q1, q2, q3, Q, Q12, Q23) -> float:
        out = 1.0
        for c, k in enumerate(self.kinds):
            out *= self._recouple1(k, q1[c], q2[c], q3[c], Q[c], Q12[c], Q23[c])
        return out

    def _turn1(self, k, ql, qr, qf):
        if k == SU2:  # SU2.hpp:143-146
            return self._swap1(k, ql, qr, qf) * math.sqrt(qf / ql) if self._tri1(k, ql, qr, qf) else 0.0
        return 1.0 if self._tri1(k, ql, qr, qf) else 0.0  # U1.hpp:159

    def coeff_turn(self, ql, qr, qf) -> float:
        out = 1.0
        for c, k in enumerate(self.kinds):
            out *= self._turn1(k, ql[c], qr[c], qf[c])
        return out

    def coeff_FS(self, q) -> float:
        out = 1.0
        for k, x in zip(self.kinds, q):
            if k == SU2:  # SU2.cpp:35-39
                out *= -1.0 if x % 2 == 0 else 1.0
        return out

    def coeff_twist(self, q) -> float:
        out = 1.0
        for k, x in zip(self.kinds, q):
            if k == FU1:  # U1.hpp:130-134
                out *= -1.0 if x % 2 != 0 else 1.0
        return out
