"""Bit-exact restatement of the reference's random data generation (oracle).

Tensor::setRandom (src/Core/Tensor.cpp:32-43) -> MatrixInterface::setRandom
(src/Interfaces/MatrixInterface_Eigen_impl.cpp:51-64) draws
std::uniform_real_distribution<double>(-1, 1) from a std::mt19937, block by block in storage
order, column-major inside a block.  libstdc++ implements the distribution through
std::generate_canonical<double, 53>, which consumes two 32-bit outputs per double:
u = (r1 + r2 * 2^32) / 2^64 in double arithmetic, then x = u * (b - a) + a.
N(0,1) data (not a reference facility, used by the PEPS configurations) follows libstdc++'s
std::normal_distribution (Marsaglia polar method; the second variate of each pair is cached).
"""
from __future__ import annotations

import math

import numpy as np


class MT19937:
    """std::mt19937 seeded with an integer (init_genrand), raw 32-bit outputs."""

    def __init__(self, seed: int = 5489):
        self._bg = np.random.MT19937()
        self._bg._legacy_seeding(int(seed) & 0xFFFFFFFF)
        self._buf = np.empty(0, dtype=np.uint64)
        self._pos = 0

    def raw(self, n: int) -> np.ndarray:
        return self._bg.random_raw(n).astype(np.uint64)

    def __call__(self) -> int:
        return int(self._bg.random_raw())


def canonical(eng: MT19937, n: int) -> np.ndarray:
    r = eng.raw(2 * n).astype(np.float64).reshape(n, 2)
    s = r[:, 0] + r[:, 1] * 4294967296.0
    u = s / 18446744073709551616.0
    u[u >= 1.0] = np.nextafter(1.0, 0.0)
    return u


def uniform(eng: MT19937, n: int, a: float = -1.0, b: float = 1.0) -> np.ndarray:
    return canonical(eng, n) * (b - a) + a


class Normal:
    """std::normal_distribution<double>(mean, stddev) as implemented in libstdc++."""

    def __init__(self, mean: float = 0.0, stddev: float = 1.0):
        self.mean, self.stddev = mean, stddev
        self.saved = None

    def draw(self, eng: MT19937) -> float:
        if self.saved is not None:
            ret = self.saved
            self.saved = None
        else:
            while True:
                x = 2.0 * float(canonical(eng, 1)[0]) - 1.0
                y = 2.0 * float(canonical(eng, 1)[0]) - 1.0
                r2 = x * x + y * y
                if not (r2 > 1.0 or r2 == 0.0):
                    break
            mult = math.sqrt(-2 * math.log(r2) / r2)
            self.saved = x * mult
            ret = y * mult
        return ret * self.stddev + self.mean

    def sample(self, eng: MT19937, n: int) -> np.ndarray:
        return np.array([self.draw(eng) for _ in range(n)], dtype=np.float64)
