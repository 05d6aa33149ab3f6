"""Pins of the reverse-pass oracle (oracle/ad.py) on the CPU.

* permute_adj: the transpose identity <permute(X), Y> = <X, permute_adj(Y)> with the oracle's
  forward permute (itself pinned against the reference's structure and coefficient tables), for
  U(1), SU(2) and fermionic U(1)xU(1), every permutation and shift of rank-3/rank-4 tensors;
  plus permute_adj(permute(X)) == X exactly for U(1) (an abelian permute is a signed permutation).
* svd_bwd_block: central finite differences of gauge-invariant functions of numpy's SVD
  (tall, wide, square), which a dropped projector term, a transposed F or a wrong sign fails.
* mul_bwd: finite differences of <W, A B>.
"""
import itertools
import random

import numpy as np
import pytest

import xped_amd.tensor as xt
from oracle import ad as oad
from oracle import tensor as oten
from oracle.symmetry import Sym

from .helpers import KINDS, both_basis, rand_entries


def otensor(pcode, legs, rank, seed):
    # sorted legs, as Xped's bases are (Qbasis::sort); conj() sorts too, so an unsorted leg that
    # crosses the domain/codomain boundary would come back from permute_adj in another order
    bases = [both_basis(pcode, e)[0] for e in legs]
    for b in bases:
        b.sort()
    return oten.Tensor(Sym(*KINDS[pcode]), bases[:rank], bases[rank:]).set_random(seed)


@pytest.mark.parametrize("pcode", [xt.U1, xt.SU2, xt.FU1xU1], ids=["U1", "SU2", "FU1xU1"])
def test_permute_adj_is_transpose(pcode):
    r = random.Random(11)
    for rank, corank in [(2, 2), (2, 1), (1, 2)]:
        legs = [rand_entries(pcode, r, 3, 3) for _ in range(rank + corank)]
        X = otensor(pcode, legs, rank, 3)
        for p in itertools.permutations(range(rank + corank)):
            for shift in range(-corank, rank + 1):
                Yx = oten.permute(X, shift, p)
                if Yx.size == 0:
                    continue
                Y = Yx.copy()
                Y.data = np.random.default_rng(len(p) + shift + 7).uniform(-1, 1, Y.size)
                A = oad.permute_adj(Y, shift, p)
                assert [tuple(q) for q in A.sector] == [tuple(q) for q in X.sector]
                lhs = float(Yx.data @ Y.data)
                rhs = float(X.data @ A.data)
                assert abs(lhs - rhs) <= 1e-12 * max(1.0, abs(lhs)), (p, shift, lhs, rhs)
                if pcode == xt.U1:
                    assert np.array_equal(oad.permute_adj(Yx, shift, p).data, X.data)


def _fd(f, A, h=1e-6):
    g = np.zeros_like(A)
    for idx in np.ndindex(A.shape):
        Ap, Am = A.copy(), A.copy()
        Ap[idx] += h
        Am[idx] -= h
        g[idx] = (f(Ap) - f(Am)) / (2 * h)
    return g


@pytest.mark.parametrize("m,n", [(7, 4), (4, 7), (5, 5)])
def test_svd_bwd_block_finite_differences(m, n):
    rng = np.random.default_rng(m * 10 + n)
    A = rng.standard_normal((m, n))
    k = min(m, n)
    W1 = rng.standard_normal((m, n))
    w2 = rng.standard_normal(k)
    W3 = rng.standard_normal((m, m))
    g = rng.uniform(0.5, 1.5, k)

    def svd(X):
        U, s, Vd = np.linalg.svd(X, full_matrices=False)
        return U, s, Vd

    # f = <W1, U Vd> + <w2, s> + <W3, U diag(g) U^T>: invariant under the sign gauge of (u_i, v_i)
    def f(X):
        U, s, Vd = svd(X)
        return float(np.sum(W1 * (U @ Vd)) + w2 @ s + np.sum(W3 * (U @ np.diag(g) @ U.T)))

    U, s, Vd = svd(A)
    dU = W1 @ Vd.T + (W3 + W3.T) @ U @ np.diag(g)
    dVd = U.T @ W1
    dS = np.diag(w2)
    got = oad.svd_bwd_block(U, s, Vd, dU, dS, dVd)
    fd = _fd(f, A)
    np.testing.assert_allclose(got, fd, rtol=0, atol=2e-6 * max(1.0, np.abs(fd).max()))


def test_svd_bwd_block_regularised_degenerate_pair():
    # equal singular values: F falls back to d / (d^2 + 1e-12) = 0 on the exact tie
    U = np.eye(3)
    s = np.array([2.0, 2.0, 1.0])
    Vd = np.eye(3)
    dU = np.zeros((3, 3))
    dVd = np.zeros((3, 3))
    dS = np.diag([1.0, 1.0, 1.0])
    got = oad.svd_bwd_block(U, s, Vd, dU, dS, dVd)
    np.testing.assert_allclose(got, np.eye(3), atol=1e-15)
    assert oad.f_regularised(np.array([0.0]))[0] == 0.0
    assert oad.f_regularised(np.array([1e-13]))[0] == pytest.approx(1e-13 / (1e-26 + 1e-12))


def test_mul_bwd_finite_differences():
    r = random.Random(5)
    legs = [rand_entries(xt.U1, r, 3, 3) for _ in range(3)]
    A = otensor(xt.U1, [legs[0], legs[1]], 1, 1)
    B = otensor(xt.U1, [legs[1], legs[2]], 1, 2)
    C = oten.mul(A, B)
    W = np.random.default_rng(3).uniform(-1, 1, C.size)

    def f_a(x):
        T = A.copy()
        T.data = x
        return float(W @ oten.mul(T, B).data)

    def f_b(x):
        T = B.copy()
        T.data = x
        return float(W @ oten.mul(A, T).data)

    dC = C.copy()
    dC.data = W.copy()
    dA, dB = A.copy(), B.copy()
    dA.data[:] = 0
    dB.data[:] = 0
    oad.mul_bwd(A, B, dC, dA, dB)
    np.testing.assert_allclose(dA.data, _fd(f_a, A.data.copy()), atol=1e-8)
    np.testing.assert_allclose(dB.data, _fd(f_b, B.data.copy()), atol=1e-8)
