"""Reverse-pass restatements of the AD tensor (oracle, test infrastructure only).

Restates, on the CPU in float64:

* Tensor::permute_adj (src/Core/Tensor.cpp:258-333);
* the operator* reverse pass (include/Xped/AD/ADTensor.hpp:745-746), per matched block;
* the full_svd_impl reverse pass (ADTensor.hpp:262-320) per block, in the reference's own form
  with the explicit projectors (1 - U U^T) and (1 - V^T V) and the 1e-12 regularised F
  (:279-282);
* the cutoff_matrices reverse pass (ADTensor.hpp:357-370).

Pinned by (tests/test_oracle_ad.py): the transpose identity <permute(X), Y> = <X, permute_adj(Y)>
against the oracle's forward permute for U(1), SU(2) and fermionic U(1)xU(1); central finite
differences of gauge-invariant functions of numpy's SVD (tall, wide and square blocks) for the
SVD adjoint; finite differences of <W, A B> for the product adjoint.
"""
from __future__ import annotations

import numpy as np

from .structure import pair_permute
from .tensor import PERMUTE_SKIP, Tensor, subblock


def inverse_permutation(p):
    inv = [0] * len(p)
    for k, x in enumerate(p):
        inv[x] = k
    return inv


def permute_adj(y: Tensor, shift: int, p) -> Tensor:
    """Y.permute_adj<shift, p...>() (Tensor.cpp:258-333): legs from p^-1 (:270-288), then for
    every output tree pair the forward images in Y, inverse-shuffled and scaled (:293-330)."""
    p = list(p)
    N = len(p)
    Ry = y.rank
    if shift == 0 and p == list(range(N)):
        return y.copy()
    pinv = inverse_permutation(p)
    newR = Ry + shift
    nd = [y.unc_cod[pinv[i] - Ry].conj() if pinv[i] >= Ry else y.unc_dom[pinv[i]] for i in range(newR)]
    nc = [y.unc_cod[pinv[i + newR] - Ry] if pinv[i + newR] >= Ry else y.unc_dom[pinv[i + newR]].conj() for i in range(N - newR)]
    out = Tensor(y.sym, nd, nc)
    for i, q in enumerate(out.sector):
        blk = out.block(i)
        for t1 in out.dom.tree(q):
            r0 = out.dom.left_offset(t1)
            for t2 in out.cod.tree(q):
                c0 = out.cod.left_offset(t2)
                dims = tuple(t1.dims) + tuple(t2.dims)
                acc = np.zeros(dims)
                for (a, b), c in pair_permute(y.sym, shift, t1, t2, p).items():
                    if abs(c) < PERMUTE_SKIP:
                        continue
                    j = y.find(a.coupled)
                    sub = subblock(y, j, a, b)
                    acc = acc + c * (np.transpose(sub, pinv) if sub.ndim else sub)
                blk[r0:r0 + t1.dim, c0:c0 + t2.dim] += np.reshape(acc, (t1.dim, t2.dim), order="F")
        out.set_block(i, blk)
    return out


def mul_bwd(A: Tensor, B: Tensor, dC: Tensor, dA: Tensor, dB: Tensor):
    """dA += dC B^T and dB += A^T dC per matched sector (ADTensor.hpp:745-746), in place."""
    for i, q in enumerate(dC.sector):
        ja, jb = A.find(q), B.find(q)
        if ja < 0 or jb < 0:
            continue
        g = dC.block(i)
        dA.set_block(ja, dA.block(ja) + g @ B.block(jb).T)
        dB.set_block(jb, dB.block(jb) + A.block(ja).T @ g)


def f_regularised(d):
    """1/d, or d / (d^2 + 1e-12) where |d| < 1e-12 (ADTensor.hpp:281-282)."""
    d = np.asarray(d, dtype=np.float64)
    small = np.abs(d) < 1.0e-12
    out = np.empty_like(d)
    out[small] = d[small] / (d[small] * d[small] + 1.0e-12)
    out[~small] = 1.0 / d[~small]
    return out


def svd_bwd_block(U, s, Vd, dU, dS, dVd):
    """The adjoint of one block A = U diag(s) Vd (ADTensor.hpp:276-316, real scalars)."""
    s = np.asarray(s, dtype=np.float64)
    S = np.diag(s)
    Sinv = np.diag(1.0 / s)
    s2 = s * s
    F = f_regularised(s2[None, :] - s2[:, None])  # F_inv(i, j) = s_j^2 - s_i^2 (vec_diff, PlainInterface_Eigen_impl.cpp:82-85)
    J = F * (U.T @ dU)
    K = F * (Vd @ dVd.T)
    dA = U @ (dS + (J + J.T) @ S + S @ (K + K.T)) @ Vd
    dA = dA + (np.eye(U.shape[0]) - U @ U.T) @ dU @ Sinv @ Vd
    dA = dA + U @ Sinv @ dVd @ (np.eye(Vd.shape[1]) - Vd.T @ Vd)
    return dA


def svd_bwd(A: Tensor, U: Tensor, S: Tensor, Vd: Tensor, dU: Tensor, dS: Tensor, dVd: Tensor) -> Tensor:
    """full_svd_impl reverse pass: the adjoint of A (a fresh tensor with A's structure)."""
    dA = Tensor(A.sym, A.unc_dom, A.unc_cod, sectors=A.sector)
    for i, q in enumerate(A.sector):
        js = S.find(q)
        if js < 0:
            continue
        ju, jv = U.find(q), Vd.find(q)
        s = np.diag(S.block(js)).copy()
        dA.set_block(i, svd_bwd_block(U.block(ju), s, Vd.block(jv), dU.block(ju), dS.block(js), dVd.block(jv)))
    return dA


def cutoff_bwd(dUt: Tensor, dSt: Tensor, dVt: Tensor, dU: Tensor, dS: Tensor, dVd: Tensor):
    """Adds the truncated adjoints into the leading columns / corner / rows (ADTensor.hpp:357-370)."""
    for i, q in enumerate(dUt.sector):
        ju, js, jv = dU.find(q), dS.find(q), dVd.find(q)
        a = dUt.block(i)
        k = a.shape[1]
        bu = dU.block(ju)
        bu[:, :k] += a
        dU.set_block(ju, bu)
        bs = dS.block(js)
        bs[:k, :k] += dSt.block(dSt.find(q))
        dS.set_block(js, bs)
        bv = dVd.block(jv)
        bv[:k, :] += dVt.block(dVt.find(q))
        dVd.set_block(jv, bv)
