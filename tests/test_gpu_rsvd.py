"""Randomized block SVD and its Cholesky-inverse building block on the GPU, judged by the full FP64
block SVD (needs an MI355X).

Tolerances (stated per test): with one or more power iterations the kept singular values agree
with the full tSVD's to 1e-12 relative to the largest and the truncated products U S Vdag to 1e-9
(one iteration) or 1e-11 (two); U and Vdag are orthonormal to 1e-12; the inverse Cholesky factor
satisfies X^T G X = I to 1e-10 on the independent columns.
"""
import ctypes as C

import math

import numpy as np
import pytest

import xped_amd.tensor as xt
from xped_amd import _lib
from xped_amd._lib import lib

from .helpers import structure_of

pytestmark = pytest.mark.gpu


def _spectral_tensor(rt, sectors, rho, seed):
    """U(1) tensor (a, b | c) whose sector blocks are Q1 diag(scale_q * rho^j) Q2^T with random
    orthogonal Q1, Q2: the exact singular values are known."""
    r = np.random.default_rng(seed)
    legs = []
    for _ in range(3):
        b = xt.Qbasis(xt.U1)
        for q, d in sectors:
            b.push_back((q,), d)
        legs.append(b)
    t = xt.Tensor(rt, legs[:2], legs[2:])
    data = np.zeros(t.size)
    exact = {}
    for q, off, rows, cols in t.sectors():
        k = min(rows, cols)
        qq = q[0] if isinstance(q, (tuple, list)) else q
        scale = 1.0 + 0.1 * qq + 0.01 * qq * qq  # distinct for +q and -q: no exact ties at the cut
        sv = scale * rho ** np.arange(k)
        Q1, _ = np.linalg.qr(r.standard_normal((rows, k)))
        Q2, _ = np.linalg.qr(r.standard_normal((cols, k)))
        block = (Q1 * sv) @ Q2.T
        data[off:off + rows * cols] = block.T.reshape(-1)
        exact[q] = sv
    t.upload(data)
    return t, exact


def _kept_ranks(S):
    return {q[0] if isinstance(q, (tuple, list)) else q: rows for q, _, rows, _ in S.sectors()}


# subspace iteration: the kept subspace is off by an angle ~ (sigma_{l+1} / sigma_k)^(2 q + 1)
# (~1.7e-4 here for q = 0), singular values by that angle squared
@pytest.mark.parametrize("power_iters,sv_tol,rec_tol", [(0, 1e-6, 1e-3), (1, 1e-12, 1e-9), (2, 1e-12, 1e-11)])
def test_randomized_matches_full_tsvd(rt, power_iters, sv_tol, rec_tol):
    t, _ = _spectral_tensor(rt, [(-2, 28), (-1, 32), (0, 36), (1, 32), (2, 28)], 0.6, 11)
    Uf, Sf, Vf, twf, _ = t.tSVD(40, 1e-10, PRESERVE_MULTIPLETS=False)
    ranks = _kept_ranks(Sf)
    U, S, V, tw, safe, ratio = t.tSVD_randomized(40, 1e-10, ranks, oversample=8, rank_factor=2.0, power_iters=power_iters, seed=5)
    assert safe, ratio
    assert structure_of(S) == structure_of(Sf)
    smax = Sf.download().max()
    # every sector is sketched with fewer columns than its full rank
    assert all(2 * k + 8 < min(rows, cols) for (q, _, rows, cols), k in zip(t.sectors(), [ranks.get(q[0], 0) for q, *_ in t.sectors()]))
    np.testing.assert_allclose(S.download(), Sf.download(), rtol=0, atol=sv_tol * smax)
    np.testing.assert_allclose(((U * S) * V).download(), ((Uf * Sf) * Vf).download(), rtol=0, atol=rec_tol * smax)
    # orthonormal columns of U and rows of Vdag
    UtU = (U.adjoint() * U).download()
    VVt = (V * V.adjoint()).download()
    eye = (Sf.adjoint() * Sf)  # same structure as the identity on the kept basis
    ident = np.zeros(eye.size)
    for _, off, rows, _ in eye.sectors():
        ident[off:off + rows * rows] = np.eye(rows).reshape(-1)
    np.testing.assert_allclose(UtU, ident, rtol=0, atol=1e-12)
    np.testing.assert_allclose(VVt, ident, rtol=0, atol=1e-12)
    # the discarded weight of the maxKeep cut is the full tSVD's, computed from fewer values
    assert tw <= twf * (1 + 1e-12) + 1e-300


def test_randomized_flags_missed_weight(rt):
    # a slowly decaying spectrum sketched with one column per sector: the missed weight exceeds
    # the smallest kept singular value and the result must not be trusted
    t, _ = _spectral_tensor(rt, [(-1, 4), (0, 6), (1, 4)], 0.95, 12)
    _, _, _, _, safe, ratio = t.tSVD_randomized(64, 1e-10, {}, oversample=1, rank_factor=0.0, power_iters=1, seed=3)
    assert not safe and ratio >= 1.0


def test_randomized_rank_deficient_sectors(rt):
    # spectra with exact zeros beyond rank 3: sketch columns beyond the rank are dependent and are
    # dropped by the Cholesky inverse; the eps cut removes their zero singular values
    t, exact = _spectral_tensor(rt, [(-1, 10), (0, 12), (1, 10)], 0.5, 13)
    data = t.download()
    r = np.random.default_rng(3)
    for q, off, rows, cols in t.sectors():
        Q1, _ = np.linalg.qr(r.standard_normal((rows, 3)))
        Q2, _ = np.linalg.qr(r.standard_normal((cols, 3)))
        data[off:off + rows * cols] = ((Q1 * np.array([3.0, 2.0, 1.0])) @ Q2.T).T.reshape(-1)
    t.upload(data)
    U, S, V, tw, safe, ratio = t.tSVD_randomized(100, 1e-10, {0: 3, 1: 3, -1: 3}, oversample=8, rank_factor=2.0, power_iters=1, seed=9)
    assert safe, ratio
    for q, off, rows, cols in S.sectors():
        assert rows == 3
        np.testing.assert_allclose(np.diag(S.download()[off:off + 9].reshape(3, 3)), [3.0, 2.0, 1.0], rtol=0, atol=1e-12)
    np.testing.assert_allclose(((U * S) * V).download(), data, rtol=0, atol=1e-12 * 3)


def test_cholesky_inverse_kernel(rt):
    L = lib()
    ctx = rt.ctx()
    r = np.random.default_rng(7)
    blocks = []
    # n = 199 is the largest block held in LDS (packed upper triangle); n = 200 and 300 run from
    # a global working copy
    for m, n in [(50, 20), (200, 130), (9, 9), (30, 6), (2, 2), (3, 3), (160, 141), (260, 180), (230, 199), (240, 200), (400, 300)]:
        Y = r.standard_normal((m, n))
        if n > 2:
            Y[:, 2] = Y[:, 0] - 2.0 * Y[:, 1]  # an exactly dependent column
        blocks.append(Y.T @ Y)
    offs, tot = [], 0
    for G in blocks:
        offs.append(tot)
        tot += G.size
    Gall = np.concatenate([G.T.reshape(-1) for G in blocks])
    dG, dX = C.c_void_p(), C.c_void_p()
    assert L.xh_alloc(ctx, 8 * tot, C.byref(dG)) == 0
    assert L.xh_alloc(ctx, 8 * tot, C.byref(dX)) == 0
    L.xh_upload(ctx, dG, Gall.ctypes.data, 8 * tot)
    items = (_lib.xh_factor_item * len(blocks))(*[_lib.xh_factor_item(o, o, 0, 0, G.shape[0], G.shape[0]) for o, G in zip(offs, blocks)])
    shift = (C.c_double * len(blocks))(*([0.0] * len(blocks)))
    assert L.xh_block_chol_inv(ctx, items, len(blocks), dG, dX, shift) == 0
    Xall = np.empty(tot)
    L.xh_download(ctx, Xall.ctypes.data, dX, 8 * tot)
    L.xh_free(ctx, dG)
    L.xh_free(ctx, dX)
    for o, G in zip(offs, blocks):
        n = G.shape[0]
        X = Xall[o:o + n * n].reshape(n, n).T
        assert np.array_equal(X, np.triu(X))
        live = [j for j in range(n) if j != 2] if n > 2 else list(range(n))
        if n > 2:
            assert np.all(X[2, :] == 0.0) and np.all(X[:, 2] == 0.0)
        XtGX = X.T @ G @ X
        np.testing.assert_allclose(XtGX[np.ix_(live, live)], np.eye(len(live)), rtol=0, atol=1e-10)


@pytest.mark.parametrize("shape,sv_tol", [((450, 400, 30, 20), 1e-12), ((1842, 1844, 1200, 1900), 4e-12)], ids=["450x400", "c4-size"])
def test_large_block_full_svd(rt, shape, sv_tol):
    # blocks with more than 64 columns take the block-Jacobi path (bjsvd.hip), up to the largest
    # projector-product blocks of the D = 8, chi = 256 CTMRG step (1842 x 1844).  Reconstruction to
    # 1e-12 (times sqrt of the size); singular values to 1e-12 sigma_max of numpy's at 450 x 400, and
    # to 4e-12 sigma_max at 1842 x 1844: the one-sided Jacobi stop (pairs orthogonal to
    # 8 sqrt(rows) sqrt(64) eps, above the rounding noise of the Gram GEMMs) leaves 1.6e-12 sigma_max
    # there (measured), where LAPACK's bidiagonal QR stays within a small multiple of eps sigma_max
    r = np.random.default_rng(21)
    m0, n0, m1, n1 = shape
    a, b = xt.Qbasis(xt.U1), xt.Qbasis(xt.U1)
    a.push_back((0,), m0)
    a.push_back((1,), m1)
    b.push_back((0,), n0)
    b.push_back((1,), n1)
    t = xt.Tensor(rt, [a], [b])
    data = r.uniform(-1, 1, t.size)
    t.upload(data)
    U, S, V, tw, _ = t.tSVD(10 ** 6, 0.0, PRESERVE_MULTIPLETS=False)
    np.testing.assert_allclose(((U * S) * V).download(), data, rtol=0, atol=1e-12 * math.sqrt(max(m0, n0, m1, n1)) * 2)
    for q, off, rows, cols in t.sectors():
        blk = data[off:off + rows * cols].reshape(cols, rows).T
        ref = np.linalg.svd(blk, compute_uv=False)
        got = [s for (qq, o2, r2, c2) in S.sectors() if qq == q for s in np.diag(S.download()[o2:o2 + r2 * r2].reshape(r2, r2))]
        np.testing.assert_allclose(sorted(got, reverse=True), ref, rtol=0, atol=sv_tol * ref[0])
