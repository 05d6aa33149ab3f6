"""Grouped FP64 block SVD (xh_block_svd) against numpy's LAPACK SVD, through the C ABI (needs an
MI355X).  One call mixes the paths: small blocks on the single-workgroup Jacobi kernel, blocks
with more than 64 columns (or too large for LDS) on block Jacobi, tall and wide shapes, a
rank-deficient block and a graded spectrum spanning twelve decades.

Tolerances: singular values within 1e-12 * sigma_max * sqrt(n) of LAPACK's (FP64 backward
stability); U diag(S) Vt reproduces the block to 1e-12 * ||A||_2 * sqrt(n); the columns of U with
nonzero sigma and the matching rows of Vt are orthonormal to 1e-12 * sqrt(n).
"""
import ctypes as C
import math

import numpy as np
import pytest

from xped_amd import _lib
from xped_amd._lib import lib

pytestmark = pytest.mark.gpu


def _blocks():
    r = np.random.default_rng(31)
    out = []
    for m, n in [(3, 5), (40, 40), (64, 70), (65, 64), (130, 97), (97, 130), (300, 256), (5000, 40), (40, 3000), (513, 511)]:
        out.append(r.uniform(-1, 1, (m, n)))
    # rank 5 block of 200 x 150
    X = r.standard_normal((200, 5)) @ r.standard_normal((5, 150))
    out.append(X)
    # graded spectrum 1 .. 1e-12
    k = 160
    Q1, _ = np.linalg.qr(r.standard_normal((180, k)))
    Q2, _ = np.linalg.qr(r.standard_normal((k, k)))
    out.append((Q1 * np.logspace(0, -12, k)) @ Q2.T)
    return out


def test_grouped_block_svd_vs_lapack(rt):
    L = lib()
    ctx = rt.ctx()
    blocks = _blocks()
    items, ao, uo, so, vo = [], 0, 0, 0, 0
    for A in blocks:
        m, n = A.shape
        k = min(m, n)
        items.append(_lib.xh_factor_item(ao, uo, so, vo, m, n))
        ao += m * n
        uo += m * k
        so += k
        vo += k * n
    Aall = np.concatenate([A.T.reshape(-1) for A in blocks])
    ptrs = [C.c_void_p() for _ in range(4)]
    for p, sz in zip(ptrs, (ao, uo, so, vo)):
        assert L.xh_alloc(ctx, 8 * sz, C.byref(p)) == 0
    dA, dU, dS, dV = ptrs
    L.xh_upload(ctx, dA, Aall.ctypes.data, 8 * ao)
    arr = (_lib.xh_factor_item * len(items))(*items)
    assert L.xh_block_svd(ctx, arr, len(items), dA, dU, dS, dV) == 0, L.xh_last_error(ctx)
    U, S, V = np.empty(uo), np.empty(so), np.empty(vo)
    L.xh_download(ctx, U.ctypes.data, dU, 8 * uo)
    L.xh_download(ctx, S.ctypes.data, dS, 8 * so)
    L.xh_download(ctx, V.ctypes.data, dV, 8 * vo)
    for p in ptrs:
        L.xh_free(ctx, p)
    for it, A in zip(items, blocks):
        m, n = A.shape
        k = min(m, n)
        u = U[it.u_off:it.u_off + m * k].reshape(k, m).T
        s = S[it.s_off:it.s_off + k]
        vt = V[it.vt_off:it.vt_off + k * n].reshape(n, k).T
        ref = np.linalg.svd(A, compute_uv=False)
        scale = ref[0] * math.sqrt(k)
        assert np.all(np.diff(s) <= 0.0), (m, n)
        np.testing.assert_allclose(s, ref, rtol=0, atol=1e-12 * scale, err_msg=str((m, n)))
        np.testing.assert_allclose((u * s) @ vt, A, rtol=0, atol=1e-12 * scale, err_msg=str((m, n)))
        live = s > 1e-13 * ref[0]
        np.testing.assert_allclose(u[:, live].T @ u[:, live], np.eye(int(live.sum())), rtol=0, atol=1e-12 * math.sqrt(k))
        np.testing.assert_allclose(vt[live] @ vt[live].T, np.eye(int(live.sum())), rtol=0, atol=1e-12 * math.sqrt(k))
