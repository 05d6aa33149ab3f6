"""Reverse-mode AD on the HIP path (needs an MI355X).

Primitive parity against the oracle (oracle/ad.py, pinned on the CPU by tests/test_oracle_ad.py):
permute_adj is data movement with the forward coefficients, bit-exact for U(1) and fermionic
U(1)xU(1) (one source per element) and within 1e-13 absolute for SU(2); the product and SVD
adjoints are FP64 GEMM chains, within 1e-12 (product) and 1e-9 (SVD, where 1/(s_j^2 - s_i^2)
amplifies rounding) relative to the largest entry.

Tape tests mirror the reference's tests/ad.cpp (check_ad_vs_fd, tests/ad_tests.hpp:10-39): the
gradient of a scalar function of a random tensor (sorted random bases, as Qbasis::setRandom +
sort) against central finite differences, here directional (three random unit directions,
h = 1e-5, relative tolerance 1e-6).  The reference's tSVD cases read single entries of U or Vdag,
which depend on the sign gauge of the SVD; these tests use gauge-invariant functions of the
truncated factors instead (U S Vdag, U Vdag, the kept singular values).
"""
import itertools
import random

import numpy as np
import pytest

import xped_amd.tensor as xt
from oracle import ad as oad
from oracle import tensor as oten
from oracle.symmetry import Sym
from xped_amd._lib import XpedError
from xped_amd.ad import ADTensor, Tape

from .helpers import KINDS, both_basis, rand_entries, structure_of

pytestmark = pytest.mark.gpu


def sorted_both(rt, pcode, legs_entries, rank, seed):
    """(oracle tensor, product tensor) with sorted legs and the same uniform data."""
    pairs = [both_basis(pcode, e) for e in legs_entries]
    for o, p in pairs:
        o.sort()
        p.sort()
    ob = [p[0] for p in pairs]
    pb = [p[1] for p in pairs]
    o = oten.Tensor(Sym(*KINDS[pcode]), ob[:rank], ob[rank:]).set_random(seed)
    t = xt.Tensor(rt, pb[:rank], pb[rank:])
    t.upload(o.data)
    return o, t


@pytest.mark.parametrize("pcode", [xt.U1, xt.FU1xU1, xt.SU2], ids=["U1", "FU1xU1", "SU2"])
def test_permute_adj_vs_oracle(rt, pcode):
    r = random.Random(21)
    for rank, corank in [(2, 2), (2, 1), (1, 2)]:
        legs = [rand_entries(pcode, r, 3, 4) for _ in range(rank + corank)]
        ox, tx = sorted_both(rt, pcode, legs, rank, 5)
        for p in itertools.permutations(range(rank + corank)):
            for shift in range(-corank, rank + 1):
                ty = tx.permute(shift, p)
                if ty.size == 0:
                    continue
                oy = oten.permute(ox, shift, p)
                oy.data = np.random.default_rng(ty.size + shift).uniform(-1, 1, ty.size)
                ty.upload(oy.data)
                got = ty.permute_adj(shift, p)
                want = oad.permute_adj(oy, shift, p)
                assert structure_of(got) == structure_of(tx), (p, shift)
                if pcode == xt.SU2:
                    np.testing.assert_allclose(got.download(), want.data, rtol=0, atol=1e-13)
                else:
                    assert np.array_equal(got.download(), want.data), (p, shift)


def test_permute_adj_large_subblocks_exact(rt):
    # degeneracies >= 16: the adjoint plan goes through the LDS transpose kernel as well
    r = random.Random(22)
    legs = [rand_entries(xt.U1, r, 3, 40, mindim=16) for _ in range(4)]
    ox, tx = sorted_both(rt, xt.U1, legs, 2, 6)
    for p, shift in [((1, 0, 2, 3), 0), ((2, 3, 0, 1), 0), ((3, 1, 2, 0), 1), ((1, 2, 3, 0), -1), ((0, 3, 2, 1), 2)]:
        ty = tx.permute(shift, p)
        back = ty.permute_adj(shift, p)
        assert np.array_equal(back.download(), ox.data), (p, shift)


def test_mul_bwd_vs_oracle(rt):
    r = random.Random(23)
    e = [rand_entries(xt.U1, r, 4, 30) for _ in range(4)]
    oa, ta = sorted_both(rt, xt.U1, [e[0], e[1], e[2]], 2, 1)
    ob, tb = sorted_both(rt, xt.U1, [e[2], e[3]], 1, 2)
    tc = ta.mul(tb)
    oc = oten.mul(oa, ob)
    oc.data = np.random.default_rng(3).uniform(-1, 1, oc.size)
    tc.upload(oc.data)
    # accumulate onto non-zero adjoints
    da0 = np.random.default_rng(4).uniform(-1, 1, oa.size)
    db0 = np.random.default_rng(5).uniform(-1, 1, ob.size)
    tda, tdb = ta.zeros_like().upload(da0), tb.zeros_like().upload(db0)
    xt.mul_bwd(ta, tb, tc, tda, tdb)
    oda, odb = oa.copy(), ob.copy()
    oda.data, odb.data = da0.copy(), db0.copy()
    oad.mul_bwd(oa, ob, oc, oda, odb)
    for got, want in [(tda.download(), oda.data), (tdb.download(), odb.data)]:
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 * np.abs(want).max())


@pytest.mark.parametrize("pcode", [xt.U1, xt.SU2], ids=["U1", "SU2"])
def test_svd_bwd_vs_oracle(rt, pcode):
    r = random.Random(24)
    # blocks up to ~90 x 90 (large enough for the GEMM-based block Jacobi path) plus tall/wide ones
    legs = [rand_entries(pcode, r, 3, 12 if pcode == xt.U1 else 5) for _ in range(3)]
    ox, tx = sorted_both(rt, pcode, legs, 2, 7)
    U, S, V = tx.full_svd_impl()
    rng = np.random.default_rng(8)
    dU = U.zeros_like().upload(rng.uniform(-1, 1, U.size))
    dV = V.zeros_like().upload(rng.uniform(-1, 1, V.size))
    dS = S.zeros_like().upload(rng.uniform(-1, 1, S.size))
    dA = tx.zeros_like()
    xt.svd_bwd(U, S, V, dU, dS, dV, dA)

    # the oracle works per block on the device factors, read through the product's sector table
    def blocks(t):
        d = t.download()
        return {q: d[o:o + rr * cc].reshape(cc, rr).T for q, o, rr, cc in t.sectors()}

    bu, bs, bv, bdu, bds, bdv = (blocks(x) for x in (U, S, V, dU, dS, dV))
    got = blocks(dA)
    assert got, "no sectors"
    for q, o, rr, cc in tx.sectors():
        s = np.diag(bs[q]).copy()
        want = oad.svd_bwd_block(bu[q], s, bv[q], bdu[q], bds[q], bdv[q])
        np.testing.assert_allclose(got[q], want, rtol=0, atol=1e-9 * max(1.0, np.abs(want).max()))


def test_cutoff_bwd_scatters(rt):
    r = random.Random(25)
    legs = [rand_entries(xt.U1, r, 3, 8) for _ in range(3)]
    ox, tx = sorted_both(rt, xt.U1, legs, 2, 9)
    U, S, V = tx.full_svd_impl()
    Ut, St, Vt, tw, en = tx.cutoff_matrices(U, S, V, 10, 1e-10, 0.0, 0.0, False)
    rng = np.random.default_rng(10)
    dUt, dSt, dVt = (x.zeros_like().upload(rng.uniform(-1, 1, x.size)) for x in (Ut, St, Vt))
    dU, dS, dV = U.zeros_like(), S.zeros_like(), V.zeros_like()
    xt.cutoff_bwd(dUt, dSt, dVt, dU, dS, dV)

    def blocks(t):
        d = t.download()
        return {q: d[o:o + rr * cc].reshape(cc, rr).T for q, o, rr, cc in t.sectors()}

    for small, full, mode in [(dUt, dU, "cols"), (dSt, dS, "corner"), (dVt, dV, "rows")]:
        bs_, bf = blocks(small), blocks(full)
        for q, f in bf.items():
            want = np.zeros_like(f)
            if q in bs_:
                a = bs_[q]
                want[: a.shape[0], : a.shape[1]] = a
            np.testing.assert_array_equal(f, want)
    # the truncated factors are the leading parts of the full ones
    bu, but = blocks(U), blocks(Ut)
    for q, a in but.items():
        np.testing.assert_array_equal(a, bu[q][:, : a.shape[1]])


# ------------------------------------------------------------------ tape vs finite differences
def _random_t(rt, pcode, seed, nsec=3, maxdim=3):
    r = random.Random(seed)
    e1, e2 = rand_entries(pcode, r, nsec, maxdim), rand_entries(pcode, r, nsec, maxdim)
    _, t = sorted_both(rt, pcode, [e1, e2, e1, e2], 2, seed)
    return t


def check_ad_vs_fd(rt, f, t, ndir=3, h=1e-5, rtol=1e-6):
    tape = Tape()
    x = ADTensor(tape, t)
    res = f(x)
    tape.grad(res)
    g = x.adj.download()
    base = t.download()
    rng = np.random.default_rng(99)
    for _ in range(ndir):
        v = rng.standard_normal(base.size)
        v /= np.linalg.norm(v)
        vals = []
        for sgn in (1.0, -1.0):
            tt = t.zeros_like().upload(base + sgn * h * v)
            vals.append(f(ADTensor(Tape(), tt)).val)
        fd = (vals[0] - vals[1]) / (2 * h)
        ad = float(g @ v)
        assert abs(ad - fd) <= rtol * max(1.0, abs(fd), np.abs(g).max()), (ad, fd)
    return g


FUNCS = {
    "coeff": lambda t: t.coeff(0, 0, 0),
    "squaredNorm": lambda t: t.mul(t.adjoint()).trace(),
    "trace": lambda t: t.trace(),
    "twist": lambda t: t.twist(0).coeff(0, 0, 0),
    "norm": lambda t: t.norm(),
    "maxNorm": lambda t: t.maxNorm(),
    "mul_self": lambda t: t.mul(t).coeff(0, 0, 0),
    "permute": lambda t: t.permute(-1, (3, 1, 2, 0)).coeff(0, 0, 0),
    "contract": lambda t: t.contract(t, (-1, 1, 2, -2), (2, 1, -3, -4), 2).coeff(0, 0, 0),
    "contract_u1": lambda t: t.contract(t, (-1, -2, 1, 2), (1, 2, -3, -4), 2).coeff(0, 0, 0),
    "diag_chain": lambda t: t.mul(t.adjoint()).tSVD(50, 1e-10, False)[1].diag_sqrt().diag_inv().trace(),
}


@pytest.mark.parametrize("pcode", [xt.U1, xt.SU2], ids=["U1", "SU2"])
@pytest.mark.parametrize("name", list(FUNCS))
def test_tape_vs_finite_differences(rt, pcode, name):
    # the reference contracts leg 1 with leg 1 for SU(2) (self-dual legs) and uses the
    # domain-with-codomain pattern for U(1) (tests/ad.cpp:398-412)
    if (name == "contract" and pcode == xt.U1) or (name == "contract_u1" and pcode == xt.SU2):
        pytest.skip("contraction pattern of the other symmetry")
    t = _random_t(rt, pcode, 31)
    check_ad_vs_fd(rt, FUNCS[name], t)


def _svd_reconstruct(t):
    U, S, V, _, _ = t.tSVD(6, 1e-10, False)
    return U.mul(S).mul(V).coeff(0, 0, 0)


def _svd_polar(t):
    U, S, V, _, _ = t.tSVD(6, 1e-10, False)
    return U.mul(V).mul(t.adjoint()).trace()


def _svd_values(t):
    U, S, V, _, _ = t.tSVD(6, 1e-10, False)
    return S.trace()


@pytest.mark.parametrize("pcode", [xt.U1, xt.SU2], ids=["U1", "SU2"])
@pytest.mark.parametrize("f", [_svd_reconstruct, _svd_polar, _svd_values], ids=["USV", "UV", "S"])
def test_tsvd_tape_vs_finite_differences(rt, pcode, f):
    t = _random_t(rt, pcode, 32, nsec=3, maxdim=3)
    check_ad_vs_fd(rt, f, t, rtol=2e-6)


# ------------------------------------------------------------------ tQR: LQ mode, adjoint
def _qr_t(rt, pcode, shape, seed):
    """Rank (1, 1) tensor whose sectors are all tall (rows >= cols) or all wide, so that the
    middle basis of tQR has min(rows, cols) in every sector."""
    qs = [(-1,), (0,), (1,)] if pcode != xt.SU2 else [(1,), (2,), (3,)]
    dims = {"tall": [(5, 3), (4, 4), (6, 2)], "wide": [(2, 5), (3, 3), (1, 4)], "square": [(3, 3), (4, 4), (2, 2)]}[shape]
    _, t = sorted_both(rt, pcode, [[(q, d[0]) for q, d in zip(qs, dims)], [(q, d[1]) for q, d in zip(qs, dims)]], 1, seed)
    return t


def _blocks(t):
    d = t.download()
    return {q: d[o:o + r * c].reshape(c, r).T for q, o, r, c in t.sectors()}


@pytest.mark.parametrize("pcode", [xt.U1, xt.SU2], ids=["U1", "SU2"])
def test_tqr_lq_mode_returns_the_transposed_factors(rt, pcode):
    t = _qr_t(rt, pcode, "square", 41)
    Q, R = t.tQR()
    Ql, Rl = t.tQR(RETURN_LQ=True)
    assert structure_of(Ql) == structure_of(Q) and structure_of(Rl) == structure_of(R)
    bq, br, bql, brl, bt = _blocks(Q), _blocks(R), _blocks(Ql), _blocks(Rl), _blocks(t)
    for q in bq:
        np.testing.assert_array_equal(bql[q], bq[q].T)
        np.testing.assert_array_equal(brl[q], br[q].T)
        # Q^T and R^T are the LQ factors of A^T: A^T = R^T Q^T with R^T lower triangular
        np.testing.assert_allclose(brl[q] @ bql[q], bt[q].T, rtol=0, atol=1e-12 * np.abs(bt[q]).max())
        assert np.abs(np.triu(brl[q], 1)).max(initial=0.0) == 0.0
    with pytest.raises(XpedError):
        _qr_t(rt, pcode, "tall", 42).tQR(RETURN_LQ=True)


def _qr_f(which):
    def f(x):
        Q, R = x.tQR()
        rng = np.random.default_rng(5)
        wq = ADTensor(x.tape, Q.val.zeros_like().upload(rng.standard_normal(Q.val.size)))
        wr = ADTensor(x.tape, R.val.zeros_like().upload(rng.standard_normal(R.val.size)))
        fq = wq.adjoint().mul(Q).trace()
        fr = wr.adjoint().mul(R).trace()
        return {"Q": fq, "R": fr, "QR": fq + fr}[which]
    return f


@pytest.mark.parametrize("pcode", [xt.U1, xt.SU2], ids=["U1", "SU2"])
@pytest.mark.parametrize("shape", ["tall", "wide", "square"])
@pytest.mark.parametrize("which", ["Q", "R", "QR"])
def test_tqr_tape_vs_finite_differences(rt, pcode, shape, which):
    # the QR adjoint (xt_qr_bwd) through the tape: <Wq, Q(A)> + <Wr, R(A)> against central differences
    t = _qr_t(rt, pcode, shape, 43)
    check_ad_vs_fd(rt, _qr_f(which), t, rtol=2e-6)


@pytest.mark.parametrize("pcode", [xt.U1, xt.SU2], ids=["U1", "SU2"])
def test_tsvd_svspec(rt, pcode):
    # tSVD's SVspec (Tensor.cpp:620-624): the kept singular values per sector, inserted into the map
    r = random.Random(17)
    legs = [rand_entries(pcode, r, 4, 5) for _ in range(3)]
    o, t = sorted_both(rt, pcode, legs, 2, 19)
    U, S, V, tw, en, spec = t.tSVD_spec(12, 1e-10, PRESERVE_MULTIPLETS=False)
    bs = _blocks(S)
    assert set(spec) == set(bs)
    for q, b in bs.items():
        np.testing.assert_array_equal(spec[q], np.diag(b))
    oU, oS, oV, otw, oent, kept = oten.tsvd(o, 12, 1e-10, False)
    for q, v in kept.items():
        np.testing.assert_allclose(spec[tuple(q)], v, rtol=0, atol=1e-12 * oS.data.max())
    # std::map::insert keeps an existing entry
    q0 = next(iter(spec))
    again = t.tSVD_spec(12, 1e-10, SVspec={q0: np.array([-1.0])}, PRESERVE_MULTIPLETS=False)[5]
    assert list(again[q0]) == [-1.0]
    # RETURN_SPEC = false leaves the map alone
    assert t.tSVD_spec(12, 1e-10, SVspec={}, PRESERVE_MULTIPLETS=False, RETURN_SPEC=False)[5] == {}
