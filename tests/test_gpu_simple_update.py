"""Simple update (TimePropagator, nearest-neighbour terms) on the GPU against the oracle and against
dense invariants.  Needs an MI355X.

* mexp: the block matrix exponential (k_block_mexp) against scipy.linalg.expm on random blocks of
  orders 1..64 with norms from 0.05 to 40 (so that 0 to 8 squarings run), at 1e-12 relative.
* One bond update on the GPU, horizontal and vertical, checked directly (no oracle): the new
  weights are the singular values of the gate applied to the dense two-site block, and the new
  two-site block is its best rank-D approximation over the weight's max norm (1e-12).
* Whole steps against the oracle from the same seeded sites: structures bit-exact, weights and
  spectra at 1e-12, and the site tensors equal up to the sign freedom of each bond index (the
  singular vectors that define a bond index are fixed only up to sign): the signs are solved over
  GF(2) from the entries and must be consistent, after which the entries agree at 1e-12 relative.
"""
import numpy as np
import pytest
import scipy.linalg

import xped_amd.tensor as xt
from oracle import simple_update as osu
from oracle import tensor as ot
from xped_amd import configs

from .helpers import both_basis, oracle_structure, structure_of
from .test_oracle_su_pins import _apply_gate, _best_rank, _dense_heisenberg, _sites, _theta_h, _theta_v, _w

pytestmark = pytest.mark.gpu

PHYS = configs.c4_physical()


def _gpu_sites(rt, oA):
    out = []
    for o, legs in zip(oA, configs.ipeps_site_legs(configs.c4_small_virtual(), PHYS)):
        pb = [both_basis(xt.U1, e)[1] for e in legs]
        t = xt.Tensor(rt, pb[:2], pb[2:]).upload(o.data)
        assert structure_of(t) == oracle_structure(o)
        out.append(t)
    return out


def _like(o, t):
    """The oracle tensor o's structure with the GPU tensor t's data (structures must agree)."""
    assert structure_of(t) == oracle_structure(o)
    c = o.copy()
    c.data = t.download()
    return c


def _propagators(rt, Jxy, Jz, dt):
    P = both_basis(xt.U1, PHYS)
    Hg = xt.heisenberg_bond(rt, P[1], Jxy, Jz)
    Ho = osu.heisenberg_bond(_sites(configs.c4_small_virtual(), 0)[0].sym, P[0], Jxy, Jz)
    np.testing.assert_array_equal(Hg.download(), Ho.data)
    return xt.TimePropagator(rt, [Hg] * 4, [Hg] * 4, dt), osu.TimePropagator([Ho] * 4, [Ho] * 4, dt)


@pytest.mark.parametrize("m", [1, 2, 5, 16, 33, 64])
def test_mexp_matches_scipy(rt, m):
    r = np.random.default_rng(m)
    b = xt.Qbasis(xt.U1)
    b.push_back((0,), m)
    b.push_back((1,), max(1, m // 2))
    t = xt.Tensor(rt, [b], [b])
    for scale in (0.05, 1.0, 40.0):
        data = r.standard_normal(t.size) * scale / np.sqrt(m)
        t.upload(data)
        got = t.mexp(-0.7).download()
        for q, off, rows, cols in t.sectors():
            blk = data[off:off + rows * cols].reshape(cols, rows).T
            ref = scipy.linalg.expm(-0.7 * blk)
            g = got[off:off + rows * cols].reshape(cols, rows).T
            assert np.abs(g - ref).max() <= 1e-12 * np.abs(ref).max(), (m, scale, q)


@pytest.mark.parametrize("vertical", [False, True], ids=["h", "v"])
def test_bond_update_is_the_best_rank_d_gate_application(rt, vertical):
    D, dt, Jxy, Jz = 3, 0.3, 0.6, 1.4
    oA = _sites(configs.c4_small_virtual(), 21)
    tp, _ = _propagators(rt, Jxy, Jz, dt)
    U = scipy.linalg.expm(-dt * _dense_heisenberg(Jxy, Jz))
    psi = xt.WeightedPEPS(rt, _gpu_sites(rt, oA), D)
    ref = osu.WeightedPEPS(oA, D)  # identity weights; only the structures of its tensors are used

    def snapshot():
        ns = type("S", (), {})()
        ns.Gs = [_like(ref.Gs[i], psi.G(i)) for i in range(4)]
        ns.whs = [_like(ref.whs[i], psi.wh(i)) for i in range(4)]
        ns.wvs = [_like(ref.wvs[i], psi.wv(i)) for i in range(4)]
        return ns

    x, y = 0, 1
    th = (_theta_v if vertical else _theta_h)(osu, snapshot(), x, y)
    best, sv = _best_rank(_apply_gate(th, U), D)
    (tp.t_step_v if vertical else tp.t_step_h)(psi, x, y)
    # the updated sites and bond have new structures: take them from the oracle run on the same input
    oprop = osu.TimePropagator([osu.heisenberg_bond(oA[0].sym, both_basis(xt.U1, PHYS)[0], Jxy, Jz)] * 4,
                               [osu.heisenberg_bond(oA[0].sym, both_basis(xt.U1, PHYS)[0], Jxy, Jz)] * 4, dt)
    (oprop.t_step_v if vertical else oprop.t_step_h)(ref, x, y)
    snap = snapshot()
    w = _w(_like((oprop.spectrum_v if vertical else oprop.spectrum_h)[osu.idx(x, y)],
                 (tp.spectrum_v if vertical else tp.spectrum_h)(osu.idx(x, y))))
    kept = np.sort(w)[::-1]
    assert len(kept) == D
    np.testing.assert_allclose(kept, sv[:D], rtol=0, atol=1e-12 * sv[0])
    new = (_theta_v if vertical else _theta_h)(osu, snap, x, y).reshape(best.shape)
    assert np.abs(new - best / kept[0]).max() <= 1e-12 * np.abs(best / kept[0]).max()


def _legs_of(x, y):
    i = osu.idx
    return [("h", i(x - 1, y)), ("v", i(x, y)), ("h", i(x, y)), ("v", i(x, y + 1))]


def _solve_signs(pairs):
    """pairs: [(site (x, y), gpu dense, oracle dense)].  Returns {(bond, k): sign} such that
    gpu = prod of the four leg signs * oracle on every significant entry, or fails."""
    cols = {}
    basis = {}

    def col(key):
        if key not in cols:
            cols[key] = len(cols)
        return cols[key]

    for (x, y), g, o in pairs:
        legs = _legs_of(x, y)
        big = np.argwhere(np.abs(o) > 1e-3 * np.abs(o).max())
        for e in big:
            e = tuple(e)
            ratio = g[e] / o[e]
            assert abs(abs(ratio) - 1.0) < 1e-6, (x, y, e, ratio)
            mask = 0
            for leg, (kind, b) in enumerate(legs):
                mask ^= 1 << col((kind, b, int(e[leg])))
            rhs = int(ratio < 0)
            while mask:
                hb = mask.bit_length() - 1
                if hb in basis:
                    mask ^= basis[hb][0]
                    rhs ^= basis[hb][1]
                else:
                    basis[hb] = (mask, rhs)
                    break
            assert mask or not rhs, "inconsistent bond signs: the tensors differ by more than a gauge"
    val = {}
    for hb in sorted(basis):
        mask, rhs = basis[hb]
        v = rhs
        rest = mask & ~(1 << hb)
        while rest:
            lb = rest.bit_length() - 1
            v ^= val.get(lb, 0)
            rest &= ~(1 << lb)
        val[hb] = v
    return {key: -1.0 if val.get(c, 0) else 1.0 for key, c in cols.items()}


@pytest.mark.parametrize("steps,D", [(1, 3), (3, 3), (2, 2)])
def test_t_steps_match_the_oracle_up_to_bond_signs(rt, steps, D):
    dt = 0.1
    oA = _sites(configs.c4_small_virtual(), 40 + steps)
    tp, otp = _propagators(rt, 1.0, 1.0, dt)
    psi = xt.WeightedPEPS(rt, _gpu_sites(rt, oA), D)
    ref = osu.WeightedPEPS(oA, D)
    for _ in range(steps):
        tp.t_step(psi)
        otp.t_step(ref)
    for i in range(4):
        for got, want in ((psi.wh(i), ref.whs[i]), (psi.wv(i), ref.wvs[i]), (tp.spectrum_h(i), otp.spectrum_h[i]), (tp.spectrum_v(i), otp.spectrum_v[i])):
            assert structure_of(got) == oracle_structure(want)
            np.testing.assert_allclose(got.download(), want.data, rtol=0, atol=1e-12 * np.abs(want.data).max())
    pairs = []
    for x in range(2):
        for y in range(2):
            i = osu.idx(x, y)
            g = psi.G(i)
            assert structure_of(g) == oracle_structure(ref.Gs[i])
            pairs.append(((x, y), ot.plain(_like(ref.Gs[i], g)), ot.plain(ref.Gs[i])))
    signs = _solve_signs(pairs)

    def gauge(x, y, shape):
        s = np.ones(shape)
        for leg, (kind, b) in enumerate(_legs_of(x, y)):
            v = np.array([signs.get((kind, b, k), 1.0) for k in range(shape[leg])])
            sh = [1] * 5
            sh[leg] = -1
            s = s * v.reshape(sh)
        return s

    for (x, y), g, o in pairs:
        assert np.abs(g - gauge(x, y, o.shape) * o).max() <= 1e-12 * np.abs(o).max(), (x, y)
    # the weighted sites (updateAtensors) carry the same gauge
    oAw = ref.A()
    for x in range(2):
        for y in range(2):
            i = osu.idx(x, y)
            o = ot.plain(oAw[i])
            g = ot.plain(_like(oAw[i], psi.A(i)))
            assert np.abs(g - gauge(x, y, o.shape) * o).max() <= 1e-12 * np.abs(o).max(), (x, y)
