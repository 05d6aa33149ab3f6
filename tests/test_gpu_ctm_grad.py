"""Reverse-mode recording (the native tape) and gradient CTMRG on the GPU.  Needs an MI355X.

The gradient of a function is fixed by the function: every check compares the recorded reverse
pass with central finite differences of the same forward computation along random directions V,
<grad f, V> = (f(x + eV) - f(x - eV)) / 2e + O(e^2), at the relative tolerance that e = 1e-5
leaves (1e-6).  The forward computations themselves are pinned against the oracle elsewhere (the
contraction, SVD and CTMRG tests).

* A chain of the recorded operations (permute, contract, operator*, adjoint, scaled, axpby,
  truncated tSVD, diag maps, tQR, div_trace) ending in a trace.
* CTMRG: the energy after 1 and after 3 tracked grow_all moves from a fixed environment (each move a
  checkpoint that the reverse pass recomputes), D = 3, chi = 12, with respect to all four site
  tensors.  The kept counts of every projector decomposition must be the same at x +- eV (the
  truncation is then locally smooth).
"""
import numpy as np
import pytest

import xped_amd.tensor as xt
from xped_amd import configs

from .helpers import make_ipeps

pytestmark = pytest.mark.gpu

EPS = 1e-5


def _dot(a, b):
    return float(np.dot(a.download(), b.download()))


def _rand_like(t, seed):
    r = np.random.default_rng(seed)
    return t.zeros_like().upload(r.standard_normal(t.size))


def _chain(A, B):
    # every recorded operation once; f = trace((Z / |Z|^2) Y^dagger)
    X = A.contract(B, [-1, -2, 1], [1, -3, -4], 2)  # (a, b | c, d)
    Xp = X.permute(-1, [0, 1, 3, 2])  # (a, b, d | c)
    Y = Xp.adjoint().adjoint().scale(0.7).axpby(Xp, 0.3, 1.0)
    U, S, V = Y.tSVD(3, 1e-12, PRESERVE_MULTIPLETS=False)[:3]
    Sq = S.diag_sqrt()
    W = U * Sq * S.diag_inv() * S * Sq * V  # the rank-3 part of Y
    Q, R = Y.tQR()
    Z = (Q * R).axpby(W, 1.0, 0.5)
    M = Z.div_trace(Z * Z.adjoint())
    return M * Y.adjoint()


def test_native_tape_gradient_of_an_operation_chain(rt):
    leg = [((-1,), 2), ((0,), 3), ((1,), 2)]
    b = xt.Qbasis(xt.U1)
    for q, d in leg:
        b.push_back(q, d)
    A = xt.Tensor(rt, [b, b], [b])
    B = xt.Tensor(rt, [b], [b, b])
    A.setRandom(3, 1)
    B.setRandom(4, 1)
    with xt.Tape(rt) as tape:
        f = _chain(A, B)
        tape.pause()
        fv = f.trace()
        tape.seed_trace(f, 1.0)
        tape.backward()
        gA, gB = tape.adjoint(A), tape.adjoint(B)
    for seed in (1, 2):
        VA, VB = _rand_like(A, 10 * seed), _rand_like(B, 10 * seed + 1)
        fp = _chain(A.axpby(VA, 1.0, EPS), B.axpby(VB, 1.0, EPS)).trace()
        fm = _chain(A.axpby(VA, 1.0, -EPS), B.axpby(VB, 1.0, -EPS)).trace()
        fd = (fp - fm) / (2 * EPS)
        ad = _dot(gA, VA) + _dot(gB, VB)
        assert abs(fd - ad) <= 1e-6 * max(abs(fd), abs(fv)), (fd, ad)


def _kept(c):
    return [{s["q"]: s["kept"] for s in r["sectors"]} for r in c.spectra()]


@pytest.mark.parametrize("steps", [1, 3])
def test_ctm_energy_gradient_matches_finite_differences(rt, steps):
    chi = 12
    _, A = make_ipeps(rt, configs.c4_small_virtual(), configs.c4_physical(), 5)
    base = xt.CTM(rt, A, chi).init()
    base.grow_all()  # the fixed environment the tracked moves start from
    g = xt.CTM(rt, A, chi).copy_env(base)
    E, dA = g.energy_grad(steps)
    # the forward value agrees with an untracked run from the same environment
    plain = xt.CTM(rt, A, chi).copy_env(base).set_options(record_spectra=True)
    for _ in range(steps):
        plain.grow_all()
    plain.computeRDM()
    assert abs(plain.energy() - E) <= 1e-12 * abs(E)
    kept0 = _kept(plain)
    for seed in (7, 8):
        V = [_rand_like(a, seed * 10 + i) for i, a in enumerate(A)]
        vals, kepts = [], []
        for sgn in (1.0, -1.0):
            Ap = [a.axpby(v, 1.0, sgn * EPS) for a, v in zip(A, V)]
            c = xt.CTM(rt, Ap, chi).copy_env(base).set_options(record_spectra=True)
            for _ in range(steps):
                c.grow_all()
            c.computeRDM()
            vals.append(c.energy())
            kepts.append(_kept(c))
        assert kepts[0] == kept0 and kepts[1] == kept0, "a kept count changed under the perturbation"
        fd = (vals[0] - vals[1]) / (2 * EPS)
        ad = sum(_dot(d, v) for d, v in zip(dA, V))
        assert abs(fd - ad) <= 1e-6 * max(abs(fd), abs(E)), (steps, fd, ad)
