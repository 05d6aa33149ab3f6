"""Pins of the CPU oracle against golden vectors, libstdc++ and closed-form identities.

No GPU and no product code here: these tests establish that the oracle itself is right before
it is used to judge the HIP path.
"""
import itertools
import json
import math
import random
import subprocess
from pathlib import Path

import numpy as np
import pytest

from oracle import rng, structure, symmetry, tensor, wigner
from oracle.structure import FTree, Qbasis, decompose, inverse
from oracle.symmetry import FU1, SU2, U1, Sym

GOLDEN = Path(__file__).parent / "golden"


# ----------------------------------------------------------------------------- random numbers

def test_mt19937_known_answer():
    # C++ standard [rand.predef]: the 10000th output of a default-constructed mt19937 is 4123659995
    e = rng.MT19937(5489)
    r = e.raw(10000)
    assert int(r[-1]) == 4123659995
    assert int(r[0]) == 3499211612


@pytest.fixture(scope="module")
def stdrng(tmp_path_factory):
    d = tmp_path_factory.mktemp("stdrng")
    exe = d / "stdrng"
    subprocess.run(["g++", "-O2", "-std=c++17", str(GOLDEN / "stdrng.cpp"), "-o", str(exe)], check=True)
    return exe


@pytest.mark.parametrize("seed", [0, 1, 42, 5489, 123456789])
def test_uniform_and_normal_match_libstdcxx(stdrng, seed):
    n = 257
    out = subprocess.run([str(stdrng), str(seed), str(n)], check=True, capture_output=True, text=True).stdout.split()
    vals = [float.fromhex(x) for x in out]
    u_ref, g_ref = vals[:n], vals[n:]
    u = rng.uniform(rng.MT19937(seed), n)
    assert [x for x in u] == u_ref  # bit-exact
    g = rng.Normal().sample(rng.MT19937(seed), n)
    assert [x for x in g] == g_ref  # bit-exact


def test_std_sort_restatement_matches_libstdcxx(tmp_path):
    from oracle.stdsort import std_sort

    exe = tmp_path / "stdsort"
    subprocess.run(["g++", "-O2", "-std=c++17", str(GOLDEN / "stdsort.cpp"), "-o", str(exe)], check=True)
    r = random.Random(99)
    cases = []
    for n in list(range(0, 40)) + [63, 64, 65, 100, 257, 1000, 3000]:
        for levels in (2, 5, 1000):
            cases.append([r.randint(0, levels) * 0.5 for _ in range(n)])
    cases.append([1.0] * 500)
    cases.append(list(range(300)))
    cases.append(list(range(300, 0, -1)))
    inp = "".join(f"{len(c)} " + " ".join(repr(float(x)) for x in c) + "\n" for c in cases)
    out = subprocess.run([str(exe)], input=inp, capture_output=True, text=True, check=True).stdout.splitlines()
    for c, line in zip(cases, out):
        v = [(i, x) for i, x in enumerate(c)]
        std_sort(v, lambda a, b: a[1] > b[1])
        assert [i for i, _ in v] == [int(x) for x in line.split()]


# ----------------------------------------------------------------------------- Wigner symbols

@pytest.fixture(scope="module")
def wfix():
    return json.loads((GOLDEN / "wigner_fixtures.json").read_text())


def _close(a, b, rel=4e-15, ab=1e-15):
    return abs(a - b) <= max(ab, rel * abs(b))


def test_oracle_6j_matches_sympy(wfix):
    bad = [(a, v, wigner.wigner_6j(*a)) for a, v in wfix["6j"] if not _close(wigner.wigner_6j(*a), v)]
    assert not bad, bad[:5]


def test_oracle_3j_matches_sympy(wfix):
    bad = [(a, v) for a, v in wfix["3j"] if not _close(wigner.wigner_3j(*a), v)]
    assert not bad, bad[:5]


def test_oracle_9j_matches_sympy(wfix):
    bad = [(a, v) for a, v in wfix["9j"] if not _close(wigner.wigner_9j(*a), v, rel=1e-14)]
    assert not bad, bad[:5]


def test_6j_orthogonality():
    # sum_X (2X+1)(2Y+1) {a b X; c d Y}{a b X; c d Y'} = delta(Y, Y')   (doubled arguments)
    for ta, tb, tc, td in [(1, 1, 1, 1), (2, 1, 3, 2), (4, 2, 3, 3), (2, 2, 2, 2), (3, 5, 4, 2)]:
        for ty, ty2 in itertools.product(range(0, 10), repeat=2):
            s = sum((tx + 1) * (ty + 1) * wigner.wigner_6j(ta, tb, tx, tc, td, ty) * wigner.wigner_6j(ta, tb, tx, tc, td, ty2) for tx in range(0, 16))
            tri = (symmetry.Sym(SU2).triangle((ta + 1,), (td + 1,), (ty + 1,)) and symmetry.Sym(SU2).triangle((tc + 1,), (tb + 1,), (ty + 1,))
                   and (ta + td + ty) % 2 == 0 and (tc + tb + ty) % 2 == 0)
            # X must be admissible with both (a, b) and (c, d) for the sum to run over a full set
            if tri and not any(wigner.wigner_6j(ta, tb, tx, tc, td, ty) != 0 for tx in range(0, 16)):
                continue
            expect = 1.0 if (ty == ty2 and tri) else 0.0
            assert abs(s - expect) < 1e-13


def test_6j_special_value():
    # {a b c; d e 0} = delta(a,e) delta(b,d) (-1)^{a+b+c} / sqrt((2a+1)(2b+1))
    for ta, tb, tc in itertools.product(range(0, 7), repeat=3):
        if not Sym(SU2).triangle((ta + 1,), (tb + 1,), (tc + 1,)) or (ta + tb + tc) % 2:
            continue
        v = wigner.wigner_6j(ta, tb, tc, tb, ta, 0)
        ph = -1.0 if ((ta + tb + tc) // 2) % 2 else 1.0
        assert abs(v - ph / math.sqrt((ta + 1) * (tb + 1))) < 1e-15


def test_3j_orthogonality():
    for tj1, tj2 in [(1, 1), (2, 1), (2, 2), (3, 2), (4, 3)]:
        for tj3, tj3b in itertools.product(range(abs(tj1 - tj2), tj1 + tj2 + 1, 2), repeat=2):
            for tm3 in range(-min(tj3, tj3b), min(tj3, tj3b) + 1, 2):
                s = 0.0
                for tm1 in range(-tj1, tj1 + 1, 2):
                    tm2 = -tm1 - tm3
                    if abs(tm2) > tj2:
                        continue
                    s += (tj3 + 1) * wigner.wigner_3j(tj1, tj2, tj3, tm1, tm2, tm3) * wigner.wigner_3j(tj1, tj2, tj3b, tm1, tm2, tm3)
                assert abs(s - (1.0 if tj3 == tj3b else 0.0)) < 1e-14


# ----------------------------------------------------------------------------- coefficients

def test_u1_and_fermionic_coefficients_closed_form():
    s, f = Sym(U1), Sym(FU1)
    for a, b in itertools.product(range(-8, 9), repeat=2):
        c = a + b
        assert s.coeff_swap((a,), (b,), (c,)) == 1.0
        assert s.coeff_swap((a,), (b,), (c + 1,)) == 0.0
        assert f.coeff_swap((a,), (b,), (c,)) == (-1.0 if (a % 2 and b % 2) else 1.0)
        assert s.coeff_turn((a,), (b,), (c,)) == 1.0 and s.coeff_FS((a,)) == 1.0
        assert f.coeff_twist((a,)) == (-1.0 if a % 2 else 1.0) and s.coeff_twist((a,)) == 1.0
        for d in range(-3, 4):
            # recouple = 1 iff every triangle of the 6j holds
            q1, q2, q3, Q = (a,), (b,), (d,), (a + b + d,)
            assert s.coeff_recouple(q1, q2, q3, Q, (a + b,), (b + d,)) == 1.0
            assert s.coeff_recouple(q1, q2, q3, Q, (a + b + 1,), (b + d,)) == 0.0


def test_su2_swap_and_turn_values():
    s = Sym(SU2)
    # two spin-1/2 (q=2): singlet (q=1) is antisymmetric, triplet (q=3) symmetric
    assert s.coeff_swap((2,), (2,), (1,)) == -1.0
    assert s.coeff_swap((2,), (2,), (3,)) == 1.0
    assert s.coeff_FS((2,)) == -1.0 and s.coeff_FS((3,)) == 1.0
    # turn of a leg: coeff_swap * sqrt(qf/ql)
    assert abs(s.coeff_turn((3,), (2,), (2,)) - s.coeff_swap((3,), (2,), (2,)) * math.sqrt(2 / 3)) < 1e-16


# ----------------------------------------------------------------------------- bases and trees

def rand_basis(sym, r, nsec, maxdim):
    b = Qbasis(sym, 1)
    for _ in range(nsec):
        if sym.kinds == (SU2,):
            q = (r.randint(1, 4),)
        elif sym.nq == 2:
            q = (r.randint(-3, 3), r.randint(-2, 2))
        else:
            q = (r.randint(-4, 4),)
        b.push_back(q, r.randint(1, maxdim))
    return b


SYMS = [Sym(U1), Sym(SU2), Sym(FU1, U1), Sym(FU1)]


@pytest.mark.parametrize("sym", SYMS, ids=lambda s: str(s.kinds))
def test_oracle_fulldim_of_product(sym):
    # tests/basis.cpp:44-91
    r = random.Random(7)
    for _ in range(10):
        B, C = rand_basis(sym, r, 4, 3), rand_basis(sym, r, 4, 3)
        assert B.combine(C).full_dim() == B.full_dim() * C.full_dim()


def test_oracle_push_back_merges():
    # tests/basis.cpp:93-142
    s = Sym(U1)
    a, b = Qbasis(s), Qbasis(s)
    a.push_back((1,), 1)
    a.push_back((1,), 3)
    b.push_back((1,), 4)
    assert a.data[0][:3] == b.data[0][:3] and a.data[0][3][0].key() == b.data[0][3][0].key()


def brute_force_sectors(sym, legs):
    """Independent enumerator: coupled sector dims of a product of bases (abelian only)."""
    out = {}
    for combo in itertools.product(*[[(q, d) for q, _, d, _ in b.data] for b in legs]):
        q = tuple(sum(x[0][c] for x in combo) for c in range(sym.nq))
        d = 1
        for x in combo:
            d *= x[1]
        out[q] = out.get(q, 0) + d
    return sorted(out.items())


@pytest.mark.parametrize("sym", [Sym(U1), Sym(FU1, U1)], ids=lambda s: str(s.kinds))
def test_oracle_combine_matches_brute_force(sym):
    r = random.Random(11)
    for _ in range(5):
        legs = [rand_basis(sym, r, 3, 3) for _ in range(3)]
        c = structure.build_fusion_tree(sym, legs)
        got = [(q, d) for q, _, d, _ in c.data]
        assert got == brute_force_sectors(sym, legs)
        # offsets are prefix sums and trees of each sector are sorted and sum to the sector dim
        acc = 0
        for q, num, d, trees in c.data:
            assert num == acc
            acc += d
            assert sum(t.dim for t in trees) == d
            assert all(trees[i].less(trees[i + 1]) for i in range(len(trees) - 1))


def _rank4(sym, seed):
    r = random.Random(seed)
    legs = [rand_basis(sym, r, 3, 2) for _ in range(4)]
    return structure.build_fusion_tree(sym, legs)


@pytest.mark.parametrize("sym", SYMS, ids=lambda s: str(s.kinds))
def test_oracle_swap_is_involution(sym):
    # tests/tree_tests.hpp:1-37
    b = _rank4(sym, 3)
    for pos in range(3):
        for q, _, _, trees in b.data:
            for t in trees:
                acc = {}
                for t2, c in structure.tree_swap(sym, t, pos).items():
                    for t3, c2 in structure.tree_swap(sym, t2, pos).items():
                        acc[t3] = acc.get(t3, 0.0) + c * c2
                for t3, c in acc.items():
                    assert abs(c - (1.0 if t3 == t else 0.0)) < 1e-12


@pytest.mark.parametrize("sym", SYMS, ids=lambda s: str(s.kinds))
def test_oracle_tree_permute_round_trip(sym):
    # tests/tree_tests.hpp:39-76, all 24 permutations of 4 legs
    b = _rank4(sym, 5)
    for p in itertools.permutations(range(4)):
        for q, _, _, trees in b.data:
            for t in trees:
                acc = {}
                for t2, c in structure.tree_permute(sym, t, p).items():
                    for t3, c2 in structure.tree_permute(sym, t2, inverse(p)).items():
                        acc[t3] = acc.get(t3, 0.0) + c * c2
                for t3, c in acc.items():
                    assert abs(c - (1.0 if t3 == t else 0.0)) < 1e-12


@pytest.mark.parametrize("sym", SYMS, ids=lambda s: str(s.kinds))
def test_oracle_turn_round_trip(sym):
    # tests/tree_tests.hpp:78-110
    r = random.Random(9)
    dom = structure.build_fusion_tree(sym, [rand_basis(sym, r, 3, 2) for _ in range(2)])
    cod = structure.build_fusion_tree(sym, [rand_basis(sym, r, 3, 2) for _ in range(2)])
    for q, _, _, t1s in dom.data:
        if cod.find(q) < 0:
            continue
        for t1 in t1s:
            for t2 in cod.tree(q):
                for shift in (-2, -1, 1, 2):
                    acc = {}
                    for (a, b), c in structure.turn(sym, shift, t1, t2).items():
                        for (a2, b2), c2 in structure.turn(sym, -shift, a, b).items():
                            acc[(a2, b2)] = acc.get((a2, b2), 0.0) + c * c2
                    for k, c in acc.items():
                        assert abs(c - (1.0 if k == (t1, t2) else 0.0)) < 1e-12


def test_decompose_reproduces_permutation():
    for p in itertools.permutations(range(5)):
        a = list(range(5))
        for j in decompose(p):
            a[j], a[j + 1] = a[j + 1], a[j]
        assert a == list(p)


# ----------------------------------------------------------------------------- tensors

def make_tensor(sym, seed, rank, corank, nsec=3, maxdim=3):
    r = random.Random(seed)
    legs = [rand_basis(sym, r, nsec, maxdim) for _ in range(rank + corank)]
    t = tensor.Tensor(sym, legs[:rank], legs[rank:])
    t.set_random(seed)
    return t


def test_oracle_u1_permute_dense_equivalence():
    # tests/tensor_tests.hpp:1-48: plain(T.permute<shift,p>()) == shuffle<p>(plain(T)), all 24
    # permutations x shifts -2..2 of a rank-(2,2) tensor
    sym = Sym(U1)
    t = make_tensor(sym, 21, 2, 2)
    dense = tensor.plain(t)
    for p in itertools.permutations(range(4)):
        for shift in (-2, -1, 0, 1, 2):
            tp = tensor.permute(t, shift, p)
            assert np.array_equal(tensor.plain(tp), np.transpose(dense, p))


@pytest.mark.parametrize("sym", SYMS, ids=lambda s: str(s.kinds))
def test_oracle_permute_round_trip(sym):
    t = make_tensor(sym, 33, 2, 2)
    for p in itertools.permutations(range(4)):
        for shift in (-2, -1, 0, 1, 2):
            tp = tensor.permute(t, shift, p)
            back = tensor.permute(tp, -shift, inverse(p))
            assert back.sector == t.sector
            np.testing.assert_allclose(back.data, t.data, rtol=0, atol=1e-13)


def test_oracle_su2_norm_is_invariant_under_permute():
    sym = Sym(SU2)
    t = make_tensor(sym, 44, 2, 2)
    n0 = tensor.squared_norm(t)
    for p in itertools.permutations(range(4)):
        for shift in (-1, 0, 1):
            assert abs(tensor.squared_norm(tensor.permute(t, shift, p)) - n0) < 1e-12 * n0


def test_get_permutations_examples():
    # A(l,p|r) . B(r,p'|r') over r -> (l,p|p',r')
    assert structure.get_permutations([-1, -2, 1], 2, [1, -3, -4], 2, 2) == ([0, 1, 2], 0, [0, 1, 2], 1, [0, 1, 2, 3], 0)
    # matrix product A(i|k) B(k|j)
    assert structure.get_permutations([-1, 1], 1, [1, -2], 1, 1) == ([0, 1], 0, [0, 1], 0, [0, 1], 0)
    # transpose contraction with reversed result order: A(k|i) B(j|k) -> (j|i)
    p1, s1, p2, s2, pres, sres = structure.get_permutations([1, -2], 1, [-1, 1], 1, 1)
    assert p1 == [1, 0] and p2 == [1, 0] and pres == [1, 0]


def test_truncation_rule_hand_built():
    s = Sym(U1)
    sv = [((0,), 1.0), ((0,), 0.5), ((1,), 0.8), ((1,), 0.1), ((-1,), 0.3)]
    kept, tw, ent = tensor.truncation_rule(s, sv, 3, 0.0, False)
    assert kept == {(0,): [1.0, 0.5], (1,): [0.8]}
    assert tw == 0.3 ** 2 + 0.1 ** 2
    # eps removes after maxKeep; truncWeight only counts the maxKeep cut
    kept, tw, _ = tensor.truncation_rule(s, sv, 4, 0.4, False)
    assert kept == {(0,): [1.0, 0.5], (1,): [0.8]} and tw == 0.1 ** 2
    # multiplet rule: relative gap > 0.1 cuts the last multiplet
    sv2 = [((0,), 1.0), ((1,), 0.5), ((-1,), 0.49)]
    kept, _, _ = tensor.truncation_rule(s, sv2, 3, 0.0, True)
    assert kept == {(0,): [1.0]}
    # all below eps: the first (pre-sort) entry survives
    kept, _, _ = tensor.truncation_rule(s, [((2,), 1e-3), ((0,), 2e-3)], 5, 1.0, False)
    assert kept == {(2,): [1e-3]}
    # SU(2): degeneracy weights the discarded weight and the entropy
    su = Sym(SU2)
    kept, tw, ent = tensor.truncation_rule(su, [((3,), 0.9), ((2,), 0.2)], 1, 0.0, False)
    assert tw == 2 * 0.2 ** 2 and ent == -3 * 0.9 * 0.9 * math.log(0.81)


def test_oracle_tsvd_reconstructs():
    sym = Sym(U1)
    t = make_tensor(sym, 55, 2, 2)
    U, S, V, allSV = tensor.full_svd(t)
    rec = tensor.mul(tensor.mul(U, S), V)
    np.testing.assert_allclose(rec.data, t.data, atol=1e-13)


# ---------------------------------------------------------------- rank-predictor network oracle
def test_mlp_oracle_gradient_matches_finite_differences():
    from oracle import mlp as om

    r = np.random.default_rng(4)
    nin, w = 5, 7
    n = om.layout(nin, w)["b3"][1]
    p = r.standard_normal(n) * 0.5
    X = r.standard_normal((11, nin))
    T = r.standard_normal(11)
    _, g = om.loss_and_grad(p, X, T, nin, w)
    h = 1e-6
    fd = np.array([(om.loss_and_grad(p + h * e, X, T, nin, w)[0] - om.loss_and_grad(p - h * e, X, T, nin, w)[0]) / (2 * h) for e in np.eye(n)])
    np.testing.assert_allclose(g, fd, rtol=1e-6, atol=1e-8)


def test_mlp_oracle_forward_by_hand():
    from oracle import mlp as om

    # nin = 1, width = 2: h1 = relu(x W1 + b1), h2 = relu(h1 W2 + b2), y = h2 . w3 + b3
    W1, b1 = np.array([[1.0, -2.0]]), np.array([0.5, 0.25])
    W2, b2 = np.array([[1.0, 0.0], [3.0, -1.0]]), np.array([0.0, 1.0])
    w3, b3 = np.array([2.0, -1.0]), 0.125
    p = np.concatenate([W1.ravel(), b1, W2.ravel(), b2, w3, [b3]])
    # x = 1: pre1 = (1.5, -1.75) -> h1 = (1.5, 0); pre2 = (1.5, 1.0) -> h2 = (1.5, 1.0); y = 3 - 1 + 0.125
    y, _ = om.forward(p, np.array([[1.0]]), 1, 2)
    assert y[0] == 2.125


def test_mlp_oracle_adam_first_step_moves_by_lr():
    from oracle import mlp as om

    g = np.array([0.3, -2.0, 1e-3])
    p1, _, _ = om.adam_step(np.zeros(3), g, np.zeros(3), np.zeros(3), 1, 0.01)
    np.testing.assert_allclose(p1, -0.01 * np.sign(g), rtol=1e-4)
