"""Host structure engine (C++ product, through the C ABI) vs the oracle: bit-exact integer
structure -- bases, fusion trees, sector lists, offsets and permute plans -- plus the
symmetry coefficients and Wigner symbols.  Runs on a structure-only runtime (no GPU)."""
import ctypes as C
import itertools
import json
import math
import random
from pathlib import Path

import numpy as np
import pytest

from oracle import structure as ost
from oracle import tensor as oten
from oracle.structure import Qbasis as OQbasis
from oracle.symmetry import FU1, SU2, U1, Sym

import xped_amd.tensor as xt
from xped_amd._lib import lib

GOLDEN = Path(__file__).parent / "golden"

SYMCASES = [((U1,), xt.U1), ((SU2,), xt.SU2), ((FU1, U1), xt.FU1xU1), ((FU1,), xt.FU1)]


def rand_legs(kinds, seed, n, nsec=3, maxdim=3):
    r = random.Random(seed)
    legs = []
    for _ in range(n):
        entries = []
        for _ in range(nsec):
            if kinds == (SU2,):
                q = (r.randint(1, 4),)
            elif len(kinds) == 2:
                q = (r.randint(-3, 3), r.randint(-2, 2))
            else:
                q = (r.randint(-4, 4),)
            entries.append((q, r.randint(1, maxdim)))
        legs.append(entries)
    return legs


def both_bases(kinds, pcode, entries):
    o = OQbasis(Sym(*kinds), 1)
    p = xt.Qbasis(pcode)
    for q, d in entries:
        o.push_back(q, d)
        p.push_back(q, d)
    return o, p


def assert_same_basis(o: OQbasis, p: xt.Qbasis):
    pe = p.entries()
    assert [(e[0], e[1], e[2]) for e in pe] == [(tuple(e[0]), e[1], e[2]) for e in o.data]
    assert p.info()["conj"] == o.conj_flag
    for i, e in enumerate(o.data):
        ptrees = p.trees(i)
        assert len(ptrees) == len(e[3])
        for pt, ot in zip(ptrees, e[3]):
            assert pt["unc"] == tuple(tuple(x) for x in ot.unc)
            assert pt["inter"] == tuple(tuple(x) for x in ot.inter)
            assert pt["coupled"] == tuple(ot.coupled)
            assert pt["dims"] == tuple(ot.dims)
            assert pt["dual"] == tuple(ot.dual)


# ----------------------------------------------------------------------------- Wigner / coefficients

def test_product_wigner_matches_sympy(xlib):
    fx = json.loads((GOLDEN / "wigner_fixtures.json").read_text())
    L = lib()
    for key, fn, tol in (("6j", L.xt_wigner_6j, 4e-15), ("3j", L.xt_wigner_3j, 4e-15), ("9j", L.xt_wigner_9j, 2e-14)):
        bad = [(a, v) for a, v in fx[key] if abs(fn(*a) - v) > max(1e-15, tol * abs(v))]
        assert not bad, (key, bad[:5])


@pytest.mark.parametrize("kinds,pcode", SYMCASES, ids=["U1", "SU2", "FU1xU1", "FU1"])
def test_product_coefficients_match_oracle(xlib, kinds, pcode):
    s = Sym(*kinds)
    L = lib()
    r = random.Random(3)

    def q2(q):
        q = tuple(q) + (0,) * (2 - len(q))
        return (C.c_int * 2)(*q)

    def rq():
        if kinds == (SU2,):
            return (r.randint(1, 6),)
        if len(kinds) == 2:
            return (r.randint(-3, 3), r.randint(-3, 3))
        return (r.randint(-5, 5),)

    for _ in range(3000):
        a, b = rq(), rq()
        for c in s.combine(a, b) + [rq()]:
            assert L.xt_coeff(pcode[0], pcode[1], 0, q2(a), q2(b), q2(c), None, None, None) == s.coeff_swap(a, b, c)
            assert L.xt_coeff(pcode[0], pcode[1], 2, q2(a), q2(b), q2(c), None, None, None) == s.coeff_turn(a, b, c)
        qs = [rq() for _ in range(6)]
        v = L.xt_coeff(pcode[0], pcode[1], 1, *[q2(q) for q in qs])
        w = s.coeff_recouple(*qs)
        assert abs(v - w) <= 4e-15 * max(1.0, abs(w))
        assert L.xt_coeff(pcode[0], pcode[1], 3, q2(a), None, None, None, None, None) == s.coeff_FS(a)
        assert L.xt_coeff(pcode[0], pcode[1], 4, q2(a), None, None, None, None, None) == s.coeff_twist(a)


# ----------------------------------------------------------------------------- bases

@pytest.mark.parametrize("kinds,pcode", SYMCASES, ids=["U1", "SU2", "FU1xU1", "FU1"])
def test_bases_combine_conj_bit_exact(xlib, kinds, pcode):
    for seed in range(6):
        legs = rand_legs(kinds, seed, 3)
        ob, pb = zip(*[both_bases(kinds, pcode, e) for e in legs])
        for o, p in zip(ob, pb):
            assert_same_basis(o, p)
            assert_same_basis(o.conj(), p.conj())
            assert_same_basis(o.conj().conj(), p.conj().conj())
        oc = ob[0].combine(ob[1]).combine(ob[2].conj())
        pc = pb[0].combine(pb[1]).combine(pb[2].conj())
        assert_same_basis(oc, pc)
        assert_same_basis(oc.forget_history(), pc.forgetHistory())
        assert pc.fullDim() == oc.full_dim() == ob[0].full_dim() * ob[1].full_dim() * ob[2].full_dim()


@pytest.mark.parametrize("kinds,pcode", SYMCASES, ids=["U1", "SU2", "FU1xU1", "FU1"])
def test_tree_permute_matches_oracle(xlib, kinds, pcode):
    s = Sym(*kinds)
    legs = rand_legs(kinds, 17, 4, nsec=2, maxdim=2)
    ob, pb = zip(*[both_bases(kinds, pcode, e) for e in legs])
    o4 = ost.build_fusion_tree(s, list(ob))
    p4 = pb[0].combine(pb[1]).combine(pb[2]).combine(pb[3])
    assert_same_basis(o4, p4)
    L = lib()
    cap = 64
    nq = len(kinds)
    for p in itertools.permutations(range(4)):
        for i, e in enumerate(o4.data):
            for t, tree in enumerate(e[3]):
                n = C.c_int()
                unc = (C.c_int * (2 * 10 * cap))()
                inter = (C.c_int * (2 * 10 * cap))()
                cp = (C.c_int * (2 * cap))()
                co = (C.c_double * cap)()
                rc = L.xt_tree_permute(p4.h, i, t, (C.c_int * 4)(*p), 4, cap, C.byref(n), unc, inter, cp, co)
                assert rc == 0
                got = {}
                for k in range(n.value):
                    u = tuple(tuple(unc[2 * 10 * k + 2 * j: 2 * 10 * k + 2 * j + nq]) for j in range(4))
                    it = tuple(tuple(inter[2 * 10 * k + 2 * j: 2 * 10 * k + 2 * j + nq]) for j in range(2))
                    got[(u, it)] = co[k]
                exp = {(tt.unc, tt.inter): c for tt, c in ost.tree_permute(s, tree, p).items()}
                assert set(got) == set(exp)
                for k in exp:
                    assert abs(got[k] - exp[k]) <= 1e-13, (p, k)


# ----------------------------------------------------------------------------- tensors and plans

def both_tensors(kinds, pcode, rt, seed, rank, corank, nsec=3, maxdim=3):
    legs = rand_legs(kinds, seed, rank + corank, nsec, maxdim)
    ob, pb = zip(*[both_bases(kinds, pcode, e) for e in legs])
    ot = oten.Tensor(Sym(*kinds), list(ob[:rank]), list(ob[rank:]))
    pt = xt.Tensor(rt, list(pb[:rank]), list(pb[rank:]))
    return ot, pt


def oracle_plan(t: oten.Tensor, shift, p):
    """Gather plan derived from the oracle: per output subblock (dst offset, rows, cols) the list
    of (source offset, coefficient) in accumulation order (Tensor.cpp:179-212)."""
    out = oten.permute(t, shift, p)  # structure + validation
    R = t.rank
    tasks = {}
    order = []
    for so, q in enumerate(out.sector):
        r0 = 0
        for a in out.dom.tree(q):
            c0 = 0
            for b in out.cod.tree(q):
                key = out.offsets[so] + r0 + c0 * out.rows[so]
                tasks[key] = (a.dim, b.dim, [])
                order.append(key)
                c0 += b.dim
            r0 += a.dim
    split = oten._is_split(t, shift, list(p))
    for i, q in enumerate(t.sector):
        r0 = 0
        for t1 in t.dom.tree(q):
            c0 = 0
            for t2 in t.cod.tree(q):
                src = t.offsets[i] + r0 + c0 * t.rows[i]
                if split:
                    pairs = [((a, b), ca * cb) for a, ca in ost.tree_permute(t.sym, t1, list(p[:R])).items()
                             for b, cb in ost.tree_permute(t.sym, t2, [x - R for x in p[R:]]).items()]
                else:
                    pairs = list(ost.pair_permute(t.sym, shift, t1, t2, list(p)).items())
                for (a, b), c in pairs:
                    if abs(c) < 1e-10:
                        continue
                    so = out.find(a.coupled)
                    key = out.offsets[so] + out.dom.left_offset(a) + out.cod.left_offset(b) * out.rows[so]
                    tasks[key][2].append((src, c))
                c0 += t2.dim
            r0 += t1.dim
    return out, [(k,) + tasks[k][:2] + (tasks[k][2],) for k in order]


@pytest.mark.parametrize("kinds,pcode", SYMCASES, ids=["U1", "SU2", "FU1xU1", "FU1"])
def test_tensor_structure_and_permute_plans(xlib, rt_struct, kinds, pcode):
    exact = kinds != (SU2,)
    for seed, (rank, corank) in zip(range(3), [(2, 2), (2, 1), (1, 3)]):
        ot, pt = both_tensors(kinds, pcode, rt_struct, 100 + seed, rank, corank)
        assert [(s[0], s[1], s[2], s[3]) for s in pt.sectors()] == [(tuple(q), o, r, c) for q, o, r, c in ot.structure()]
        N = rank + corank
        for p in itertools.permutations(range(N)):
            for shift in range(-corank, rank + 1):
                out, tasks = oracle_plan(ot, shift, p)
                pp = pt.permute(shift, p)
                assert [(s[0], s[1], s[2], s[3]) for s in pp.sectors()] == [(tuple(q), o, r, c) for q, o, r, c in out.structure()]
                plan = pt.permute_plan(shift, p)
                assert list(plan["task_dst"]) == [t[0] for t in tasks]
                assert list(plan["task_rows"]) == [t[1] for t in tasks]
                assert list(plan["task_cols"]) == [t[2] for t in tasks]
                assert list(plan["task_nsrc"]) == [len(t[3]) for t in tasks]
                srcs = [s for t in tasks for s in t[3]]
                assert list(plan["src_off"]) == [s[0] for s in srcs]
                if exact:
                    assert list(plan["src_coeff"]) == [s[1] for s in srcs]
                else:
                    np.testing.assert_allclose(plan["src_coeff"], [s[1] for s in srcs], rtol=0, atol=1e-13)


def test_get_permutations_matches_oracle(xlib):
    cases = [([-1, -2, 1], 2, [1, -3, -4], 2, 2), ([-1, 1], 1, [1, -2], 1, 1), ([1, -2], 1, [-1, 1], 1, 1),
             ([-1, -2, -3, -4, 1], 2, [-5, -6, 1, -7, -8], 3, 8), ([1, -1, -2, -3, 2, -4, -5, -6], 4, [-7, 1, 2], 1, 6),
             ([2, -1, 1, -2], 2, [1, 2, -3], 2, 2), ([-2, 1, -1], 1, [-3, 1], 1, 2)]
    for a1, r1, a2, r2, rr in cases:
        assert xt.get_permutations(a1, r1, a2, r2, rr) == ost.get_permutations(a1, r1, a2, r2, rr)


def test_truncation_rule_matches_oracle(xlib):
    r = random.Random(5)
    for trial in range(200):
        kinds, pcode = SYMCASES[trial % 2]
        s = Sym(*kinds)
        sv = []
        for _ in range(r.randint(1, 40)):
            q = (r.randint(1, 4),) if kinds == (SU2,) else (r.randint(-3, 3),)
            # coarse values create exact ties between sectors
            sv.append((q, r.choice([0.5, 0.25, 1.0, r.random(), r.random() * 1e-11])))
        mk = r.randint(1, 30)
        eps = r.choice([0.0, 1e-10, 0.3])
        pm = r.random() < 0.5
        kept_o, tw_o, en_o = oten.truncation_rule(s, sv, mk, eps, pm, 0.25, 0.5)
        kept_p, tw_p, en_p = xt.truncation_rule(pcode, sv, mk, eps, pm, 0.25, 0.5)
        assert kept_p == {q: len(v) for q, v in kept_o.items()}
        assert tw_p == tw_o or abs(tw_p - tw_o) <= 1e-15 * abs(tw_o)
        assert en_p == en_o or abs(en_p - en_o) <= 1e-15 * max(1.0, abs(en_o))


@pytest.mark.parametrize("kinds,pcode", SYMCASES, ids=["U1", "SU2", "FU1xU1", "FU1"])
def test_permute_plans_reused_across_dimension_changes(xlib, kinds, pcode):
    """A runtime keeps the dimension-free part of each permute plan by the charges of the input legs;
    a tensor with the same charges and other dimensions (a truncation changed the sector sizes)
    gets its plan from that part.  The result must equal the oracle's plan for the new dimensions."""
    rt = xt.Runtime(-1)
    exact = kinds != (SU2,)
    rank, corank = 2, 2
    legs = rand_legs(kinds, 7, rank + corank, 3, 3)
    for scale in (1, 2, 3):
        # same charges in the same order, dimensions scaled
        scaled = [[(q, d * scale + (i % 2)) for i, (q, d) in enumerate(e)] for e in legs]
        ob, pb = zip(*[both_bases(kinds, pcode, e) for e in scaled])
        ot = oten.Tensor(Sym(*kinds), list(ob[:rank]), list(ob[rank:]))
        pt = xt.Tensor(rt, list(pb[:rank]), list(pb[rank:]))
        for p, shift in [((1, 0, 2, 3), 0), ((3, 1, 2, 0), 1), ((2, 3, 0, 1), -1), ((0, 2, 1, 3), 0)]:
            out, tasks = oracle_plan(ot, shift, p)
            plan = pt.permute_plan(shift, p)
            assert list(plan["task_dst"]) == [t[0] for t in tasks]
            assert list(plan["task_rows"]) == [t[1] for t in tasks]
            assert list(plan["task_cols"]) == [t[2] for t in tasks]
            assert list(plan["task_nsrc"]) == [len(t[3]) for t in tasks]
            srcs = [s for t in tasks for s in t[3]]
            assert list(plan["src_off"]) == [s[0] for s in srcs]
            if exact:
                assert list(plan["src_coeff"]) == [s[1] for s in srcs]
            else:
                np.testing.assert_allclose(plan["src_coeff"], [s[1] for s in srcs], rtol=0, atol=1e-13)
    st = rt.host_stats()
    assert st["n_permute_plans"] == 12 and st["symbolic_hits"] == 8


def test_plan_cache_is_bounded_and_rebuilds_evicted_plans(xlib):
    """The runtime drops the least recently used half of its permute plans past the bound; a
    dropped plan is rebuilt on its next use and equals the oracle's plan."""
    rt = xt.Runtime(-1)
    rt.set_max_plans(8)
    kinds, pcode = (U1,), xt.U1
    rank, corank = 2, 2
    legs = rand_legs(kinds, 9, rank + corank, 3, 3)
    ob, pb = zip(*[both_bases(kinds, pcode, e) for e in legs])
    ot = oten.Tensor(Sym(*kinds), list(ob[:rank]), list(ob[rank:]))
    pt = xt.Tensor(rt, list(pb[:rank]), list(pb[rank:]))
    cases = [(p, shift) for p in itertools.permutations(range(4)) for shift in (-1, 0, 1)][:40]
    for p, shift in cases:
        pt.permute_plan(shift, p)
        assert rt.cache_sizes()[0] <= 8
    nperm, _, evicted = rt.cache_sizes()
    assert evicted >= len(cases) - 8
    # the first plans were dropped; rebuilding them gives the oracle's plans
    for p, shift in cases[:3]:
        out, tasks = oracle_plan(ot, shift, p)
        plan = pt.permute_plan(shift, p)
        assert list(plan["task_dst"]) == [t[0] for t in tasks]
        assert list(plan["src_off"]) == [s[0] for t in tasks for s in t[3]]
    with pytest.raises(Exception):
        rt.set_max_plans(2)
