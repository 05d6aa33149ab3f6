"""Integer structure of block-sparse tensors (oracle): fusion trees, bases, tree-pair moves.

Restates src/Core/FusionTree.cpp, src/Core/Qbasis.cpp, src/Core/TensorHelper.cpp:14-44,
src/Core/treepair.cpp:35-217 and include/Xped/Util/Permutations.hpp:195-225.
"""
from __future__ import annotations

from dataclasses import dataclass, field, replace
from typing import Dict, List, Tuple

from .symmetry import Sym

TREE_EPS = 1.0e-12  # ScalarTraits<double>::epsilon() (ScalarTraits.hpp:17)


@dataclass(frozen=True)
class FTree:
    unc: tuple
    inter: tuple
    coupled: tuple
    dims: tuple
    dual: tuple

    @property
    def rank(self):
        return len(self.unc)

    @property
    def dim(self):
        d = 1
        for x in self.dims:
            d *= x
        return d

    def key(self):
        # FusionTree::operator== without pedantic asserts ignores IS_DUAL (FusionTree.cpp:66-83)
        return (self.unc, self.inter, self.coupled, self.dims)

    def less(self, o):
        # FusionTree::operator< (FusionTree.cpp:32-64)
        return (self.unc, self.inter, self.coupled) < (o.unc, o.inter, o.coupled)

    def enlarge(self, other, coupled):
        # FusionTree::enlarge (FusionTree.cpp:309-324)
        r = self.rank + 1
        inter = self.inter + ((self.coupled,) if r >= 3 else ())
        return FTree(self.unc + other.unc, inter, coupled, self.dims + other.dims, self.dual + other.dual)


def compute_intermediates(s: Sym, unc):
    """FusionTree::computeIntermediates (FusionTree.cpp:136-152), abelian only."""
    r = len(unc)
    if r == 0:
        return (), tuple(0 for _ in range(s.nq))
    inter = []
    for i in range(max(r - 2, 0)):
        inter.append(s.combine(unc[0], unc[1])[0] if i == 0 else s.combine(inter[i - 1], unc[i + 1])[0])
    if r < 2:
        coupled = unc[0]
    elif r == 2:
        coupled = s.combine(unc[0], unc[1])[0]
    else:
        coupled = s.combine(inter[-1], unc[-1])[0]
    return tuple(inter), coupled


def decompose(pi):
    """Adjacent-swap decomposition (Permutations.hpp:195-225)."""
    a = list(range(len(pi)))
    out = []
    if a == list(pi):
        return out
    for i in range(len(pi)):
        index = a.index(pi[i])
        if index == 0:
            continue
        for j in range(index - 1, i - 1, -1):
            a[j], a[j + 1] = a[j + 1], a[j]
            out.append(j)
    assert a == list(pi)
    return out


def inverse(pi):
    inv = [0] * len(pi)
    for i, x in enumerate(pi):
        inv[x] = i
    return inv


class Qbasis:
    """Qbasis<Symmetry, depth> (Qbasis.hpp:37-268)."""

    def __init__(self, sym: Sym, depth: int = 1):
        self.sym = sym
        self.depth = depth
        self.conj_flag = False
        self.curr_dim = 0
        self.data: List[list] = []  # [q, num, dim, trees]

    def copy(self):
        b = Qbasis(self.sym, self.depth)
        b.conj_flag = self.conj_flag
        b.curr_dim = self.curr_dim
        b.data = [[e[0], e[1], e[2], list(e[3])] for e in self.data]
        return b

    def find(self, q):
        for i, e in enumerate(self.data):
            if e[0] == q:
                return i
        return -1

    def inner_dim(self, q):
        i = self.find(q)
        return 0 if i < 0 else self.data[i][2]

    def dim(self):
        return sum(e[2] for e in self.data)

    def full_dim(self):
        return sum(e[2] * self.sym.degeneracy(e[0]) for e in self.data)

    def qs(self):
        return [e[0] for e in self.data]

    def tree(self, q):
        return self.data[self.find(q)][3]

    def push_back(self, q, inner):
        """Qbasis.cpp:34-73"""
        q = tuple(q)
        if inner == 0:
            return
        i = self.find(q)
        if i < 0:
            if self.depth == 1:
                t = FTree((q,), (), q, (inner,), (self.conj_flag,))
            else:
                t = FTree((), (), q, (), ())
            self.data.append([q, self.curr_dim, inner, [t]])
        else:
            self.data[i][2] += inner
            for j in range(i + 1, len(self.data)):
                self.data[j][1] += inner
            if self.depth == 1:
                t = self.data[i][3][0]
                self.data[i][3][0] = replace(t, dims=(t.dims[0] + inner,))
        self.curr_dim += inner

    def push_back_trees(self, q, inner, trees):
        assert self.find(q) < 0
        self.data.append([tuple(q), self.curr_dim, inner, list(trees)])
        self.curr_dim += inner

    def sort(self):
        """Qbasis.cpp:306-331"""
        if len(self.data) < 2:
            return
        s = self.sym
        if self.conj_flag:
            keyf = lambda e: tuple(s.conj(e[0]))
        else:
            keyf = lambda e: tuple(e[0])
        self.data.sort(key=keyf)
        acc = 0
        for e in self.data:
            e[1] = acc
            acc += e[2]

    def conj(self):
        """Qbasis.cpp:418-439"""
        out = Qbasis(self.sym, 1)
        out.conj_flag = not self.conj_flag
        for q, num, dim, trees in self.data:
            old = trees[0]
            cq = self.sym.conj(q)
            out.push_back_trees(cq, dim, [FTree((cq,), (), cq, old.dims, (not old.dual[0],))])
        out.sort()
        return out

    def combine(self, other):
        """Qbasis.cpp:441-495"""
        out = Qbasis(self.sym, self.depth + 1)
        trees: Dict[tuple, list] = {}
        for q1, _, _, t1s in self.data:
            for q2, _, _, t2s in other.data:
                for q in self.sym.combine(q1, q2):
                    for t1 in t1s:
                        for t2 in t2s:
                            tt = t1.enlarge(t2, q)
                            lst = trees.setdefault(q, [])
                            if all(x.key() != tt.key() for x in lst):
                                lst.append(tt)
        for q, lst in trees.items():
            lst.sort(key=lambda t: (t.unc, t.inter, t.coupled))
            out.push_back_trees(q, sum(t.dim for t in lst), lst)
        out.sort()
        return out

    def forget_history(self):
        out = Qbasis(self.sym, 1)
        for q, _, dim, _ in self.data:
            out.push_back(q, dim)
        return out

    def left_offset(self, tree: FTree):
        """Qbasis.cpp:208-225"""
        out = 0
        for t in self.tree(tree.coupled):
            if t.key() == tree.key():
                return out
            out += t.dim
        raise KeyError("tree not in basis")

    def full_outer_num(self, q):
        out = 0
        for p, _, dim, _ in self.data:
            if p == q:
                return out
            out += dim * self.sym.degeneracy(p)
        return out


def build_fusion_tree(s: Sym, uncoupled: List[Qbasis]) -> Qbasis:
    """internal::build_FusionTree (TensorHelper.cpp:14-44)."""
    if not uncoupled:
        b = Qbasis(s, 0)
        b.push_back(s.vacuum(), 1)
        return b
    b = uncoupled[0]
    for u in uncoupled[1:]:
        b = b.combine(u)
    return b


# ----------------------------------------------------------------------------- tree moves

def tree_swap(s: Sym, t: FTree, pos: int) -> Dict[FTree, float]:
    """FusionTree::swap (FusionTree.cpp:353-412)."""
    R = t.rank
    unc = list(t.unc)
    dual = list(t.dual)
    dims = list(t.dims)
    unc[pos], unc[pos + 1] = unc[pos + 1], unc[pos]
    dual[pos], dual[pos + 1] = dual[pos + 1], dual[pos]
    dims[pos], dims[pos + 1] = dims[pos + 1], dims[pos]
    out = {}
    if pos == 0:
        qf = t.coupled if R == 2 else t.inter[0]
        c = s.coeff_swap(t.unc[0], t.unc[1], qf)
        if abs(c) < TREE_EPS:
            return out
        out[FTree(tuple(unc), t.inter, t.coupled, tuple(dims), tuple(dual))] = c
        return out
    q1 = t.unc[0] if pos == 1 else t.inter[pos - 2]
    q2 = t.unc[pos]
    q3 = t.unc[pos + 1]
    Q = t.coupled if pos == R - 2 else t.inter[pos]
    Q12 = t.inter[pos - 1]
    for Q31 in s.combine(q3, q1):
        inter = list(t.inter)
        inter[pos - 1] = Q31
        cgc = s.coeff_swap(Q12, q3, Q) * s.coeff_recouple(q3, q1, q2, Q, Q31, Q12) * s.coeff_swap(q3, q1, Q31)
        if abs(cgc) < TREE_EPS:
            continue
        out[FTree(tuple(unc), tuple(inter), t.coupled, tuple(dims), tuple(dual))] = cgc
    return out


def tree_permute(s: Sym, t: FTree, pi) -> Dict[FTree, float]:
    """FusionTree::permute (FusionTree.cpp:326-351)."""
    out = {t: 1.0}
    for pos in decompose(pi):
        tmp: Dict[FTree, float] = {}
        for tree, c in out.items():
            for tree2, c2 in tree_swap(s, tree, pos).items():
                if tree2 not in tmp:
                    tmp[tree2] = c * c2
                elif abs(tmp[tree2] + c * c2) < TREE_EPS:
                    del tmp[tree2]
                else:
                    tmp[tree2] += c * c2
        out = tmp
    return out


def turn_right(s: Sym, t1: FTree, t2: FTree):
    """treepair::turn_right (treepair.cpp:35-88)."""
    R, C = t1.rank, t2.rank
    if C == 1:
        a = s.vacuum()
    elif C == 2:
        a = t2.unc[0]
    else:
        a = t2.inter[-1]
    inter1 = t1.inter + ((t1.coupled,) if R + 1 > 2 else ())
    t1p = FTree(t1.unc + (s.conj(t2.unc[-1]),), inter1, a, t1.dims + (t2.dims[-1],), t1.dual + (not t2.dual[-1],))
    t2p = FTree(t2.unc[:-1], t2.inter[:-1] if C > 2 else (), a, t2.dims[:-1], t2.dual[:-1])
    b = t2.unc[-1]
    c = t2.coupled
    coeff = s.coeff_turn(a, b, c)
    if t2.dual[-1]:
        coeff *= s.coeff_FS(s.conj(b))
    if abs(coeff) < TREE_EPS:
        return {}
    return {(t1p, t2p): coeff}


def turn_left(s: Sym, t1: FTree, t2: FTree):
    return {(b, a): c for (a, b), c in turn_right(s, t2, t1).items()}


def turn(s: Sym, shift: int, t1: FTree, t2: FTree):
    """treepair::turn (treepair.cpp:103-141)."""
    if shift == 0:
        return {(t1, t2): 1.0}
    out = {}
    first = turn_left(s, t1, t2) if shift > 0 else turn_right(s, t1, t2)
    nxt = shift - 1 if shift > 0 else shift + 1
    for (a, b), c1 in first.items():
        for pr, c2 in turn(s, nxt, a, b).items():
            out.setdefault(pr, c1 * c2)
    return out


def pair_permute(s: Sym, shift: int, t1: FTree, t2: FTree, p):
    """treepair::permute (treepair.cpp:143-217)."""
    R, C = t1.rank, t2.rank
    newR, newC = R - shift, C + shift
    if not s.non_abelian and not s.fermionic:
        du, dd, dl = [], [], []
        for r in range(newR):
            src = p[r]
            du.append(t1.unc[src] if src < R else s.conj(t2.unc[src - R]))
            dd.append(t1.dims[src] if src < R else t2.dims[src - R])
            dl.append(t1.dual[src] if src < R else (not t2.dual[src - R]))
        cu, cd, cl = [], [], []
        for c in range(newC):
            src = p[c + newR]
            cu.append(s.conj(t1.unc[src]) if src < R else t2.unc[src - R])
            cd.append(t1.dims[src] if src < R else t2.dims[src - R])
            cl.append((not t1.dual[src]) if src < R else t2.dual[src - R])
        di, dc = compute_intermediates(s, du)
        ci, cc = compute_intermediates(s, cu)
        return {(FTree(tuple(du), di, dc, tuple(dd), tuple(dl)), FTree(tuple(cu), ci, cc, tuple(cd), tuple(cl))): 1.0}
    N = R + C
    pi_id = list(range(N))
    for i in range(C):
        pi_id[i + R] = (C - 1) - i + R
    pi_tmp = [pi_id[p[i]] for i in range(N)]
    pi_corr = pi_tmp[:newR] + [pi_tmp[(newC - 1) - i + newR] for i in range(newC)]
    out = {}
    for (t1p, triv), c1 in turn(s, -C, t1, t2).items():
        for t1pp, c2 in tree_permute(s, t1p, pi_corr).items():
            for fin, c3 in turn(s, newC, t1pp, triv).items():
                out[fin] = out.get(fin, 0.0) + c1 * c2 * c3
    return out


def get_permutations(a1, rank1, a2, rank2, rankres):
    """util::constFct::get_permutations (Constfct.hpp:55-109) at run time."""
    r1, r2 = len(a1), len(a2)
    max1 = max(0, max(a1))
    max2 = max(0, max(a2))
    assert max1 == max2
    p2 = [0] * r2
    found2 = [-1] * r2
    for ind in range(max2):
        val = list(a2).index(ind + 1)
        found2[ind] = val
        p2[ind] = val
    for ind in range(max2, r2):
        val = next(i for i in range(r2) if i not in found2)
        found2[ind] = val
        p2[ind] = val
    p1 = [0] * r1
    found1 = [-1] * r1
    for ind in range(max1):
        val = list(a1).index(ind + 1)
        found1[ind] = val
        p1[r1 - max1 + ind] = val
    for ind in range(max1, r1):
        val = next(i for i in range(r1) if i not in found1)
        found1[ind] = val
        p1[ind - max1] = val
    tmp = [x for x in a1 if x < 0] + [x for x in a2 if x < 0]
    pres = sorted(range(len(tmp)), key=lambda i: -tmp[i])
    shift1 = rank1 - r1 + max1
    shift2 = rank2 - max2
    shiftres = r1 - max1 - rankres
    return p1, shift1, p2, shift2, pres, shiftres
