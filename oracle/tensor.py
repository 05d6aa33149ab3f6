"""Block-sparse tensor data operations (oracle), FP64 on the CPU with numpy/scipy.

Restates src/Core/Tensor.cpp (permute_impl :83-214, tSVD/full_svd_impl/cutoff_matrices
:478-633, tQR :757-780, teigh :782-803, twist :357-377), src/Core/TensorBase.cpp (trace/norms
:26-64, operator* :237-288), Tensor.hpp:328-351 (contract) and the storage layout of
include/Xped/Core/storage/StorageType_contiguous.hpp:79-97,161-171.
"""
from __future__ import annotations

import math
from typing import Dict, List, Optional, Tuple

import numpy as np
import scipy.linalg

from . import rng
from .structure import Qbasis, build_fusion_tree, get_permutations, pair_permute, tree_permute
from .stdsort import std_sort
from .symmetry import SU2, Sym, phase
from .wigner import wigner_3j

PERMUTE_SKIP = 1.0e-10  # Tensor.cpp:127,193


class Tensor:
    def __init__(self, sym: Sym, dom: List[Qbasis], cod: List[Qbasis], sectors: Optional[list] = None):
        self.sym = sym
        self.unc_dom = [b.copy() for b in dom]
        self.unc_cod = [b.copy() for b in cod]
        self.dom = build_fusion_tree(sym, self.unc_dom)
        self.cod = build_fusion_tree(sym, self.unc_cod)
        if sectors is None:  # StorageType::resize
            sectors = [q for q in self.dom.qs() if self.cod.find(q) >= 0]
        self.sector = [tuple(q) for q in sectors]
        self.offsets, self.rows, self.cols = [], [], []
        cur = 0
        for q in self.sector:
            r, c = self.dom.inner_dim(q), self.cod.inner_dim(q)
            self.offsets.append(cur)
            self.rows.append(r)
            self.cols.append(c)
            cur += r * c
        self.size = cur
        self.data = np.zeros(cur, dtype=np.float64)

    @property
    def rank(self):
        return len(self.unc_dom)

    @property
    def corank(self):
        return len(self.unc_cod)

    def find(self, q):
        q = tuple(q)
        for i, s in enumerate(self.sector):
            if s == q:
                return i
        return -1

    def block(self, i) -> np.ndarray:
        """Column-major view of block i."""
        o = self.offsets[i]
        return self.data[o:o + self.rows[i] * self.cols[i]].reshape(self.cols[i], self.rows[i]).T

    def set_block(self, i, m):
        o = self.offsets[i]
        self.data[o:o + self.rows[i] * self.cols[i]] = np.asarray(m, dtype=np.float64).T.reshape(-1)

    def set_random(self, seed: int, dist: int = 0):
        eng = rng.MT19937(seed)
        if dist == 0:
            self.data = rng.uniform(eng, self.size)
        else:
            self.data = rng.Normal().sample(eng, self.size)
        return self

    def copy(self):
        t = Tensor.__new__(Tensor)
        t.__dict__.update(self.__dict__)
        t.data = self.data.copy()
        return t

    def structure(self):
        return [(q, o, r, c) for q, o, r, c in zip(self.sector, self.offsets, self.rows, self.cols)]


def subblock(t: Tensor, i: int, t1, t2) -> np.ndarray:
    """Tensor::subBlock (Tensor.cpp:962-981): the (t1, t2) submatrix reshaped column-major."""
    r0 = t.dom.left_offset(t1)
    c0 = t.cod.left_offset(t2)
    m = t.block(i)[r0:r0 + t1.dim, c0:c0 + t2.dim]
    dims = tuple(t1.dims) + tuple(t2.dims)
    return np.reshape(m, dims, order="F") if dims else m.reshape(())


def _is_split(t: Tensor, shift: int, p) -> bool:
    if shift != 0:
        return False
    R = t.rank
    return all(x < R for x in p[:R]) and all(x >= R for x in p[R:])


def permute(t: Tensor, shift: int, p) -> Tensor:
    """Tensor::permute (Tensor.cpp:335-355) with both implementations of permute_impl."""
    p = list(p)
    R, C = t.rank, t.corank
    if shift == 0 and p == list(range(R + C)):
        return t.copy()
    newR, newC = R - shift, C + shift
    split = _is_split(t, shift, p)
    if split:
        nd = [t.unc_dom[p[i]] for i in range(newR)]
        nc = [t.unc_cod[p[i + newR] - R] for i in range(newC)]
    else:
        nd = [t.unc_cod[p[i] - R].conj() if p[i] >= R else t.unc_dom[p[i]] for i in range(newR)]
        nc = [t.unc_cod[p[i + newR] - R] if p[i + newR] >= R else t.unc_dom[p[i + newR]].conj() for i in range(newC)]
    out = Tensor(t.sym, nd, nc)
    pd = p[:R]
    pc = [x - R for x in p[R:]]
    for i, q in enumerate(t.sector):
        for t1 in t.dom.tree(q):
            for t2 in t.cod.tree(q):
                sub = subblock(t, i, t1, t2)
                sh = np.transpose(sub, p) if sub.ndim else sub
                if split:
                    pairs = [((a, b), ca * cb) for a, ca in tree_permute(t.sym, t1, pd).items()
                             for b, cb in tree_permute(t.sym, t2, pc).items()]
                else:
                    pairs = list(pair_permute(t.sym, shift, t1, t2, p).items())
                for (a, b), c in pairs:
                    if abs(c) < PERMUTE_SKIP:
                        continue
                    so = out.find(a.coupled)
                    r0 = out.dom.left_offset(a)
                    c0 = out.cod.left_offset(b)
                    blk = out.block(so)
                    blk[r0:r0 + a.dim, c0:c0 + b.dim] += c * np.reshape(sh, (a.dim, b.dim), order="F")
                    out.set_block(so, blk)
    return out


def block_gemm(a: np.ndarray, b: np.ndarray, c: np.ndarray) -> np.ndarray:
    """The per-sector product of operator* (MatrixInterface::prod, MatrixInterface_Eigen_impl.cpp:155-160):
    one FP64 BLAS GEMM, c = a @ b."""
    return np.matmul(a, b, out=c)


def mul(l: Tensor, r: Tensor) -> Tensor:
    """TensorBase::operator* (TensorBase.cpp:237-288)."""
    out = Tensor(l.sym, l.unc_dom, r.unc_cod)
    for i, q in enumerate(l.sector):
        j = r.find(q)
        if j < 0:
            continue
        o = out.find(q)
        a, b = l.block(i), r.block(j)
        out.set_block(o, out.block(o) + block_gemm(a, b, np.empty((a.shape[0], b.shape[1]))))
    return out


def contract(a: Tensor, a1, b: Tensor, a2, rankres: int) -> Tensor:
    """Tensor::contract (Tensor.hpp:328-351)."""
    p1, s1, p2, s2, pres, sres = get_permutations(a1, a.rank, a2, b.rank, rankres)
    return permute(mul(permute(a, s1, p1), permute(b, s2, p2)), sres, pres)


def adjoint(t: Tensor) -> Tensor:
    """AdjointOp evaluated into a Tensor (AdjointOp.hpp:25-69, Tensor.hpp:446-452)."""
    out = Tensor(t.sym, t.unc_cod, t.unc_dom, sectors=t.sector)
    for i in range(len(t.sector)):
        out.set_block(i, t.block(i).T)
    return out


def twist(t: Tensor, leg: int) -> Tensor:
    """Tensor::twist (Tensor.cpp:357-377)."""
    if not t.sym.fermionic:
        return t.copy()
    out = t.copy()
    for i, q in enumerate(t.sector):
        blk = out.block(i).copy()
        r0 = 0
        for t1 in t.dom.tree(q):
            c0 = 0
            for t2 in t.cod.tree(q):
                c = t.sym.coeff_twist(t1.unc[leg]) if leg < t.rank else t.sym.coeff_twist(t2.unc[leg - t.rank])
                blk[r0:r0 + t1.dim, c0:c0 + t2.dim] *= c
                c0 += t2.dim
            r0 += t1.dim
        out.set_block(i, blk)
    return out


def trace(t: Tensor) -> float:
    return sum(float(np.trace(t.block(i))) * t.sym.degeneracy(q) for i, q in enumerate(t.sector))


def squared_norm(t: Tensor) -> float:
    return sum(float(np.sum(t.block(i) ** 2)) * t.sym.degeneracy(q) for i, q in enumerate(t.sector))


def max_norm(t: Tensor) -> float:
    return float(np.max(np.abs(t.data))) if t.size else 0.0


def truncation_rule(sym: Sym, allSV, maxKeep, eps_svd, preserve_multiplets, truncWeight=0.0, entropy=0.0):
    """cutoff_matrices selection (Tensor.cpp:535-598)."""
    allSV = list(allSV)
    n = len(allSV)
    first = allSV[0]
    # std::sort is not stable: ties are broken exactly as libstdc++'s introsort breaks them
    std_sort(allSV, lambda a, b: a[1] > b[1])
    for q, s in allSV[maxKeep:]:
        truncWeight += sym.degeneracy(q) * math.pow(abs(s), 2.0)
    allSV = allSV[:min(maxKeep, n)]
    if eps_svd > 0.0:
        allSV = [x for x in allSV if not (x[1] < eps_svd)]
    if preserve_multiplets:
        end = -1
        for i in range(len(allSV) - 1, 0, -1):
            rel = 2 * (allSV[i - 1][1] - allSV[i][1]) / (allSV[i - 1][1] + allSV[i][1])
            if rel > 0.1:
                end = i
                break
        if end != -1:
            allSV = allSV[:end]
    if not allSV:
        allSV = [first]
    kept: Dict[tuple, list] = {}
    for q, s in allSV:
        kept.setdefault(q, []).append(s)
        # std::log(0) is -inf in C++ (the reference asserts against all-zero spectra, Tensor.cpp:587)
        entropy += -sym.degeneracy(q) * s * s * (math.log(s * s) if s * s > 0 else float("-inf"))
    kept = dict(sorted(kept.items()))
    return kept, truncWeight, entropy


def full_svd(t: Tensor):
    """full_svd_impl (Tensor.cpp:494-521) with LAPACK gesvd per sector."""
    middle = Qbasis(t.sym, 1)
    for i, q in enumerate(t.sector):
        middle.push_back(q, min(t.rows[i], t.cols[i]))
    U = Tensor(t.sym, t.unc_dom, [middle], sectors=t.sector)
    S = Tensor(t.sym, [middle], [middle], sectors=t.sector)
    V = Tensor(t.sym, [middle], t.unc_cod, sectors=t.sector)
    allSV = []
    for i, q in enumerate(t.sector):
        u, s, vt = scipy.linalg.svd(t.block(i), full_matrices=False, lapack_driver="gesvd")
        U.set_block(i, u)
        S.set_block(i, np.diag(s))
        V.set_block(i, vt)
        allSV += [(q, float(x)) for x in s]
    return U, S, V, allSV


def tsvd(t: Tensor, maxKeep, eps_svd, preserve_multiplets=False, truncWeight=0.0, entropy=0.0):
    """Tensor::tSVD (Tensor.cpp:478-633)."""
    U, S, V, allSV = full_svd(t)
    kept, tw, ent = truncation_rule(t.sym, allSV, maxKeep, eps_svd, preserve_multiplets, truncWeight, entropy)
    tb = Qbasis(t.sym, 1)
    for q, v in kept.items():
        tb.push_back(q, len(v))
    order = list(kept.keys())
    Ut = Tensor(t.sym, t.unc_dom, [tb], sectors=order)
    St = Tensor(t.sym, [tb], [tb], sectors=order)
    Vt = Tensor(t.sym, [tb], t.unc_cod, sectors=order)
    for o, q in enumerate(order):
        i = t.find(q)
        n = len(kept[q])
        Ut.set_block(o, U.block(i)[:, :n])
        St.set_block(o, S.block(i)[:n, :n])
        Vt.set_block(o, V.block(i)[:n, :])
    return Ut, St, Vt, tw, ent, kept


def tqr(t: Tensor):
    """Tensor::tQR (Tensor.cpp:757-780) with LAPACK Householder QR (same reflector sign rule
    as Eigen::HouseholderQR)."""
    middle = t.cod.forget_history() if t.dom.dim() > t.cod.dim() else t.dom.forget_history()
    Q = Tensor(t.sym, t.unc_dom, [middle], sectors=t.sector)
    R = Tensor(t.sym, [middle], t.unc_cod, sectors=t.sector)
    for i, q in enumerate(t.sector):
        m, n = t.rows[i], t.cols[i]
        mid = middle.inner_dim(q)
        qf, rf = scipy.linalg.qr(t.block(i), mode="full")
        qo = np.zeros((m, mid))
        k = min(m, mid)
        qo[:, :k] = qf[:, :k]
        ro = np.zeros((mid, n))
        kr = min(mid, rf.shape[0])
        ro[:kr, :] = np.triu(rf)[:kr, :]
        Q.set_block(i, qo)
        R.set_block(i, ro)
    return Q, R


def teigh(t: Tensor):
    """Tensor::teigh (Tensor.cpp:782-803)."""
    dh, ch = t.dom.forget_history(), t.cod.forget_history()
    D = Tensor(t.sym, [dh], [ch], sectors=t.sector)
    V = Tensor(t.sym, t.unc_dom, [ch], sectors=t.sector)
    for i, q in enumerate(t.sector):
        w, v = np.linalg.eigh(t.block(i), UPLO="L")
        D.set_block(i, np.diag(w))
        V.set_block(i, v)
    return D, V


def cgc(sym: Sym, q1, q2, q3) -> np.ndarray:
    """Clebsch-Gordan tensor of one fusion vertex (SU2.cpp:70-89): indices (m1, m2, m3), each
    running from the highest weight down (2m = q - 1 - 2i), entries 3j(q1 q2 q3; m1 m2 -m3) *
    (-1)^((q1 - q2 + 2m3)/2) * sqrt(q3) with the sympy-pinned exact 3j symbols.  Abelian
    components contribute the 1 x 1 x 1 unit."""
    out = np.ones((1, 1, 1))
    for c, k in enumerate(sym.kinds):
        if k != SU2:
            continue
        d1, d2, d3 = q1[c], q2[c], q3[c]
        v = np.zeros((d1, d2, d3))
        for i1 in range(d1):
            for i2 in range(d2):
                for i3 in range(d3):
                    m1, m2, m3 = d1 - 1 - 2 * i1, d2 - 1 - 2 * i2, d3 - 1 - 2 * i3
                    v[i1, i2, i3] = wigner_3j(d1 - 1, d2 - 1, d3 - 1, m1, m2, -m3) * phase((d1 - d2 + m3) // 2) * math.sqrt(d3)
        # product symmetries: the degeneracy index of a product irrep runs over the SU(2) factor
        out = np.einsum("abc,ijk->aibjck", out, v).reshape(out.shape[0] * d1, out.shape[1] * d2, out.shape[2] * d3)
    return out


def one_j(sym: Sym, q) -> np.ndarray:
    """The 1j tensor of a dual leg (SU2.cpp:41-61): entry (i, j) = (-1)^((q - 2m_i)/2) for
    m_i = -m_j, truncating the halved exponent as C++ integer division does."""
    out = np.ones((1, 1))
    for c, k in enumerate(sym.kinds):
        if k != SU2:
            continue
        d = q[c]
        v = np.zeros((d, d))
        for i in range(d):
            for j in range(d):
                im, jm = d - 1 - 2 * i, d - 1 - 2 * j
                if im == -jm:
                    v[i, j] = phase((d - im) // 2)
        out = np.kron(out, v)
    return out


def tree_tensor(sym: Sym, tree) -> np.ndarray:
    """FusionTree::asTensor (FusionTree.cpp:154-307): the fold of vertex tensors over the
    intermediates, legs (m_1 .. m_r, m_coupled), with the 1j tensor applied to every dual leg."""
    r = tree.rank
    if r == 0:
        return np.ones((1,))
    if r == 1:
        out = np.eye(sym.degeneracy(tree.unc[0]))
    else:
        outs = list(tree.inter) + [tree.coupled]
        out = cgc(sym, tree.unc[0], tree.unc[1], outs[0])
        for k in range(2, r):
            out = np.tensordot(out, cgc(sym, outs[k - 2], tree.unc[k], outs[k - 1]), axes=([k], [0]))
    for i in range(r):
        if tree.dual[i]:
            out = np.moveaxis(np.tensordot(one_j(sym, tree.unc[i]), out, axes=([1], [i])), 0, i)
    return out


def plain(t: Tensor) -> np.ndarray:
    """Dense expansion Tensor::plainTensor (Tensor.cpp:1133-1310).  Every leg is sorted; within a
    leg, sector q starts at full_outer_num(q) and the index of (inner state a, weight m) is
    a * degeneracy(q) + m.  A (domain tree, codomain tree) subblock of sector q contributes
    sub[a.., b..] * sum_mc T1[m.., mc] T2[n.., mc] with T1, T2 the trees' Clebsch-Gordan
    tensors (tree_tensor).  For abelian symmetries every tensor is the unit and this is the plain
    placement of subblocks."""
    legs = t.unc_dom + t.unc_cod
    sorted_legs = []
    for b in legs:
        c = b.copy()
        c.sort()
        sorted_legs.append(c)
    out = np.zeros([b.full_dim() for b in sorted_legs])
    R, C = t.rank, t.corank
    for i, q in enumerate(t.sector):
        for t1 in t.dom.tree(q):
            T1 = tree_tensor(t.sym, t1)
            for t2 in t.cod.tree(q):
                sub = subblock(t, i, t1, t2)
                G = np.tensordot(T1, tree_tensor(t.sym, t2), axes=([R], [C]))  # (m.., n..)
                full = np.multiply.outer(sub, G)
                n = R + C
                # interleave (a_i, m_i) -> a_i * deg_i + m_i on every leg
                full = np.transpose(full, [x for i in range(n) for x in (i, n + i)]) if n else full
                shape, idx = [], []
                for leg, qq in enumerate(tuple(t1.unc) + tuple(t2.unc)):
                    b = sorted_legs[leg]
                    deg = t.sym.degeneracy(qq)
                    shape.append(b.inner_dim(qq) * deg)
                    o = b.full_outer_num(qq)
                    idx.append(slice(o, o + b.inner_dim(qq) * deg))
                out[tuple(idx)] += full.reshape(shape) if n else full
    return out
