"""A symmetry-free dense CTMRG written from the geometry of the 2x2-cell double-layer network, the
checker of the oracle's block-sparse CTMRG (oracle/ctm.py).

Site tensors are the dense expansions A[l, t, r, b, s] of the U(1) site tensors (oracle plain()).
Bond (x, y)-(x+1, y) joins A(x, y).r with A(x+1, y).l, bond (x, y)-(x, y+1) joins A(x, y).b with
A(x, y+1).t; the norm network pairs every ket A with a bra A on the same site.  Environment legs,
here (not the reference's order): C1[r, d], C2[l, d], C3[l, u], C4[u, r], T1[l, r, d, d'],
T2[u, d, l, l'], T3[l, r, u, u'], T4[u, d, r, r'] (primed: bra).

init(): the FROM_A environment (CTM.cpp:137-240): every corner and edge tensor is one site's
double layer with its outward ket and bra legs joined (traced), so the network of an RDM is the
finite patch of sites around the bond with traced boundary.

grow_all(): the four moves of CTM.cpp:386-420 in the reference's order (columns x = 0, 1 for the
left and right moves, rows y = 0, 1 for the top and bottom moves, all sites of a column or row
from the tensors before that column or row), each absorbing one row or column of sites WITHOUT
projectors: the absorbed legs are fused in the order (old, ket, bra), so bonds grow by D^2 per
absorbed site.  Where the reference's projectors are exact (chi above every bond dimension, full-
rank corner products), its step computes the same network.

bond_energies(): <S_i . S_j> of the 4 horizontal and 4 vertical bonds of the cell from the full
networks around each bond (physical index 0 = 2Sz -1, 1 = +1)."""
import numpy as np

LX = LY = 2


def idx(x, y):
    return (x % LX) * LY + (y % LY)


def net(pairs, out):
    """Contract tensors with named legs in the listed order, two at a time (legs named twice are
    summed): pairs = [(array, [names])], out = [names]."""
    acc, acc_l = pairs[0]
    for arr, labels in pairs[1:]:
        assert arr.ndim == len(labels), (arr.shape, labels)
        keep = [n for n in acc_l if n not in labels] + [n for n in labels if n not in acc_l]
        ids = {n: i for i, n in enumerate(dict.fromkeys(acc_l + labels))}
        acc = np.einsum(acc, [ids[n] for n in acc_l], arr, [ids[n] for n in labels], [ids[n] for n in keep])
        acc_l = keep
    return np.einsum(acc, list(range(len(acc_l))), [acc_l.index(n) for n in out])


def heisenberg():
    sz = np.diag([-0.5, 0.5])
    sp = np.array([[0.0, 0.0], [1.0, 0.0]])  # S+ |down> = |up>
    return (np.kron(sz, sz) + 0.5 * (np.kron(sp, sp.T) + np.kron(sp.T, sp))).reshape(2, 2, 2, 2)


class DenseCTM:
    def __init__(self, A_dense):
        self.A = list(A_dense)

    def a(self, x, y):
        A = self.A[idx(x, y)]
        return np.einsum("ltrbs,LTRBs->lLtTrRbB", A, A)

    def a_open(self, x, y):
        A = self.A[idx(x, y)]
        return np.einsum("ltrbs,LTRBS->lLtTrRbBsS", A, A)

    @staticmethod
    def _fuse(t, groups):
        """reshape consecutive axis groups into one axis each"""
        s = t.shape
        dims, k = [], 0
        for g in groups:
            d = 1
            for _ in range(g):
                d *= s[k]
                k += 1
            dims.append(d)
        return t.reshape(dims)

    @staticmethod
    def _n(t):
        return t / np.abs(t).max()

    def init(self):
        self.C1, self.C2, self.C3, self.C4, self.T1, self.T2, self.T3, self.T4 = ({} for _ in range(8))
        f = self._fuse
        for x in range(LX):
            for y in range(LY):
                a, p = self.a(x, y), idx(x, y)
                self.C1[p] = f(np.einsum("llttrRbB->rRbB", a), [2, 2])
                self.C2[p] = f(np.einsum("lLttrrbB->lLbB", a), [2, 2])
                self.C3[p] = f(np.einsum("lLtTrrbb->lLtT", a), [2, 2])
                self.C4[p] = f(np.einsum("lltTrRbb->tTrR", a), [2, 2])
                self.T1[p] = f(np.einsum("lLttrRbB->lLrRbB", a), [2, 2, 1, 1])
                self.T2[p] = f(np.einsum("lLtTrrbB->tTbBlL", a), [2, 2, 1, 1])
                self.T3[p] = f(np.einsum("lLtTrRbb->lLrRtT", a), [2, 2, 1, 1])
                self.T4[p] = f(np.einsum("lltTrRbB->tTbBrR", a), [2, 2, 1, 1])
        return self

    def left_move(self):
        f, n = self._fuse, self._n
        for x in range(LX):
            C1n, T4n, C4n = {}, {}, {}
            for y in range(LY):
                C1n[idx(x, y - 1)] = n(f(np.einsum("od,olkK->ldkK", self.C1[idx(x - 1, y - 1)], self.T1[idx(x, y - 1)]), [1, 3]))
                T4n[idx(x, y)] = n(f(np.einsum("udlL,lLtTrRbB->utTdbBrR", self.T4[idx(x - 1, y)], self.a(x, y)), [3, 3, 1, 1]))
                C4n[idx(x, y + 1)] = n(f(np.einsum("uo,orvV->uvVr", self.C4[idx(x - 1, y + 1)], self.T3[idx(x, y + 1)]), [3, 1]))
            for y in range(LY):
                p = idx(x, y)
                self.C1[p], self.T4[p], self.C4[p] = C1n[p], T4n[p], C4n[p]

    def right_move(self):
        f, n = self._fuse, self._n
        for x in range(LX):
            C2n, T2n, C3n = {}, {}, {}
            for y in range(LY):
                C2n[idx(x, y - 1)] = n(f(np.einsum("lokK,oe->lekK", self.T1[idx(x, y - 1)], self.C2[idx(x + 1, y - 1)]), [1, 3]))
                T2n[idx(x, y)] = n(f(np.einsum("lLtTrRbB,udrR->utTdbBlL", self.a(x, y), self.T2[idx(x + 1, y)]), [3, 3, 1, 1]))
                C3n[idx(x, y + 1)] = n(f(np.einsum("louU,ov->lvuU", self.T3[idx(x, y + 1)], self.C3[idx(x + 1, y + 1)]), [1, 3]))
            for y in range(LY):
                p = idx(x, y)
                self.C2[p], self.T2[p], self.C3[p] = C2n[p], T2n[p], C3n[p]

    def top_move(self):
        f, n = self._fuse, self._n
        for y in range(LY):
            C1n, T1n, C2n = {}, {}, {}
            for x in range(LX):
                C1n[idx(x - 1, y)] = n(f(np.einsum("ro,odsS->rsSd", self.C1[idx(x - 1, y - 1)], self.T4[idx(x - 1, y)]), [3, 1]))
                T1n[idx(x, y)] = n(f(np.einsum("mntT,lLtTrRbB->mlLnrRbB", self.T1[idx(x, y - 1)], self.a(x, y)), [3, 3, 1, 1]))
                C2n[idx(x + 1, y)] = n(f(np.einsum("mo,odlL->mlLd", self.C2[idx(x + 1, y - 1)], self.T2[idx(x + 1, y)]), [3, 1]))
            for x in range(LX):
                p = idx(x, y)
                self.C1[p], self.T1[p], self.C2[p] = C1n[p], T1n[p], C2n[p]

    def bottom_move(self):
        f, n = self._fuse, self._n
        for y in range(LY):
            C4n, T3n, C3n = {}, {}, {}
            for x in range(LX):
                C4n[idx(x - 1, y)] = n(f(np.einsum("uorR,om->umrR", self.T4[idx(x - 1, y)], self.C4[idx(x - 1, y + 1)]), [1, 3]))
                T3n[idx(x, y)] = n(f(np.einsum("lLtTrRbB,mnbB->mlLnrRtT", self.a(x, y), self.T3[idx(x, y + 1)]), [3, 3, 1, 1]))
                C3n[idx(x + 1, y)] = n(f(np.einsum("uolL,mo->mlLu", self.T2[idx(x + 1, y)], self.C3[idx(x + 1, y + 1)]), [3, 1]))
            for x in range(LX):
                p = idx(x, y)
                self.C4[p], self.T3[p], self.C3[p] = C4n[p], T3n[p], C3n[p]

    def grow_all(self):
        self.left_move()
        self.right_move()
        self.top_move()
        self.bottom_move()
        return self

    def _bond(self, N):
        H = heisenberg()
        return float(np.einsum("abAB,ABab->", N, H) / np.einsum("abab->", N))

    def bond_energies(self):
        """[horizontal bonds at idx 0..3] + [vertical bonds at idx 0..3]"""
        eh, ev = [], []
        for x in range(LX):
            for y in range(LY):
                bottom = net([(self.C4[idx(x - 1, y + 1)], ["t4d", "c4r"]), (self.T3[idx(x, y + 1)], ["c4r", "bb", "1b", "1B"]),
                              (self.T3[idx(x + 1, y + 1)], ["bb", "bc", "2b", "2B"]), (self.C3[idx(x + 2, y + 1)], ["bc", "t2d"])],
                             ["t4d", "1b", "1B", "2b", "2B", "t2d"])
                N = net([(self.C1[idx(x - 1, y - 1)], ["c1r", "c1d"]), (self.T1[idx(x, y - 1)], ["c1r", "tt", "1t", "1T"]),
                         (self.T1[idx(x + 1, y - 1)], ["tt", "tc", "2t", "2T"]), (self.C2[idx(x + 2, y - 1)], ["tc", "c2d"]),
                         (self.T4[idx(x - 1, y)], ["c1d", "t4d", "1l", "1L"]),
                         (self.a_open(x, y), ["1l", "1L", "1t", "1T", "12", "12'", "1b", "1B", "s1", "S1"]),
                         (self.a_open(x + 1, y), ["12", "12'", "2t", "2T", "2r", "2R", "2b", "2B", "s2", "S2"]),
                         (self.T2[idx(x + 2, y)], ["c2d", "t2d", "2r", "2R"]), (bottom, ["t4d", "1b", "1B", "2b", "2B", "t2d"])],
                        ["s1", "s2", "S1", "S2"])
                eh.append(self._bond(N))
        for x in range(LX):
            for y in range(LY):
                bottom = net([(self.C4[idx(x - 1, y + 2)], ["t4d", "c4r"]), (self.T3[idx(x, y + 2)], ["c4r", "bc", "2b", "2B"]),
                              (self.C3[idx(x + 1, y + 2)], ["bc", "t2d"])], ["t4d", "2b", "2B", "t2d"])
                N = net([(self.C1[idx(x - 1, y - 1)], ["c1r", "c1d"]), (self.T1[idx(x, y - 1)], ["c1r", "tc", "1t", "1T"]),
                         (self.C2[idx(x + 1, y - 1)], ["tc", "c2d"]), (self.T4[idx(x - 1, y)], ["c1d", "ll", "1l", "1L"]),
                         (self.a_open(x, y), ["1l", "1L", "1t", "1T", "1r", "1R", "12", "12'", "s1", "S1"]),
                         (self.T2[idx(x + 1, y)], ["c2d", "rr", "1r", "1R"]), (self.T4[idx(x - 1, y + 1)], ["ll", "t4d", "2l", "2L"]),
                         (self.a_open(x, y + 1), ["2l", "2L", "12", "12'", "2r", "2R", "2b", "2B", "s2", "S2"]),
                         (self.T2[idx(x + 1, y + 1)], ["rr", "t2d", "2r", "2R"]), (bottom, ["t4d", "2b", "2B", "t2d"])],
                        ["s1", "s2", "S1", "S2"])
                ev.append(self._bond(N))
        return eh + ev
