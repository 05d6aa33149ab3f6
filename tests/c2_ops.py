"""The oracle's operations for the shared C2 index lists (xped_amd/su2_layer.py)."""
from oracle import ctm as octm
from oracle import tensor as oten


class OracleOps:
    def __init__(self, sym):
        self.sym = sym

    def contract(self, a, a1, b, a2, rank):
        return oten.contract(a, a1, b, a2, rank)

    def permute(self, t, shift, p):
        return oten.permute(t, shift, p)

    def adjoint(self, t):
        return oten.adjoint(t)

    def twist(self, t, leg):
        return oten.twist(t, leg)

    def leg(self, t, i):
        return (t.unc_dom + t.unc_cod)[i]

    def fuse(self, a, b):
        return a.combine(b).forget_history()

    def identity(self, dom, cod):
        return octm.identity(self.sym, dom, cod)
