"""Reverse-mode AD tensors: the mirror of Xped's ``Tensor<..., ENABLE_AD=true>``
(include/Xped/AD/ADTensor.hpp) on the device path.

The reference wraps each value in a stan-math ``vari_value<Tensor>`` and registers one reverse
callback per operation (``stan::math::reverse_pass_callback``); ``stan::math::grad(res.vi_)`` then
runs the callbacks in reverse order.  Here a :class:`Tape` holds the callbacks, an
:class:`ADTensor` holds a device value and a device adjoint, and every callback is a call into the
C ABI (``xt_permute_adj``, ``xt_mul_bwd``, ``xt_svd_bwd``, ``xt_cutoff_bwd``, ``xt_diag_bwd``,
``xt_trace_bwd``, ``xt_max_norm_bwd``, ``xt_adjoint``, ``xt_twist``, ``xt_axpby``), so the reverse
pass runs on the GPU.  Adjoints accumulate (``+=``) as in the reference.

Operations and their reverse passes (ADTensor.hpp line numbers):

* ``adjoint`` (:138-150), ``permute`` (:171-185), ``contract`` (:193-219, composed of permutes and
  ``operator*`` exactly as the reference composes it), ``operator*`` (:703-763);
* ``tSVD`` = ``full_svd_impl`` (:237-329) + ``cutoff_matrices`` (:331-374);
* ``norm`` (:492-499, ``sqrt((t * t.adjoint()).trace())``), ``maxNorm`` (:450-476),
  ``trace`` (:531-548), ``diag_sqrt`` / ``diag_inv`` (:560-590), ``twist`` (:592-602);
* ``coeff`` (the scalar entry used by the reference's tests, ADTensor.hpp:152-169);
* ``tQR``, which the reference does not record (its AD tensor has no tQR); its reverse pass is the
  standard QR adjoint (``xt_qr_bwd``).
"""
from __future__ import annotations

import math
from typing import Callable, List, Optional, Sequence

from .tensor import Tensor, cutoff_bwd, diag_bwd, get_permutations, max_norm_bwd, mul_bwd, qr_bwd, svd_bwd, trace_bwd


class Tape:
    """The reverse-callback stack (stan-math's ChainableStack for one nested region)."""

    def __init__(self):
        self.callbacks: List[Callable[[], None]] = []

    def record(self, fn: Callable[[], None]):
        self.callbacks.append(fn)

    def grad(self, root: "Var"):
        """stan::math::grad(root.vi_): seed d root = 1 and run the callbacks in reverse order."""
        root.adj += 1.0
        for fn in reversed(self.callbacks):
            fn()
        self.callbacks.clear()


class Var:
    """A scalar with an adjoint (stan::math::var_value<double>)."""

    def __init__(self, tape: Tape, val: float):
        self.tape, self.val, self.adj = tape, float(val), 0.0

    def __add__(self, other: "Var") -> "Var":
        out = Var(self.tape, self.val + other.val)

        def back(a=self, b=other, o=out):
            a.adj += o.adj
            b.adj += o.adj

        self.tape.record(back)
        return out

    def sqrt(self) -> "Var":
        out = Var(self.tape, math.sqrt(self.val))

        def back(x=self, o=out):
            x.adj += o.adj * 0.5 / o.val

        self.tape.record(back)
        return out


class ADTensor:
    """A device tensor value with a lazily allocated device adjoint."""

    def __init__(self, tape: Tape, val: Tensor):
        self.tape, self.val = tape, val
        self._adj: Optional[Tensor] = None

    @property
    def adj(self) -> Tensor:
        if self._adj is None:
            self._adj = self.val.zeros_like()
        return self._adj

    def accumulate(self, t: Tensor):
        # a fresh buffer either way: t may share its storage with another adjoint (an identity
        # permute returns its input), and the in-place reverse passes write into self.adj
        self._adj = t.scale(1.0) if self._adj is None else t.axpby(self._adj, 1.0, 1.0)

    def _new(self, val: Tensor) -> "ADTensor":
        return ADTensor(self.tape, val)

    # ---- structure-changing operations
    def adjoint(self) -> "ADTensor":
        out = self._new(self.val.adjoint())
        self.tape.record(lambda x=self, o=out: x.accumulate(o.adj.adjoint()))
        return out

    def permute(self, shift: int, p: Sequence[int]) -> "ADTensor":
        out = self._new(self.val.permute(shift, p))
        p = list(p)
        self.tape.record(lambda x=self, o=out: x.accumulate(o.adj.permute_adj(shift, p)))
        return out

    def __mul__(self, other):
        if isinstance(other, ADTensor):
            return self.mul(other)
        return self.scale(float(other))

    def mul(self, other: "ADTensor", precision: str = "fp64") -> "ADTensor":
        out = self._new(self.val.mul(other.val, precision))

        def back(a=self, b=other, o=out):
            mul_bwd(a.val, b.val, o.adj, a.adj, b.adj, precision)

        self.tape.record(back)
        return out

    def contract(self, other: "ADTensor", a1: Sequence[int], a2: Sequence[int], rankres: int, precision: str = "fp64") -> "ADTensor":
        """ADTensor::contract (ADTensor.hpp:193-219): permute, operator*, permute, each recorded."""
        perms = get_permutations(a1, self.val.rank, a2, other.val.rank, rankres)
        p1, shift1, p2, shift2, pres, shiftres = perms
        left = self.permute(shift1, p1)
        right = other.permute(shift2, p2)
        return left.mul(right, precision).permute(shiftres, pres)

    def scale(self, alpha: float) -> "ADTensor":
        out = self._new(self.val.scale(alpha))
        self.tape.record(lambda x=self, o=out: x.accumulate(o.adj.scale(alpha)))
        return out

    def twist(self, leg: int) -> "ADTensor":
        out = self._new(self.val.twist(leg))
        self.tape.record(lambda x=self, o=out: x.accumulate(o.adj.twist(leg)))
        return out

    def diag_sqrt(self) -> "ADTensor":
        out = self._new(self.val.diag_sqrt())
        self.tape.record(lambda x=self, o=out: diag_bwd(o.val, o.adj, x.adj, "sqrt"))
        return out

    def diag_inv(self) -> "ADTensor":
        out = self._new(self.val.diag_inv())
        self.tape.record(lambda x=self, o=out: diag_bwd(o.val, o.adj, x.adj, "inv"))
        return out

    # ---- SVD
    def full_svd_impl(self):
        U, S, V = self.val.full_svd_impl()
        aU, aS, aV = self._new(U), self._new(S), self._new(V)

        def back(x=self, u=aU, s=aS, v=aV):
            svd_bwd(u.val, s.val, v.val, u.adj, s.adj, v.adj, x.adj)

        self.tape.record(back)
        return aU, aS, aV

    def cutoff_matrices(self, U: "ADTensor", S: "ADTensor", V: "ADTensor", maxKeep: int, eps_svd: float, PRESERVE_MULTIPLETS: bool = True):
        Ut, St, Vt, tw, en = self.val.cutoff_matrices(U.val, S.val, V.val, maxKeep, eps_svd, 0.0, 0.0, PRESERVE_MULTIPLETS)
        aU, aS, aV = self._new(Ut), self._new(St), self._new(Vt)

        def back(u=U, s=S, v=V, tu=aU, ts=aS, tv=aV):
            cutoff_bwd(tu.adj, ts.adj, tv.adj, u.adj, s.adj, v.adj)

        self.tape.record(back)
        return aU, aS, aV, tw, en

    def tSVD(self, maxKeep: int, eps_svd: float, PRESERVE_MULTIPLETS: bool = True):
        """ADTensor::tSVD (ADTensor.hpp:221-233): full_svd_impl then cutoff_matrices."""
        U, S, V = self.full_svd_impl()
        return self.cutoff_matrices(U, S, V, maxKeep, eps_svd, PRESERVE_MULTIPLETS)

    # ---- QR (no AD counterpart in the reference: xt_qr_bwd is the adjoint a recorded tQR needs)
    def tQR(self):
        Q, R = self.val.tQR()
        aQ, aR = self._new(Q), self._new(R)

        def back(x=self, q=aQ, r=aR):
            qr_bwd(x.val, q.val, r.val, q.adj, r.adj, x.adj)

        self.tape.record(back)
        return aQ, aR

    # ---- scalars
    def trace(self) -> Var:
        out = Var(self.tape, self.val.trace())
        self.tape.record(lambda x=self, o=out: trace_bwd(x.adj, o.adj))
        return out

    def squaredNorm(self) -> Var:
        return self.mul(self.adjoint()).trace()

    def norm(self) -> Var:
        return self.squaredNorm().sqrt()

    def maxNorm(self) -> Var:
        out = Var(self.tape, self.val.maxNorm())
        self.tape.record(lambda x=self, o=out: max_norm_bwd(x.val, x.adj, o.adj))
        return out

    def coeff(self, block: int, row: int, col: int) -> Var:
        idx = self.val.flat_index(block, row, col)
        out = Var(self.tape, self.val.coeff(block, row, col))
        self.tape.record(lambda x=self, o=out: x.adj.add_at(idx, o.adj))
        return out
