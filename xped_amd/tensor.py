"""Python mirror of the reference's Xped::Tensor / Qbasis interface over the C ABI.

Names, argument meaning and error behaviour follow include/Xped/Core/Tensor.hpp:59-444,
TensorBase.hpp:47-135 and Qbasis.hpp:37-268.  All data lives on the GPU; structural queries
also work on a structure-only runtime (``Runtime(device=-1)``), which has no data.
"""
from __future__ import annotations

import ctypes as C
from typing import Dict, List, Sequence, Tuple

import numpy as np

from . import _lib
from ._lib import XH_BF16X3, XH_FP32, XH_FP64, XH_OZAKI, check, lib

# symmetry kinds (component codes of the C ABI)
U1 = (1, 0)          # U1<SpinU1> / U1<ChargeU1>
FU1 = (2, 0)         # U1<FChargeU1>
SU2 = (3, 0)         # SU2<SpinSU2>
FU1xU1 = (2, 1)      # Combined<U1<FChargeU1>, U1<SpinU1>>
U1xU1 = (1, 1)

PRECISION = {"fp64": XH_FP64, "fp32": XH_FP32, "bf16x3": XH_BF16X3, "ozaki": XH_OZAKI}


def _q2(q) -> C.Array:
    q = tuple(q) if isinstance(q, (tuple, list)) else (q,)
    a = (C.c_int * 2)(0, 0)
    for i, x in enumerate(q):
        a[i] = int(x)
    return a


def _ints(v) -> C.Array:
    v = list(v)
    return (C.c_int * max(1, len(v)))(*v)


def _check_xh(rc, ctx, what):
    if rc != _lib.XH_OK:
        msg = lib().xh_last_error(ctx)
        raise _lib.XpedError(rc, f"{what}: {msg.decode() if msg else ''}")


class Graph:
    """A captured sequence of grouped launches on a runtime's stream (xh_graph_*): launch()
    replays it on the buffers named at capture time."""

    def __init__(self, rt: "Runtime", h):
        self.rt, self.h = rt, h

    def launch(self):
        _check_xh(lib().xh_graph_launch(self.rt.ctx(), self.h), self.rt.ctx(), "graph launch")

    def close(self):
        if self.h:
            lib().xh_graph_destroy(self.rt.ctx(), self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Runtime:
    """Device context plus plan caches (one per host thread)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        rc = lib().xt_rt_create(device, C.byref(h))
        check(rc, None, "xt_rt_create")
        self.h = h
        self.device = device

    @property
    def has_device(self) -> bool:
        return self.device >= 0

    def ctx(self):
        return lib().xt_rt_ctx(self.h)

    def counters(self):
        f, b = C.c_double(), C.c_double()
        ng, npm = C.c_long(), C.c_long()
        lib().xt_rt_counters(self.h, C.byref(f), C.byref(b), C.byref(ng), C.byref(npm))
        return {"gemm_flops": f.value, "permute_bytes": b.value, "n_gemm": ng.value, "n_permute": npm.value}

    def host_stats(self):
        a, b = C.c_double(), C.c_double()
        na, nb = C.c_long(), C.c_long()
        lib().xt_rt_host_stats(self.h, C.byref(a), C.byref(b), C.byref(na), C.byref(nb))
        t = (C.c_double * 8)()
        lib().xt_rt_host_timers(self.h, t, 8)
        return {"permute_plan_s": a.value, "gemm_plan_s": b.value, "n_permute_plans": na.value, "n_gemm_plans": nb.value,
                "breakdown_s": [round(x, 4) for x in list(t[:5]) + list(t[6:8])], "symbolic_hits": int(t[5])}

    def device_cache(self, limit_bytes=None):
        """Counters of the context's cache of freed device blocks: allocations it served (hits), those
        that reached the HIP runtime (misses) and the bytes it holds idle.  With limit_bytes, first
        bounds the idle bytes (0 keeps none; the default bound is 4 GiB)."""
        if limit_bytes is not None:
            _check_xh(lib().xh_set_device_cache(self.ctx(), int(limit_bytes)), self.ctx(), "device cache")
        h, m, b = C.c_int64(), C.c_int64(), C.c_size_t()
        _check_xh(lib().xh_device_cache_stats(self.ctx(), C.byref(h), C.byref(m), C.byref(b)), self.ctx(), "device cache")
        return {"hits": h.value, "misses": m.value, "idle_bytes": b.value}

    def graph_begin(self):
        """Start capturing the runtime's stream into a HIP graph (see Graph)."""
        _check_xh(lib().xh_graph_begin(self.ctx()), self.ctx(), "graph capture")

    def graph_end(self) -> "Graph":
        """End the capture and instantiate it; a capture that recorded work which allocates or
        synchronises (a new plan, a split-K GEMM) raises instead."""
        g = C.c_void_p()
        _check_xh(lib().xh_graph_end(self.ctx(), C.byref(g)), self.ctx(), "graph capture")
        return Graph(self, g)

    def marker(self, id: int):
        """An empty kernel of `id` workgroups on the runtime's stream: a mark in a kernel trace."""
        check(lib().xt_rt_marker(self.h, int(id)), self.h, "marker")

    def gemm_timing(self, enable: bool, reset: bool = False):
        """HIP-event timing of every grouped GEMM launch; returns (kernel ms, flops, logged contraction blocks)."""
        ms, fl, n = C.c_double(), C.c_double(), C.c_long()
        lib().xt_rt_gemm_timing(self.h, int(enable), int(reset), C.byref(ms), C.byref(fl), C.byref(n))
        return ms.value, fl.value, n.value

    def permute_timing(self, enable: bool, reset: bool = False):
        """HIP-event timing of every permute launch; returns (kernel ms, bytes moved, launches)."""
        ms, b, n = C.c_double(), C.c_double(), C.c_int64()
        check(lib().xh_permute_timing(self.ctx(), int(enable), int(reset), C.byref(ms), C.byref(b), C.byref(n)), self.h, "permute timing")
        return ms.value, b.value, n.value

    def gemm_launches(self) -> int:
        """Number of xh_gemm calls timed since the last reset."""
        ms, fl, n = C.c_double(), C.c_double(), C.c_int64()
        lib().xh_gemm_timing(self.ctx(), -1, 0, C.byref(ms), C.byref(fl), C.byref(n))
        return n.value

    def gemm_log(self) -> np.ndarray:
        """Block shapes (m, n, k) of the contraction GEMMs logged while timing was on."""
        ms, fl, nn = C.c_double(), C.c_double(), C.c_long()
        lib().xt_rt_gemm_timing(self.h, -1, 0, C.byref(ms), C.byref(fl), C.byref(nn))
        out = np.zeros((max(nn.value, 1), 3), np.int32)
        lib().xt_rt_gemm_log(self.h, out.ctypes.data_as(C.POINTER(C.c_int32)), nn.value)
        return out[:nn.value]

    def synchronize(self):
        if self.has_device:
            lib().xh_synchronize(self.ctx())

    def clear_caches(self):
        check(lib().xt_rt_clear_caches(self.h), self.h, "clear_caches")

    def set_max_plans(self, n: int):
        """Bound of the permute and GEMM plan caches (least recently used half dropped past it)."""
        check(lib().xt_rt_set_max_plans(self.h, int(n)), self.h, "set_max_plans")

    def fuse_permute(self, enable=None):
        """contract() with the left operand's permute read inside skinny FP64 products (the gather
        GEMM, on by default; bit-identical either way).  enable: True / 1 (gather), 2 (gather and
        the product's own permute folded into its stores), False / 0 (off), None (keep).
        Returns the number of contractions run fused so far."""
        n = C.c_long()
        e = -1 if enable is None else (2 if enable == 2 else (1 if enable else 0))
        check(lib().xt_rt_set_fuse_permute(self.h, e, C.byref(n)), self.h, "fuse_permute")
        return n.value

    def cache_sizes(self):
        """(cached permute plans, cached GEMM plans, plans dropped so far)."""
        a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
        check(lib().xt_rt_cache_sizes(self.h, C.byref(a), C.byref(b), C.byref(c)), self.h, "cache_sizes")
        return a.value, b.value, c.value

    def __del__(self):
        h = getattr(self, "h", None)
        f = _lib.native("xt_rt_destroy") if (_lib is not None and h) else None
        if f is not None:
            f(h)
            self.h = None


class Qbasis:
    """Qbasis<Symmetry, depth> handle."""

    def __init__(self, sym=U1, handle=None):
        self.sym = tuple(sym)
        if handle is None:
            h = C.c_void_p()
            check(lib().xt_basis_create(self.sym[0], self.sym[1], C.byref(h)), None, "xt_basis_create")
            handle = h
        self.h = handle

    def __del__(self):
        h = getattr(self, "h", None)
        f = _lib.native("xt_basis_destroy") if (_lib is not None and h) else None
        if f is not None:
            f(h)
            self.h = None

    def push_back(self, q, dim: int):
        check(lib().xt_basis_push_back(self.h, _q2(q), int(dim)), None, "push_back")
        return self

    def _new(self, fn, *args):
        h = C.c_void_p()
        check(fn(self.h, *args, C.byref(h)), None, fn.__name__)
        return Qbasis(self.sym, h)

    def conj(self) -> "Qbasis":
        return self._new(lib().xt_basis_conj)

    def combine(self, other: "Qbasis") -> "Qbasis":
        return self._new(lib().xt_basis_combine, other.h)

    def forgetHistory(self) -> "Qbasis":
        return self._new(lib().xt_basis_forget_history)

    def copy(self) -> "Qbasis":
        return self._new(lib().xt_basis_copy)

    def sort(self):
        check(lib().xt_basis_sort(self.h), None, "sort")
        return self

    def info(self):
        d, n, conj = C.c_int(), C.c_int(), C.c_int()
        dim, fd = C.c_size_t(), C.c_size_t()
        lib().xt_basis_info(self.h, C.byref(d), C.byref(n), C.byref(dim), C.byref(fd), C.byref(conj))
        return {"depth": d.value, "nsec": n.value, "dim": dim.value, "full_dim": fd.value, "conj": bool(conj.value)}

    def dim(self):
        return self.info()["dim"]

    def fullDim(self):
        return self.info()["full_dim"]

    def nq(self):
        return 2 if self.sym[1] else 1

    def entries(self):
        """[(q, num, dim, ntrees)] in basis order."""
        out = []
        n = self.info()["nsec"]
        q = (C.c_int * 2)()
        num, dim = C.c_size_t(), C.c_size_t()
        nt = C.c_int()
        for i in range(n):
            lib().xt_basis_entry(self.h, i, q, C.byref(num), C.byref(dim), C.byref(nt))
            out.append((tuple(q[: self.nq()]), num.value, dim.value, nt.value))
        return out

    def trees(self, i: int):
        """Fusion trees of sector i: list of dicts (unc, inter, coupled, dims, dual, dim)."""
        nq = self.nq()
        ents = self.entries()
        res = []
        for t in range(ents[i][3]):
            rank = C.c_int()
            unc = (C.c_int * 40)()
            inter = (C.c_int * 40)()
            cp = (C.c_int * 2)()
            dims = (C.c_uint32 * 20)()
            dual = (C.c_int * 20)()
            dim = C.c_uint64()
            check(lib().xt_basis_tree(self.h, i, t, C.byref(rank), unc, inter, cp, dims, dual, C.byref(dim)), None, "tree")
            r = rank.value
            res.append({
                "unc": tuple(tuple(unc[2 * k: 2 * k + nq]) for k in range(r)),
                "inter": tuple(tuple(inter[2 * k: 2 * k + nq]) for k in range(max(r - 2, 0))),
                "coupled": tuple(cp[:nq]),
                "dims": tuple(dims[:r]),
                "dual": tuple(bool(x) for x in dual[:r]),
                "dim": dim.value,
            })
        return res


def _basis_sym(h):
    k0, k1 = C.c_int(), C.c_int()
    lib().xt_basis_sym(h, C.byref(k0), C.byref(k1))
    return (k0.value, k1.value)


class Tensor:
    """Tensor<double, Rank, CoRank, Symmetry> with GPU-resident data."""

    def __init__(self, rt: Runtime, domain: Sequence[Qbasis] = (), codomain: Sequence[Qbasis] = (), handle=None):
        self.rt = rt
        if handle is None:
            legs = list(domain) + list(codomain)
            arr = (C.c_void_p * max(1, len(legs)))(*[b.h.value if isinstance(b.h, C.c_void_p) else b.h for b in legs])
            h = C.c_void_p()
            check(lib().xt_tensor_create(rt.h, len(domain), len(codomain), arr, C.byref(h)), rt.h, "Tensor()")
            handle = h
        self.h = handle
        r, c, n = C.c_int(), C.c_int(), C.c_int()
        s = C.c_size_t()
        lib().xt_tensor_shape(self.h, C.byref(r), C.byref(c), C.byref(n), C.byref(s))
        self.rank, self.corank, self.nsec, self.size = r.value, c.value, n.value, s.value
        leg = self.leg(0) if self.rank + self.corank else None
        self.sym = leg.sym if leg is not None else U1

    def serialize(self) -> bytes:
        """The tensor's storage archive (xt_tensor_serialize)."""
        n = C.c_size_t()
        check(lib().xt_tensor_serialize(self.h, None, 0, C.byref(n)), self.rt.h, "serialize")
        buf = (C.c_char * max(n.value, 1))()
        check(lib().xt_tensor_serialize(self.h, buf, n.value, C.byref(n)), self.rt.h, "serialize")
        return bytes(buf[:n.value])

    @classmethod
    def deserialize(cls, rt: Runtime, data: bytes) -> "Tensor":
        h = C.c_void_p()
        check(lib().xt_tensor_deserialize(rt.h, data, len(data), C.byref(h)), rt.h, "deserialize")
        return cls(rt, handle=h)

    def save(self, path: str):
        check(lib().xt_tensor_save(self.h, str(path).encode()), self.rt.h, "save")

    @classmethod
    def load(cls, rt: Runtime, path: str) -> "Tensor":
        h = C.c_void_p()
        check(lib().xt_tensor_load(rt.h, str(path).encode(), C.byref(h)), rt.h, "load")
        return cls(rt, handle=h)

    @classmethod
    def identity(cls, rt: Runtime, domain: Sequence[Qbasis], codomain: Sequence[Qbasis]) -> "Tensor":
        """Tensor::Identity (Tensor.hpp:224-236): every block the rows x cols identity (ones on the
        main diagonal), written on the host and uploaded."""
        t = cls(rt, domain, codomain)
        data = np.zeros(t.size)
        for _, off, r, c in t.sectors():
            k = min(r, c)
            data[off + np.arange(k) * (r + 1)] = 1.0
        return t.upload(data)

    def div_trace(self, s: "Tensor") -> "Tensor":
        """self * (1 / s.trace()) (degeneracy-weighted trace), recorded with both adjoints."""
        h = C.c_void_p()
        check(lib().xt_div_trace(self.h, s.h, C.byref(h)), self.rt.h, "div_trace")
        return self._wrap(h)

    def mexp(self, factor: float) -> "Tensor":
        """TensorBase::mexp(factor) (TensorBase.cpp:180-183): exp(factor * block) of every square
        block (order <= 64)."""
        h = C.c_void_p()
        check(lib().xt_tensor_mexp(self.h, float(factor), C.byref(h)), self.rt.h, "mexp")
        return self._wrap(h)

    def __del__(self):
        h = getattr(self, "h", None)
        f = _lib.native("xt_tensor_destroy") if (_lib is not None and h) else None
        if f is not None:
            f(h)
            self.h = None

    def _wrap(self, h):
        return Tensor(self.rt, handle=h)

    # ---- structure
    def leg(self, i: int) -> Qbasis:
        h = C.c_void_p()
        check(lib().xt_tensor_leg(self.h, i, C.byref(h)), self.rt.h, "leg")
        return Qbasis(_basis_sym(h), h)

    def uncoupledDomain(self):
        return [self.leg(i) for i in range(self.rank)]

    def uncoupledCodomain(self):
        return [self.leg(self.rank + i) for i in range(self.corank)]

    def coupledDomain(self) -> Qbasis:
        h = C.c_void_p()
        check(lib().xt_tensor_coupled(self.h, 0, C.byref(h)), self.rt.h, "coupledDomain")
        return Qbasis(self.sym, h)

    def coupledCodomain(self) -> Qbasis:
        h = C.c_void_p()
        check(lib().xt_tensor_coupled(self.h, 1, C.byref(h)), self.rt.h, "coupledCodomain")
        return Qbasis(self.sym, h)

    def plain_size(self) -> int:
        """Number of entries of the dense (symmetry-free) tensor: the product of the legs' full dims."""
        n = 1
        for i in range(self.rank + self.corank):
            n *= self.leg(i).fullDim()
        return n

    def sectors(self):
        """[(q, offset, rows, cols)] in storage order."""
        out = []
        q = (C.c_int * 2)()
        o, r, c = C.c_size_t(), C.c_size_t(), C.c_size_t()
        for i in range(self.nsec):
            lib().xt_tensor_sector(self.h, i, q, C.byref(o), C.byref(r), C.byref(c))
            out.append((tuple(q[: 2 if self.sym[1] else 1]), o.value, r.value, c.value))
        return out

    # ---- data
    def upload(self, data: np.ndarray):
        a = np.ascontiguousarray(data, dtype=np.float64)
        check(lib().xt_tensor_upload(self.h, a.ctypes.data_as(C.POINTER(C.c_double)), a.size), self.rt.h, "upload")
        return self

    def download(self) -> np.ndarray:
        a = np.empty(self.size, dtype=np.float64)
        check(lib().xt_tensor_download(self.h, a.ctypes.data_as(C.POINTER(C.c_double)), a.size), self.rt.h, "download")
        return a

    def setRandom(self, seed: int, dist: int = 0):
        check(lib().xt_tensor_set_random(self.h, seed & 0xFFFFFFFF, dist), self.rt.h, "setRandom")
        return self

    def device_ptr(self) -> int:
        return lib().xt_tensor_device_ptr(self.h) or 0

    # ---- operations
    def permute(self, shift: int, p: Sequence[int]) -> "Tensor":
        h = C.c_void_p()
        check(lib().xt_permute(self.h, shift, _ints(p), len(p), C.byref(h)), self.rt.h, "permute")
        return self._wrap(h)

    def permute_plan(self, shift: int, p: Sequence[int]):
        nt, ns, ne = C.c_int64(), C.c_int64(), C.c_int64()
        check(lib().xt_permute_plan_stats(self.h, shift, _ints(p), len(p), C.byref(nt), C.byref(ns), C.byref(ne)), self.rt.h, "plan")
        n_t, n_s = nt.value, ns.value
        td = np.zeros(max(n_t, 1), np.int64)
        tr = np.zeros(max(n_t, 1), np.int64)
        tc = np.zeros(max(n_t, 1), np.int64)
        tn = np.zeros(max(n_t, 1), np.int32)
        so = np.zeros(max(n_s, 1), np.int64)
        sc = np.zeros(max(n_s, 1), np.float64)
        I64 = C.POINTER(C.c_int64)
        check(lib().xt_permute_plan_dump(self.h, shift, _ints(p), len(p), td.ctypes.data_as(I64), tr.ctypes.data_as(I64), tc.ctypes.data_as(I64),
                                         tn.ctypes.data_as(C.POINTER(C.c_int32)), so.ctypes.data_as(I64), sc.ctypes.data_as(C.POINTER(C.c_double))),
              self.rt.h, "plan dump")
        return {"task_dst": td[:n_t], "task_rows": tr[:n_t], "task_cols": tc[:n_t], "task_nsrc": tn[:n_t], "src_off": so[:n_s], "src_coeff": sc[:n_s],
                "elems": ne.value}

    def __mul__(self, other):
        if isinstance(other, Tensor):
            return self.mul(other)
        return self.scale(float(other))

    __rmul__ = lambda self, other: self.scale(float(other))

    def mul(self, other: "Tensor", precision: str = "fp64") -> "Tensor":
        h = C.c_void_p()
        check(lib().xt_mul(self.h, other.h, PRECISION[precision], C.byref(h)), self.rt.h, "operator*")
        return self._wrap(h)

    def mul_into(self, other: "Tensor", out: "Tensor", precision: str = "fp64") -> "Tensor":
        """operator* written into `out` (which must have the product's structure); no allocation."""
        check(lib().xt_mul_into(self.h, other.h, out.h, PRECISION[precision]), self.rt.h, "operator* into")
        return out

    def contract(self, other: "Tensor", a1: Sequence[int], a2: Sequence[int], rankres: int, precision: str = "fp64") -> "Tensor":
        h = C.c_void_p()
        check(lib().xt_contract(self.h, _ints(a1), len(a1), other.h, _ints(a2), len(a2), rankres, PRECISION[precision], C.byref(h)), self.rt.h,
              "contract")
        return self._wrap(h)

    def tSVD(self, maxKeep: int, eps_svd: float, truncWeight: float = 0.0, entropy: float = 0.0, PRESERVE_MULTIPLETS: bool = True):
        U, S, V = C.c_void_p(), C.c_void_p(), C.c_void_p()
        tw, en = C.c_double(truncWeight), C.c_double(entropy)
        check(lib().xt_tsvd(self.h, int(maxKeep), float(eps_svd), int(bool(PRESERVE_MULTIPLETS)), C.byref(U), C.byref(S), C.byref(V), C.byref(tw),
                            C.byref(en)), self.rt.h, "tSVD")
        return self._wrap(U), self._wrap(S), self._wrap(V), tw.value, en.value

    def tSVD_spec(self, maxKeep: int, eps_svd: float, truncWeight: float = 0.0, entropy: float = 0.0, SVspec: Dict = None,
                  PRESERVE_MULTIPLETS: bool = True, RETURN_SPEC: bool = True):
        """The full tSVD overload (Tensor.hpp:353-362): returns (U, S, V, truncWeight, entropy, SVspec),
        SVspec a dict {q: kept singular values} extended as std::map::insert does."""
        L = lib()
        sp = C.c_void_p()
        check(L.xt_spec_create(C.byref(sp)), self.rt.h, "spec")
        try:
            U, S, V = C.c_void_p(), C.c_void_p(), C.c_void_p()
            tw, en = C.c_double(truncWeight), C.c_double(entropy)
            check(L.xt_tsvd_spec(self.h, int(maxKeep), float(eps_svd), C.byref(tw), C.byref(en), sp, int(bool(PRESERVE_MULTIPLETS)),
                                 int(bool(RETURN_SPEC)), C.byref(U), C.byref(S), C.byref(V)), self.rt.h, "tSVD")
            out = dict(SVspec or {})
            n = C.c_int()
            L.xt_spec_size(sp, C.byref(n))
            for i in range(n.value):
                q, cnt = (C.c_int * 2)(), C.c_int()
                L.xt_spec_entry(sp, i, q, C.byref(cnt), None, 0)
                vals = (C.c_double * max(cnt.value, 1))()
                L.xt_spec_entry(sp, i, q, C.byref(cnt), vals, cnt.value)
                key = (q[0],) if self.sym[1] == 0 else (q[0], q[1])
                out.setdefault(key, np.array(vals[:cnt.value]))
        finally:
            L.xt_spec_destroy(sp)
        return self._wrap(U), self._wrap(S), self._wrap(V), tw.value, en.value, out

    def tSVD_randomized(self, maxKeep: int, eps_svd: float, ranks: Dict, oversample: int = 8, rank_factor: float = 2.0, power_iters: int = 1,
                        sketch_precision: str = "fp64", seed: int = 1):
        """Randomized truncated block SVD; returns (U, S, Vdag, truncWeight, safe, residual_ratio)."""
        n = len(ranks)
        qs = (C.c_int * (2 * max(n, 1)))()
        rk = (C.c_int * max(n, 1))()
        for i, (q, r) in enumerate(ranks.items()):
            qq = (q,) if isinstance(q, int) else tuple(q)
            qs[2 * i] = qq[0]
            qs[2 * i + 1] = qq[1] if len(qq) > 1 else 0
            rk[i] = int(r)
        U, S, V = C.c_void_p(), C.c_void_p(), C.c_void_p()
        tw, safe, ratio = C.c_double(0.0), C.c_int(), C.c_double()
        check(lib().xt_tsvd_randomized(self.h, int(maxKeep), float(eps_svd), qs, rk, n, int(oversample), float(rank_factor), int(power_iters),
                                       PRECISION[sketch_precision], int(seed), C.byref(U), C.byref(S), C.byref(V), C.byref(tw), C.byref(safe),
                                       C.byref(ratio)), self.rt.h, "tSVD_randomized")
        return self._wrap(U), self._wrap(S), self._wrap(V), tw.value, bool(safe.value), ratio.value

    # ---- the halves of tSVD and the reverse-mode primitives (include/Xped/AD/ADTensor.hpp)
    def zeros_like(self) -> "Tensor":
        return self._unary(lib().xt_tensor_zeros_like)

    def permute_adj(self, shift: int, p: Sequence[int]) -> "Tensor":
        """Tensor::permute_adj (Tensor.cpp:258-333): the transpose of X -> X.permute(shift, p)."""
        h = C.c_void_p()
        check(lib().xt_permute_adj(self.h, shift, _ints(p), len(p), C.byref(h)), self.rt.h, "permute_adj")
        return self._wrap(h)

    def full_svd_impl(self):
        """Untruncated (U, Sigma, Vdag) per sector (Tensor.cpp:494-521)."""
        U, S, V = C.c_void_p(), C.c_void_p(), C.c_void_p()
        check(lib().xt_svd_full(self.h, C.byref(U), C.byref(S), C.byref(V)), self.rt.h, "full_svd_impl")
        return self._wrap(U), self._wrap(S), self._wrap(V)

    def cutoff_matrices(self, U, S, V, maxKeep: int, eps_svd: float, truncWeight: float = 0.0, entropy: float = 0.0, PRESERVE_MULTIPLETS: bool = True):
        """cutoff_matrices (Tensor.cpp:523-633) applied to the full factors of this tensor."""
        Ut, St, Vt = C.c_void_p(), C.c_void_p(), C.c_void_p()
        tw, en = C.c_double(truncWeight), C.c_double(entropy)
        check(lib().xt_cutoff(self.h, U.h, S.h, V.h, int(maxKeep), float(eps_svd), int(bool(PRESERVE_MULTIPLETS)), C.byref(Ut), C.byref(St), C.byref(Vt),
                              C.byref(tw), C.byref(en)), self.rt.h, "cutoff_matrices")
        return self._wrap(Ut), self._wrap(St), self._wrap(Vt), tw.value, en.value

    def flat_index(self, block: int, row: int, col: int) -> int:
        """Storage index of block(block)(row, col) (blocks column-major)."""
        q, off, r, c = self.sectors()[block]
        if not (0 <= row < r and 0 <= col < c):
            raise IndexError("coefficient outside the block")
        return off + row + col * r

    def coeff(self, block: int, row: int, col: int) -> float:
        return float(self.download()[self.flat_index(block, row, col)])

    def add_at(self, index: int, value: float):
        """In place: entry `index` of the storage += value (on the device)."""
        if not 0 <= index < self.size:
            raise IndexError("index outside the tensor")
        check(lib().xh_add_at(self.rt.ctx(), C.c_void_p(self.device_ptr()), int(index), float(value)), self.rt.h, "add_at")
        return self

    def tQR(self, RETURN_LQ: bool = False):
        """Tensor::tQR(RETURN_LQ) (Tensor.cpp:757-780)."""
        Q, R = C.c_void_p(), C.c_void_p()
        check(lib().xt_tqr(self.h, int(bool(RETURN_LQ)), C.byref(Q), C.byref(R)), self.rt.h, "tQR")
        return self._wrap(Q), self._wrap(R)

    def teigh(self):
        D, V = C.c_void_p(), C.c_void_p()
        check(lib().xt_teigh(self.h, C.byref(D), C.byref(V)), self.rt.h, "teigh")
        return self._wrap(D), self._wrap(V)

    def _unary(self, fn, *args):
        h = C.c_void_p()
        check(fn(self.h, *args, C.byref(h)), self.rt.h, fn.__name__)
        return self._wrap(h)

    def adjoint(self) -> "Tensor":
        return self._unary(lib().xt_adjoint)

    def twist(self, leg: int) -> "Tensor":
        return self._unary(lib().xt_twist, leg)

    def scale(self, alpha: float) -> "Tensor":
        return self._unary(lib().xt_scale, alpha)

    def axpby(self, other: "Tensor", a: float, b: float) -> "Tensor":
        h = C.c_void_p()
        check(lib().xt_axpby(self.h, other.h, a, b, C.byref(h)), self.rt.h, "axpby")
        return self._wrap(h)

    def __add__(self, other):
        return self.axpby(other, 1.0, 1.0)

    def __sub__(self, other):
        return other.axpby(self, -1.0, 1.0)

    def diag_sqrt(self):
        return self._unary(lib().xt_diag_map, 0)

    def diag_inv(self):
        return self._unary(lib().xt_diag_map, 1)

    def _scalar(self, fn):
        v = C.c_double()
        check(fn(self.h, C.byref(v)), self.rt.h, fn.__name__)
        return v.value

    def trace(self) -> float:
        return self._scalar(lib().xt_trace)

    def squaredNorm(self) -> float:
        return self._scalar(lib().xt_squared_norm)

    def norm(self) -> float:
        return float(np.sqrt(self.squaredNorm()))

    def maxNorm(self) -> float:
        return self._scalar(lib().xt_max_norm)


def qr_bwd(A: Tensor, Q: Tensor, R: Tensor, dQ: Tensor, dR: Tensor, dA: Tensor):
    """tQR reverse pass (xt_qr_bwd): dA += the QR adjoint, sector by sector."""
    check(lib().xt_qr_bwd(A.h, Q.h, R.h, dQ.h, dR.h, dA.h), A.rt.h, "qr adjoint")


def mul_bwd(A: Tensor, B: Tensor, dC: Tensor, dA: Tensor = None, dB: Tensor = None, precision: str = "fp64"):
    """operator* reverse pass (ADTensor.hpp:745-746), in place: dA += dC B^dagger, dB += A^dagger dC."""
    check(lib().xt_mul_bwd(A.h, B.h, dC.h, dA.h if dA is not None else None, dB.h if dB is not None else None, PRECISION[precision]), A.rt.h,
          "operator* adjoint")


def svd_bwd(U: Tensor, S: Tensor, V: Tensor, dU: Tensor, dS: Tensor, dV: Tensor, dA: Tensor):
    """full_svd_impl reverse pass (ADTensor.hpp:262-320), in place into dA."""
    check(lib().xt_svd_bwd(U.h, S.h, V.h, dU.h, dS.h, dV.h, dA.h), U.rt.h, "svd adjoint")


def cutoff_bwd(dUt: Tensor, dSt: Tensor, dVt: Tensor, dU: Tensor, dS: Tensor, dV: Tensor):
    """cutoff_matrices reverse pass (ADTensor.hpp:357-370), in place into dU, dS, dV."""
    check(lib().xt_cutoff_bwd(dUt.h, dSt.h, dVt.h, dU.h, dS.h, dV.h), dUt.rt.h, "cutoff adjoint")


def diag_bwd(y: Tensor, dy: Tensor, dx: Tensor, mode: str):
    """diag_sqrt / diag_inv reverse passes (ADTensor.hpp:560-590), in place into dx."""
    check(lib().xt_diag_bwd(y.h, dy.h, dx.h, {"sqrt": 0, "inv": 1}[mode]), y.rt.h, "diag adjoint")


def trace_bwd(dx: Tensor, g: float):
    """trace reverse pass (ADTensor.hpp:533-545): dx += g * WeightedIdentity."""
    check(lib().xt_trace_bwd(dx.h, float(g)), dx.rt.h, "trace adjoint")


def max_norm_bwd(x: Tensor, dx: Tensor, g: float):
    """maxNorm reverse pass (ADTensor.hpp:450-476)."""
    check(lib().xt_max_norm_bwd(x.h, dx.h, float(g)), x.rt.h, "maxNorm adjoint")


def get_permutations(a1, rank1, a2, rank2, rankres):
    """util::constFct::get_permutations through the C ABI."""
    n1, n2 = len(a1), len(a2)
    p1, p2 = (C.c_int * n1)(), (C.c_int * n2)()
    pres = (C.c_int * (n1 + n2))()
    s1, s2, sr = C.c_int(), C.c_int(), C.c_int()
    check(lib().xt_get_permutations(_ints(a1), n1, rank1, _ints(a2), n2, rank2, rankres, p1, C.byref(s1), p2, C.byref(s2), pres, C.byref(sr)),
          None, "get_permutations")
    nfree = sum(1 for x in a1 if x < 0) + sum(1 for x in a2 if x < 0)
    return list(p1), s1.value, list(p2), s2.value, list(pres)[:nfree], sr.value


def truncation_rule(sym, allSV, maxKeep, eps_svd, preserve_multiplets, truncWeight=0.0, entropy=0.0):
    n = len(allSV)
    qs = (C.c_int * (2 * max(n, 1)))()
    sv = (C.c_double * max(n, 1))()
    for i, (q, s) in enumerate(allSV):
        q = tuple(q) + (0,) * (2 - len(tuple(q)))
        qs[2 * i], qs[2 * i + 1] = q[0], q[1]
        sv[i] = s
    nsec = C.c_int()
    kq = (C.c_int * (2 * max(n, 1)))()
    kn = (C.c_int * max(n, 1))()
    tw, en = C.c_double(truncWeight), C.c_double(entropy)
    check(lib().xt_truncation_rule(sym[0], sym[1], qs, sv, n, maxKeep, eps_svd, int(preserve_multiplets), C.byref(nsec), kq, kn, C.byref(tw),
                                   C.byref(en)), None, "truncation_rule")
    nq = 2 if sym[1] else 1
    kept = {tuple(kq[2 * i: 2 * i + nq]): kn[i] for i in range(nsec.value)}
    return kept, tw.value, en.value


class CTM:
    """CTMRG environment of a 2x2-cell iPEPS (the reference's CTM<double, Symmetry, 2>)."""

    def __init__(self, rt: Runtime, A: Sequence[Tensor], chi: int):
        self.rt = rt
        self._A = list(A)  # keep the site tensors alive
        arr = (C.c_void_p * 4)(*[a.h.value if isinstance(a.h, C.c_void_p) else a.h for a in A])
        h = C.c_void_p()
        check(lib().xt_ctm_create(rt.h, arr, int(chi), C.byref(h)), rt.h, "CTM()")
        self.h = h

    def __del__(self):
        h = getattr(self, "h", None)
        f = _lib.native("xt_ctm_destroy") if (_lib is not None and h) else None
        if f is not None:
            f(h)
            self.h = None

    def set_options(self, precision="fp64", svd="full", oversample=8, power_iters=1, record_spectra=False, check_finite=False, verbose=False,
                    fresh_predictor_sketch=False, batch_contractions=False):
        flags = (int(bool(record_spectra)) | (2 if check_finite else 0) | (4 if verbose else 0) | (8 if fresh_predictor_sketch else 0)
                 | (16 if batch_contractions else 0))
        check(lib().xt_ctm_set_options(self.h, PRECISION[precision], 1 if svd == "randomized" else 0, oversample, power_iters, flags),
              self.rt.h, "set_options")
        return self

    def set_rsvd(self, oversample=8, rank_factor=2.0, power_iters=1, sketch_precision="fp64"):
        check(lib().xt_ctm_set_rsvd(self.h, int(oversample), float(rank_factor), int(power_iters), PRECISION[sketch_precision]), self.rt.h, "set_rsvd")
        return self

    def set_predictor(self, params=None, width=64, margin=1.25):
        """Learned target ranks for the randomized SVDs (see xped_amd.rank_predictor); None removes them."""
        if params is None:
            check(lib().xt_ctm_set_predictor(self.h, None, 0, int(width), float(margin)), self.rt.h, "set_predictor")
            return self
        p = np.ascontiguousarray(params, np.float32)
        check(lib().xt_ctm_set_predictor(self.h, p.ctypes.data_as(C.POINTER(C.c_float)), p.size, int(width), float(margin)), self.rt.h, "set_predictor")
        return self

    def predictor_stats(self):
        """(decompositions whose learned ranks read the previous spectrum, ones that sketched afresh)"""
        a, b = C.c_long(), C.c_long()
        lib().xt_ctm_predictor_stats(self.h, C.byref(a), C.byref(b))
        return a.value, b.value

    def rsvd_stats(self):
        """(randomized SVDs run, full-SVD fallbacks, largest missed-weight ratio)"""
        a, b, r = C.c_long(), C.c_long(), C.c_double()
        lib().xt_ctm_rsvd_stats(self.h, C.byref(a), C.byref(b), C.byref(r))
        return a.value, b.value, r.value

    def init(self):
        check(lib().xt_ctm_init(self.h), self.rt.h, "CTM::init")
        return self

    def grow_all(self):
        check(lib().xt_ctm_grow_all(self.h), self.rt.h, "grow_all")

    def computeRDM(self):
        check(lib().xt_ctm_compute_rdm(self.h), self.rt.h, "computeRDM")

    def energy(self, Jxy=1.0, Jz=1.0) -> float:
        e = C.c_double()
        check(lib().xt_ctm_energy(self.h, Jxy, Jz, C.byref(e)), self.rt.h, "energy")
        return e.value

    def step(self) -> float:
        self.grow_all()
        self.computeRDM()
        return self.energy()

    def tensor(self, which: str, site: int) -> Tensor:
        names = ["C1", "C2", "C3", "C4", "T1", "T2", "T3", "T4", "rho_h", "rho_v"]
        h = C.c_void_p()
        check(lib().xt_ctm_tensor(self.h, names.index(which), site, C.byref(h)), self.rt.h, "ctm tensor")
        return Tensor(self.rt, handle=h)

    def copy_env(self, other: "CTM"):
        """Take other's environment tensors (shared, not copied)."""
        check(lib().xt_ctm_copy_env(self.h, other.h), self.rt.h, "copy_env")
        return self

    def energy_grad(self, steps: int = 1, Jxy: float = 1.0, Jz: float = 1.0):
        """Energy after `steps` more grow_all moves and its gradient with respect to the four site
        tensors (CTMSolver::solve with AD, each move a checkpoint).  Returns (E, [dA per site])."""
        E = C.c_double()
        out = (C.c_void_p * 4)()
        check(lib().xt_ctm_energy_grad(self.h, int(steps), float(Jxy), float(Jz), C.byref(E), out), self.rt.h, "energy_grad")
        return E.value, [Tensor(self.rt, handle=C.c_void_p(out[i])) for i in range(4)]

    def svd_ms(self) -> float:
        v = C.c_double()
        lib().xt_ctm_svd_ms(self.h, C.byref(v))
        return v.value

    def spectra(self):
        n = C.c_int()
        lib().xt_ctm_num_spectra(self.h, C.byref(n))
        out = []
        for r in range(n.value):
            cap = 1 << 20
            site, d, nsec = C.c_int(), C.c_int(), C.c_int()
            qs = (C.c_int * 4096)()
            rows, cols, kept, nsv = (C.c_int * 2048)(), (C.c_int * 2048)(), (C.c_int * 2048)(), (C.c_int * 2048)()
            sv = (C.c_double * cap)()
            check(lib().xt_ctm_spectrum(self.h, r, C.byref(site), C.byref(d), C.byref(nsec), qs, rows, cols, kept, nsv, sv, cap), self.rt.h, "spectrum")
            secs, off = [], 0
            for k in range(nsec.value):
                secs.append({"q": qs[2 * k], "rows": rows[k], "cols": cols[k], "kept": kept[k], "sv": np.array(sv[off:off + nsv[k]])})
                off += nsv[k]
            out.append({"site": site.value, "dir": d.value, "sectors": secs})
        return out

    def clear_spectra(self):
        lib().xt_ctm_clear_spectra(self.h)

    def set_ranks(self, site, direction, ranks: Dict):
        n = len(ranks)
        qs = (C.c_int * (2 * max(n, 1)))()
        rk = (C.c_int * max(n, 1))()
        for i, (q, r) in enumerate(ranks.items()):
            qs[2 * i] = q if isinstance(q, int) else q[0]
            rk[i] = int(r)
        check(lib().xt_ctm_set_ranks(self.h, site, direction, qs, rk, n), self.rt.h, "set_ranks")


def _handles(ts: Sequence[Tensor]):
    return (C.c_void_p * len(ts))(*[t.h.value if isinstance(t.h, C.c_void_p) else t.h for t in ts])


class WeightedPEPS:
    """The weighted iPEPS state of the simple update: bare sites Gs, bond weights whs/wvs and the
    bond dimension D (iPEPS::initWeightTensors, iPEPS.cpp:575-591).  Site index x * 2 + y."""

    def __init__(self, rt: Runtime, A: Sequence[Tensor], D: int):
        self.rt = rt
        h = C.c_void_p()
        check(lib().xt_peps_create(rt.h, _handles(A), int(D), C.byref(h)), rt.h, "iPEPS weights")
        self.h = h

    def __del__(self):
        h = getattr(self, "h", None)
        f = _lib.native("xt_peps_destroy") if (_lib is not None and h) else None
        if f is not None:
            f(h)
            self.h = None

    def _get(self, which, site):
        h = C.c_void_p()
        check(lib().xt_peps_tensor(self.h, which, int(site), C.byref(h)), self.rt.h, "iPEPS tensor")
        return Tensor(self.rt, handle=h)

    def G(self, site) -> Tensor:
        return self._get(0, site)

    def wh(self, site) -> Tensor:
        return self._get(1, site)

    def wv(self, site) -> Tensor:
        return self._get(2, site)

    def A(self, site) -> Tensor:
        """applyWeights(Gs, sqrt weights) (iPEPS::updateAtensors, iPEPS.cpp:593-602)."""
        return self._get(3, site)


class TimePropagator:
    """Simple-update propagator exp(-dt H) for nearest-neighbour bond Hamiltonians on a 2x2 cell
    (TimePropagator.cpp: initU :453-487, t_step :78-105, t_step_h :113-190, t_step_v :192-226).
    Hh[x*2+y] acts on the bond (x,y)-(x+1,y), Hv[x*2+y] on (x,y)-(x,y+1)."""

    def __init__(self, rt: Runtime, Hh: Sequence[Tensor], Hv: Sequence[Tensor], dt: float):
        self.rt = rt
        h = C.c_void_p()
        check(lib().xt_tprop_create(rt.h, _handles(Hh), _handles(Hv), float(dt), C.byref(h)), rt.h, "TimePropagator")
        self.h = h

    def __del__(self):
        h = getattr(self, "h", None)
        f = _lib.native("xt_tprop_destroy") if (_lib is not None and h) else None
        if f is not None:
            f(h)
            self.h = None

    def t_step(self, psi: WeightedPEPS):
        check(lib().xt_tprop_step(self.h, psi.h), self.rt.h, "t_step")

    def t_step_h(self, psi: WeightedPEPS, x: int, y: int):
        check(lib().xt_tprop_bond(self.h, psi.h, 0, int(x), int(y)), self.rt.h, "t_step_h")

    def t_step_v(self, psi: WeightedPEPS, x: int, y: int):
        check(lib().xt_tprop_bond(self.h, psi.h, 1, int(x), int(y)), self.rt.h, "t_step_v")

    def _get(self, which, site):
        h = C.c_void_p()
        check(lib().xt_tprop_tensor(self.h, which, int(site), C.byref(h)), self.rt.h, "propagator tensor")
        return Tensor(self.rt, handle=h)

    def gate_h(self, site) -> Tensor:
        return self._get(0, site)

    def gate_v(self, site) -> Tensor:
        return self._get(1, site)

    def spectrum_h(self, site) -> Tensor:
        return self._get(2, site)

    def spectrum_v(self, site) -> Tensor:
        return self._get(3, site)


def heisenberg_bond(rt: Runtime, phys: Qbasis, Jxy: float = 1.0, Jz: float = 1.0) -> Tensor:
    """The spin-1/2 Heisenberg bond Hamiltonian Jz Sz.Sz + Jxy/2 (S+S- + S-S+) (Heisenberg.hpp:50-53)
    as Tensor<2,2>(p, p | p, p) on a U(1) physical leg whose labels are 2 Sz.  Entries are written
    on the host from the tree labels and uploaded (the model's input, like the reference's
    operator tables)."""
    H = Tensor(rt, [phys, phys], [phys, phys])
    data = np.zeros(H.size)
    cd, cc = H.coupledDomain(), H.coupledCodomain()
    dq = {e[0]: i for i, e in enumerate(cd.entries())}
    cq = {e[0]: i for i, e in enumerate(cc.entries())}

    def sz(tree, leg):
        lab = tree["unc"][leg][0]
        return 0.5 * (-lab if tree["dual"][leg] else lab)

    for q, off, r, c in H.sectors():
        dom, cod = cd.trees(dq[q]), cc.trees(cq[q])
        r0 = 0
        for td in dom:
            c0 = 0
            for tc in cod:
                if td["dim"] == 1 and tc["dim"] == 1:
                    c1, c2, a1, a2 = sz(td, 0), sz(td, 1), sz(tc, 0), sz(tc, 1)
                    v = 0.0
                    if c1 == a1 and c2 == a2:
                        v = Jz * a1 * a2
                    if (c1 == a1 + 1 and c2 == a2 - 1) or (c1 == a1 - 1 and c2 == a2 + 1):
                        v = 0.5 * Jxy
                    data[off + r0 + c0 * r] = v
                c0 += tc["dim"]
            r0 += td["dim"]
    return H.upload(data)


class Tape:
    """Reverse-mode recording of the native tensor operations on one runtime (xt_tape_*): the
    stan-math reverse-pass stack of the reference's AD tensors (ADTensor.hpp)."""

    def __init__(self, rt: Runtime):
        self.rt = rt
        check(lib().xt_tape_begin(rt.h), rt.h, "tape")
        self.open = True

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def pause(self, paused: bool = True):
        check(lib().xt_tape_pause(self.rt.h, int(bool(paused))), self.rt.h, "tape pause")

    def seed(self, y: Tensor, dy: Tensor, alpha: float = 1.0):
        check(lib().xt_tape_seed(self.rt.h, y.h, dy.h, float(alpha)), self.rt.h, "tape seed")

    def seed_trace(self, y: Tensor, g: float = 1.0):
        """Seed the adjoint of y.trace() (times g)."""
        check(lib().xt_tape_seed_trace(self.rt.h, y.h, float(g)), self.rt.h, "tape seed")

    def backward(self):
        check(lib().xt_tape_backward(self.rt.h), self.rt.h, "tape backward")

    def adjoint(self, x: Tensor) -> Tensor:
        h = C.c_void_p()
        check(lib().xt_tape_adjoint(self.rt.h, x.h, C.byref(h)), self.rt.h, "tape adjoint")
        return Tensor(self.rt, handle=h)

    def close(self):
        if self.open:
            lib().xt_tape_end(self.rt.h)
            self.open = False
