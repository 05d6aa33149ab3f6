"""Learned per-sector rank prediction for the randomized block SVD of the CTMRG projectors.

The network (device layer xh_mlp_*, hand-written HIP kernels; 19 -> 64 -> 64 -> 1, FP32) reads
the RANK_TOP = 16 leading singular values of a sector (from a cheap sketch), the sector size and
a candidate cut c, and predicts log(1 + #{sigma >= c}).  Bisection on c makes the predicted
counts sum to chi; the counts (times a margin) are the target ranks of the randomized SVD, whose
safety check falls back to the full FP64 SVD when a sketch was too narrow.  The features are
computed by the library itself (xt_rank_features), so training and the C++ inference inside the
CTM step see the same numbers.

Training data: the exact FP64 block spectra of full-SVD CTMRG steps (records of site, direction,
per-sector spectrum and kept count), with cuts sampled around each decomposition's true cut.
No counterpart in the reference, whose tSVD is always the full one.
"""
from __future__ import annotations

import ctypes as C
import math
from pathlib import Path
from typing import Dict, List, Sequence

import numpy as np

from . import _lib
from ._lib import check, lib

NFEAT = 19
TOP = 16
WIDTH = 64
DATA = Path(__file__).resolve().parent / "data"
SPECTRA = DATA / "ctm_spectra_d8_chi256.npz"
WEIGHTS = DATA / "ranknet_d8_chi256.npz"


def features(top: Sequence[float], dim: int, s_glob: float, log_cut: float) -> np.ndarray:
    t = np.ascontiguousarray(np.asarray(top, dtype=np.float64)[:TOP])
    out = np.zeros(NFEAT, np.float32)
    rc = lib().xt_rank_features(t.ctypes.data_as(C.POINTER(C.c_double)), int(t.size), int(dim), float(s_glob), float(log_cut),
                                out.ctypes.data_as(C.POINTER(C.c_float)))
    if rc != 0:
        raise ValueError("xt_rank_features failed")
    return out


def records_from_npz(path) -> List[dict]:
    """Decomposition records [{sv: [per-sector arrays], kept, rows, cols, q}] from a spectra dump."""
    d = np.load(path, allow_pickle=False)
    keys = sorted(set(zip(d["step"].tolist(), d["site"].tolist(), d["dir"].tolist())))
    if "len" in d:  # compact layout: leading values per sector in a padded 2-D array
        sv_of = lambda i: d["sv"][i, :d["len"][i]].astype(np.float64)
    else:
        sv_of = lambda i: d["sv"][d["off"][i]:d["off"][i + 1]]
    out = []
    for k in keys:
        idx = np.where((d["step"] == k[0]) & (d["site"] == k[1]) & (d["dir"] == k[2]))[0]
        out.append({"key": k, "q": d["q"][idx], "rows": d["rows"][idx], "cols": d["cols"][idx], "kept": d["kept"][idx],
                    "sv": [sv_of(i) for i in idx]})
    return out


def compact(path_in, path_out, keep=192):
    """Leading `keep` singular values per sector (float32) of a spectra dump, for fixtures."""
    d = np.load(path_in, allow_pickle=False)
    n = d["step"].size
    sv = np.zeros((n, keep), np.float32)
    ln = np.zeros(n, np.int32)
    for i in range(n):
        v = d["sv"][d["off"][i]:d["off"][i + 1]][:keep]
        sv[i, :v.size] = v
        ln[i] = v.size
    np.savez_compressed(path_out, step=d["step"], site=d["site"], dir=d["dir"], q=d["q"], rows=d["rows"], cols=d["cols"], kept=d["kept"], len=ln, sv=sv)


def record_cut(rec) -> float:
    vals = [sv[k - 1] for sv, k in zip(rec["sv"], rec["kept"]) if k > 0]
    return float(min(vals))


def dataset(records, cuts_per_sector=16, spread=1.0, seed=0):
    """(X, T): features with cuts sampled uniformly within `spread` decades of each record's true
    cut, targets log(1 + exact count above the cut)."""
    rng = np.random.default_rng(seed)
    X, T = [], []
    for rec in records:
        tops = [sv[:TOP] for sv in rec["sv"]]
        s_glob = max(float(t[0]) for t in tops if t.size)
        lc0 = math.log10(record_cut(rec) / s_glob)
        for sv, r, c in zip(rec["sv"], rec["rows"], rec["cols"]):
            if sv.size == 0:
                continue
            for u in rng.uniform(-spread, spread, cuts_per_sector):
                lc = lc0 + u
                X.append(features(sv[:TOP], min(int(r), int(c)), s_glob, lc))
                T.append(math.log1p(float(np.count_nonzero(sv >= s_glob * 10.0 ** lc))))
    return np.asarray(X, np.float32), np.asarray(T, np.float32)


class DeviceArray:
    """A device FP32 array on a runtime's context (the runtime is kept alive with it)."""

    def __init__(self, rt, a: np.ndarray):
        self.rt = rt
        ctx = rt.ctx()
        self.ctx = ctx
        self.n = a.size
        self.p = C.c_void_p()
        check(lib().xh_alloc(ctx, max(4 * a.size, 4), C.byref(self.p)), None, "alloc")
        if a.size:
            lib().xh_upload(ctx, self.p, np.ascontiguousarray(a, np.float32).ctypes.data, 4 * a.size)

    def get(self) -> np.ndarray:
        out = np.empty(self.n, np.float32)
        lib().xh_download(self.ctx, out.ctypes.data, self.p, 4 * self.n)
        return out

    def __del__(self):
        f = _lib.native("xh_free") if (_lib is not None and self.p) else None
        if f is not None:
            f(self.ctx, self.p)
            self.p = None


class RankNet:
    """The device network (xh_mlp) with training and inference helpers."""

    def __init__(self, rt, width: int = WIDTH, seed: int = 1, params: np.ndarray = None):
        self.rt = rt
        self.ctx = rt.ctx()
        self.width = width
        self.h = C.c_void_p()
        check(lib().xh_mlp_create(self.ctx, NFEAT, width, seed, C.byref(self.h)), None, "mlp create")
        n = C.c_int64()
        lib().xh_mlp_num_params(self.h, C.byref(n))
        self.nparams = n.value
        if params is not None:
            self.set_params(params)

    def __del__(self):
        f = _lib.native("xh_mlp_destroy") if (_lib is not None and getattr(self, "h", None)) else None
        if f is not None:
            f(self.ctx, self.h)
            self.h = None

    def params(self) -> np.ndarray:
        out = np.empty(self.nparams, np.float32)
        check(lib().xh_mlp_get_params(self.ctx, self.h, out.ctypes.data_as(C.POINTER(C.c_float))), None, "mlp params")
        return out

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, np.float32)
        if p.size != self.nparams:
            raise ValueError("parameter count mismatch")
        check(lib().xh_mlp_set_params(self.ctx, self.h, p.ctypes.data_as(C.POINTER(C.c_float))), None, "mlp set params")

    def forward(self, X: np.ndarray) -> np.ndarray:
        dX = DeviceArray(self.rt, np.ascontiguousarray(X, np.float32).ravel())
        dY = DeviceArray(self.rt, np.zeros(X.shape[0], np.float32))
        check(lib().xh_mlp_forward(self.ctx, self.h, dX.p, X.shape[0], dY.p), None, "mlp forward")
        return dY.get()

    def grad(self, X: np.ndarray, T: np.ndarray):
        dX = DeviceArray(self.rt, np.ascontiguousarray(X, np.float32).ravel())
        dT = DeviceArray(self.rt, np.ascontiguousarray(T, np.float32))
        dG = DeviceArray(self.rt, np.zeros(self.nparams, np.float32))
        loss = C.c_float()
        check(lib().xh_mlp_grad(self.ctx, self.h, dX.p, dT.p, X.shape[0], dG.p, C.byref(loss)), None, "mlp grad")
        return loss.value, dG.get()

    def train(self, X: np.ndarray, T: np.ndarray, steps: int = 2000, lr: float = 3e-3, log_every: int = 0) -> List[float]:
        """Full-batch Adam on the GPU; returns the loss after every step."""
        dX = DeviceArray(self.rt, np.ascontiguousarray(X, np.float32).ravel())
        dT = DeviceArray(self.rt, np.ascontiguousarray(T, np.float32))
        dG = DeviceArray(self.rt, np.zeros(self.nparams, np.float32))
        losses = []
        loss = C.c_float()
        for t in range(1, steps + 1):
            check(lib().xh_mlp_grad(self.ctx, self.h, dX.p, dT.p, X.shape[0], dG.p, C.byref(loss)), None, "mlp grad")
            check(lib().xh_mlp_adam_step(self.ctx, self.h, dG.p, float(lr), t), None, "adam")
            losses.append(loss.value)
            if log_every and t % log_every == 0:
                print(f"[train] step {t} loss {loss.value:.5f}", flush=True)
        return losses

    # ------------------------------------------------------------------ inference on records
    def counts(self, tops, dims, s_glob, log_cut) -> np.ndarray:
        X = np.stack([features(t, d, s_glob, log_cut) for t, d in zip(tops, dims)])
        y = self.forward(X).astype(np.float64)
        return np.minimum(np.asarray(dims, np.float64), np.maximum(0.0, np.expm1(y)))

    def ranks(self, tops, dims, chi: int, margin: float = 1.25) -> np.ndarray:
        """The CTM's inference (RankPredictor::predict): the cut where the predicted counts sum to
        chi, on a coarse grid over 14 decades and a fine grid inside the crossing cell."""
        s_glob = max(float(t[0]) for t in tops if len(t))
        n = len(tops)

        def crossing(lo, hi, m):
            grid = lo + (hi - lo) * np.arange(m + 1) / m
            X = np.stack([features(t, d, s_glob, g) for g in grid for t, d in zip(tops, dims)])
            c = np.minimum(np.tile(np.asarray(dims, np.float64), m + 1), np.maximum(0.0, np.expm1(self.forward(X).astype(np.float64))))
            tot = c.reshape(m + 1, n).sum(axis=1)
            above = np.nonzero(tot > chi)[0]
            if above.size == 0:
                return lo, lo
            k = int(above[-1])
            return grid[k], (hi if k >= m else grid[k + 1])

        lo, hi = crossing(-14.0, 0.0, 280)
        lo, hi = crossing(lo, hi, 64)
        return np.ceil(margin * self.counts(tops, dims, s_glob, hi)).astype(int)


def sketch_cost_and_safety(rec, ranks, oversample=16, rank_factor=2.0, safety=2.0):
    """For target ranks on a record with exact spectra: the range-finder panel widths
    l_q = min(ceil(rank_factor k_q) + oversample, dim_q), their cost sum_q l_q (m_q + n_q) (the
    entries of the sketch panels, a proxy for the range finder's memory traffic) and whether the
    randomized SVD's safety check would accept every sector (safety * sigma_{l_q} < sigma_cut)."""
    cut = record_cut(rec)
    cost, safe = 0.0, True
    for sv, r, c, k in zip(rec["sv"], rec["rows"], rec["cols"], ranks):
        dim = min(int(r), int(c))
        if dim == 0:
            continue
        l = int(min(math.ceil(rank_factor * max(0, int(k))) + oversample, dim))
        cost += l * (int(r) + int(c))
        if l < dim and not (safety * sv[l - 1] < cut):
            safe = False
    return cost, safe


def proportional_ranks(rec, chi):
    """The rank guess without a predictor: chi shared out in proportion to the sector sizes."""
    dims = np.array([min(int(r), int(c)) for r, c in zip(rec["rows"], rec["cols"])], np.float64)
    return np.ceil(chi * dims / dims.sum()).astype(int)


def train_default(rt, spectra=SPECTRA, out=WEIGHTS, steps=3000, lr=3e-3, seed=3, train_steps=(0,), log_every=500):
    """Train the network on the GPU on the records of the given CTM steps and save the weights
    (npz: params, width, and the training/held-out losses)."""
    recs = records_from_npz(spectra)
    train = [r for r in recs if r["key"][0] in train_steps]
    held = [r for r in recs if r["key"][0] not in train_steps]
    X, T = dataset(train, 16)
    net = RankNet(rt, seed=seed)
    losses = net.train(X, T, steps=steps, lr=lr, log_every=log_every)
    held_mse = None
    if held:
        Xh, Th = dataset(held, 16, seed=1)
        held_mse = float(np.mean((net.forward(Xh) - Th) ** 2))
    np.savez(out, params=net.params(), width=np.int32(net.width), train_loss=np.float32(losses[-1]),
             held_out_mse=np.float32(held_mse if held_mse is not None else np.nan))
    return net, losses, held_mse


def load_weights(path=WEIGHTS):
    d = np.load(path, allow_pickle=False)
    return d["params"].astype(np.float32), int(d["width"])


if __name__ == "__main__":
    import sys

    sys.path.insert(0, str(Path(__file__).resolve().parents[1]))
    from xped_amd.tensor import Runtime

    net, losses, held = train_default(Runtime(0))
    print(f"trained: final loss {losses[-1]:.5f}, held-out mse {held}")