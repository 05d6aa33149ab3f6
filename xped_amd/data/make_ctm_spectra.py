"""Writes xped_amd/data/ctm_spectra_d8_chi256.npz: the exact FP64 block spectra of the projector
SVDs of two full-SVD CTMRG steps (after the growing step) of the D = 8, chi = 256 U(1) iPEPS
workload (site tensors N(0,1) * 10 / sqrt(plain size), seed 1), 192 leading values per sector in
float32, with the sector sizes and kept counts.  These spectra train and judge the learned rank
predictor.  Needs the HIP library and a GPU:  python xped_amd/data/make_ctm_spectra.py
"""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import xped_amd.tensor as xt  # noqa: E402
from xped_amd import configs, rank_predictor  # noqa: E402
from xped_amd.ipeps import random_ipeps  # noqa: E402


def main(out=ROOT / "xped_amd" / "data" / "ctm_spectra_d8_chi256.npz", chi=256, seed=1, steps=2):
    rt = xt.Runtime(0)
    A = [a.scale(10.0) for a in random_ipeps(rt, configs.c4_virtual(), configs.c4_physical(), seed)]
    c = xt.CTM(rt, A, chi).init()
    c.grow_all()
    c.set_options(record_spectra=True)
    cols = {k: [] for k in ("step", "site", "dir", "q", "rows", "cols", "kept")}
    sv, off = [], [0]
    for s in range(steps):
        c.grow_all()
        for r in c.spectra():
            for sec in r["sectors"]:
                for k, v in (("step", s), ("site", r["site"]), ("dir", r["dir"]), ("q", sec["q"]), ("rows", sec["rows"]), ("cols", sec["cols"]),
                             ("kept", sec["kept"])):
                    cols[k].append(v)
                sv.append(sec["sv"])
                off.append(off[-1] + sec["sv"].size)
        c.clear_spectra()
        print(f"step {s} recorded", flush=True)
    tmp = Path(str(out) + ".full.npz")
    np.savez(tmp, **{k: np.array(v) for k, v in cols.items()}, off=np.array(off), sv=np.concatenate(sv))
    rank_predictor.compact(tmp, out)
    tmp.unlink()


if __name__ == "__main__":
    main()
