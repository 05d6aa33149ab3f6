"""Benchmark of the block-sparse symmetric tensor hot path on one MI355X.

Default workload (--workload c4): one CTMRG step (grow_all + computeRDM + energy, CORNER
projection) of a U(1) iPEPS on a 2x2 unit cell, virtual D = 8 over charges -3..3
({+-3:1, +-2:1, +-1:1, 0:2}), physical spin 1/2, environment chi = 256, nearest-neighbour
Heisenberg energy.  Two modes run on fresh runtimes (no shared plan caches):
  fp64 - plain FP64 GPU path: FP64 grouped GEMMs, full FP64 block SVDs (one-sided block Jacobi)
  fast - the accelerated path: FP64 grouped GEMMs, randomized block SVDs of the projector
         products with per-sector target ranks from the learned predictor (or, with
         --ranks previous, the previous step's kept counts); full FP64 SVD as fallback
Both start from the same warm-up (init from the double layer + one full-SVD grow_all reaching
chi = 256).  `value` is the contraction GEMM throughput of the fast step: the summed algorithmic
flops of its contraction products (operator* of every contract) over their HIP-event-timed kernel
time; `roofline` reports the dominant kernel, the contraction launches with matrix-core tiles,
and `skinny_contractions` the memory-bound ones (bytes against HBM), the sketch and Gram products
of the randomized SVDs are reported beside them; `ms_per_step` is the fast step's wall time; the
fp64 step time, the speedup and the energy difference are reported beside them.

--workload c3 times the grouped U(1)xU(1) fermionic contraction A(l1,l2|l3,l4).B(l3,l4|l5,l6)
(leg dim 192, 313 matched sector GEMMs) instead.

--workload copy is the HBM calibration for counter-based traffic figures: k_axpby with b = 0
copies 1 GiB (reads 1 GiB, writes 1 GiB) three times; it prints the copy bandwidth.

`roofline.traffic` (HBM bytes per launch of the dominant kernel) is read from the rocprofv3
counter summary in profiles/ when one is present for the workload (FETCH_SIZE + WRITE_SIZE,
scaled by the copy calibration), else null.  `permute_roofline` reports the permute kernels of
the same step against the HBM peak: algorithmic bytes 8 (elements read + written) over their
HIP-event kernel time, with their counter-based traffic per launch from profiles/ likewise.

Prints one JSON line.  --gpus N > 1 is refused: this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "contraction TFLOP/s + CTMRG step ms, mixed/learned vs FP64 GPU at ≤1e-10 err"
PEAKS = {"fp64": 78.6, "fp32": 157.3, "bf16x3": 2500.0 / 3.0, "ozaki": 5000.0}  # TFLOP/s, dense; BF16x3 counted as 3 BF16 products
# XH_OZAKI: 8 slices of 7 bits per row / column, the 36 slice products with s + t <= 9 on the int8
# matrix cores (TOP/s against 5.0 POP/s dense int8)
OZ_PAIRS = 36
DTYPES = {"fp64": "f64", "fp32": "f32", "bf16x3": "bf16x3/f32", "ozaki": "f64 via int8 slices (Ozaki)"}
SCALE = 10.0  # site tensors N(0,1) * SCALE / sqrt(plain size), see run_c4
# the CPU oracle's whole FP64 CTMRG step on this workload (oracle/ctm.py: numpy/OpenBLAS per block,
# written by tests/golden/make_ctm_fixture.py): far beyond a bounded sample, so it is quoted from that
# run rather than measured here
ORACLE_STEP = {"seconds": 4171.0, "cores": 8, "where": "build container, tests/golden/make_ctm_fixture.py, second step (chi = 256 reached)"}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default="c4", choices=["c4", "c3", "copy"])
    ap.add_argument("--steps", type=int, default=3, help="timed fast-mode steps (c4) or contractions (c3)")
    ap.add_argument("--warmup", type=int, default=1, help="untimed fast-mode steps after the common warm-up (c4)")
    ap.add_argument("--fp64-steps", type=int, default=4, help="steps of the plain FP64 mode (c4), compared step by step with the fast mode's")
    ap.add_argument("--chi", type=int, default=256)
    ap.add_argument("--graph", action="store_true", help="c3: also time the product replayed from a captured HIP graph")
    ap.add_argument("--autotune", action="store_true", help="pick each FP64 GEMM plan's tiling by timing both on its first call")
    ap.add_argument("--power-iters", type=int, default=6)
    ap.add_argument("--oversample", type=int, default=16)
    ap.add_argument("--rank-factor", type=float, default=2.0, help="sketch width l = ceil(rank_factor * k) + oversample")
    ap.add_argument("--sketch-precision", default="fp64", choices=["fp64", "fp32", "bf16x3", "ozaki"], help="GEMM precision of the range finder (c4)")
    ap.add_argument("--ranks", default="learned", choices=["learned", "previous"],
                    help="target ranks of the randomized SVDs: the trained predictor or the previous step's kept counts")
    ap.add_argument("--leg-dim", type=int, default=512, help="c3 leg dimension")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32", "bf16x3", "ozaki"], help="GEMM precision of the fast mode / c3")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--device-cache-gib", type=float, default=None,
                    help="bound of the idle device blocks each runtime keeps for reuse (library default 4; 0 returns every freed block)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="bound of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true", help="skip the plain FP64 mode (c4)")
    ap.add_argument("--no-su", action="store_true", help="c4: skip the simple-update section")
    ap.add_argument("--no-fuse", action="store_true", help="c4: materialise every permuted operand (no gather GEMM); same results bit for bit")
    ap.add_argument("--no-previous", action="store_true", help="c4: skip the arms with the previous step's kept counts as target ranks and with freshly sketched predictor features")
    return ap.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


# ------------------------------------------------------------------------------------------ C4
SKINNY_MAX = 32  # an item with min(m, n) <= 32 and k <= 32 runs in the memory-bound skinny kernel


def launch_split(rt, xt):
    """Kernel time, flops, algorithmic bytes and count of the timed step's GEMM launches, split by
    the tag the tensor layer gives its contraction products against the randomized SVDs' sketch,
    Gram and factor products (key 0); contraction launches with at least one matrix-core item
    (key 1) apart from those whose every item is a skinny item (key 2, memory-bound: the fused
    gather products and the other site-tensor products)."""
    import ctypes as C

    L = xt.lib()
    ctx = rt.ctx()
    n = rt.gemm_launches()
    out = {0: [0.0, 0.0, 0, 0.0], 1: [0.0, 0.0, 0, 0.0], 2: [0.0, 0.0, 0, 0.0]}
    for i in range(n):
        t, tr, ni = C.c_double(), (C.c_int32 * 3)(), C.c_int64()
        L.xh_gemm_launch_log(ctx, i, C.byref(t), tr, None, 0, C.byref(ni))
        buf = (C.c_int32 * (3 * max(ni.value, 1)))()
        L.xh_gemm_launch_log(ctx, i, C.byref(t), tr, buf, ni.value, C.byref(ni))
        fl = by = 0.0
        skinny = True
        for k in range(ni.value):
            m, nn, kk = buf[3 * k], buf[3 * k + 1], buf[3 * k + 2]
            fl += 2.0 * m * nn * kk
            by += 8.0 * (m * kk + kk * nn + m * nn)
            skinny = skinny and (min(m, nn) <= SKINNY_MAX and kk <= SKINNY_MAX)
        o = out[0 if tr[2] != 1 else (2 if skinny else 1)]
        o[0] += t.value
        o[1] += fl
        o[2] += 1
        o[3] += by
    return out


def run_mode(mode, args, xt, configs, random_ipeps, ranks=None, fresh=False):
    """One CTMRG mode on a fresh runtime; returns timings, energies and GEMM statistics."""
    rt = xt.Runtime(0)
    rt.fuse_permute(not args.no_fuse)
    if args.device_cache_gib is not None:
        rt.device_cache(int(args.device_cache_gib * 2**30))
    if args.autotune and mode == "fast":
        xt.lib().xh_set_gemm_autotune(rt.ctx(), 1)
    A = [a.scale(SCALE) for a in random_ipeps(rt, configs.c4_virtual(), configs.c4_physical(), args.seed)]
    c = xt.CTM(rt, A, args.chi).init()
    t0 = time.perf_counter()
    c.grow_all()  # common warm-up: full FP64 SVDs, chi grows from D^2 = 64 to chi
    rt.synchronize()
    log(f"[bench] {mode}: warm-up grow_all {time.perf_counter() - t0:.1f} s")
    def configure(record):
        if mode == "fast":
            c.set_options(precision=args.precision, svd="randomized", oversample=args.oversample, power_iters=args.power_iters, record_spectra=record,
                          fresh_predictor_sketch=fresh)
            c.set_rsvd(oversample=args.oversample, rank_factor=args.rank_factor, power_iters=args.power_iters, sketch_precision=args.sketch_precision)
        else:
            c.set_options(precision="fp64", svd="full", record_spectra=record)

    # the first step after the warm-up records its kept spectra (compared between the modes)
    configure(True)
    c.clear_spectra()
    ranks = ranks or args.ranks
    if mode == "fast" and ranks == "learned":
        from xped_amd import rank_predictor

        params, width = rank_predictor.load_weights()
        c.set_predictor(params, width)
    nsteps = args.steps if mode == "fast" else args.fp64_steps
    nwarm = args.warmup if mode == "fast" else 0
    energies = []
    spectra = None
    for _ in range(nwarm):
        energies.append(c.step())
        if spectra is None:
            spectra = c.spectra()
            configure(False)
    rt.synchronize()
    svd0 = c.svd_ms()
    c0 = rt.counters()
    h0 = rt.host_stats()
    f0 = rt.fuse_permute(None)
    d0 = rt.device_cache()
    t0 = time.perf_counter()
    for k in range(nsteps):
        energies.append(c.step())
        log(f"[bench] {mode}: step {k} E = {energies[-1]:.15f}")
    rt.synchronize()
    wall = time.perf_counter() - t0
    c1 = rt.counters()
    h1 = rt.host_stats()
    f1 = rt.fuse_permute(None)
    d1 = rt.device_cache()
    if spectra is None:  # no warm-up step: the first timed step recorded (16 decompositions per step)
        spectra = c.spectra()[:16]
        configure(False)
    plan_ms = 1e3 * ((h1["permute_plan_s"] - h0["permute_plan_s"]) + (h1["gemm_plan_s"] - h0["gemm_plan_s"])) / nsteps
    out = {"ms_per_step": 1e3 * wall / nsteps, "host_plan_ms_per_step": plan_ms, "spectra": spectra,
           "plans_per_step": ((h1["n_permute_plans"] - h0["n_permute_plans"]) + (h1["n_gemm_plans"] - h0["n_gemm_plans"])) / nsteps,
           "host_breakdown_ms_per_step": [round(1e3 * (b - a) / nsteps, 1) for a, b in zip(h0["breakdown_s"], h1["breakdown_s"])],
           "symbolic_plan_hits_per_step": (h1["symbolic_hits"] - h0["symbolic_hits"]) / nsteps, "svd_ms_per_step": (c.svd_ms() - svd0) / nsteps, "energies": energies,
           "predictor_features_previous_sketch": list(c.predictor_stats()),
           "gemm_gflop_per_step": (c1["gemm_flops"] - c0["gemm_flops"]) / nsteps / 1e9,
           "permute_gb_per_step": (c1["permute_bytes"] - c0["permute_bytes"]) / nsteps / 1e9, "rsvd": c.rsvd_stats(), "host": rt.host_stats(),
           "fused_contractions_per_step": (f1 - f0) / nsteps,
           # device allocations per step served by the context's block cache / passed to the HIP runtime
           "device_allocs_cached_missed_per_step": [(d1["hits"] - d0["hits"]) / nsteps, (d1["misses"] - d0["misses"]) / nsteps]}
    if mode == "fast":
        # one more step with every GEMM and permute launch timed by HIP events on the context stream; marker
        # kernels of 1 and 2 workgroups delimit it in a kernel trace, and inside it markers of 3 and 4
        # workgroups bracket each contraction product (the tensor layer's tagged GEMM calls)
        rt.gemm_timing(True, reset=True)
        rt.permute_timing(True, reset=True)
        rt.marker(1)
        c.step()
        rt.marker(2)
        ms, fl, nlog = rt.gemm_timing(False)
        out["gemm_split"] = launch_split(rt, xt)
        out["gemm_launches"] = rt.gemm_launches()
        out["permute_ms"], out["permute_bytes"], out["permute_launches"] = rt.permute_timing(False, reset=True)
        out["gemm_kernel_ms"] = ms
        out["gemm_timed_flops"] = fl
        out["gemm_launches_blocks"] = nlog
        out["gemm_shapes"] = rt.gemm_log()
        if args.autotune:
            out["autotune"] = autotune_stats(rt)
    return out


def compare_spectra(fast, ref):
    """Kept singular values of the first compared step, decomposition by decomposition (same site
    and direction order in both modes): the largest difference relative to the decomposition's
    largest singular value, and the number of sectors whose kept count differs."""
    import numpy as np

    if not fast or not ref:
        return None
    worst, mism, n = 0.0, 0, 0
    for a, b in zip(fast, ref):
        if (a["site"], a["dir"]) != (b["site"], b["dir"]):
            return {"error": "decomposition order differs"}
        kb = {s["q"]: np.sort(s["sv"])[::-1][: s["kept"]] for s in b["sectors"] if s["kept"] > 0}
        ka = {s["q"]: np.sort(s["sv"])[::-1][: s["kept"]] for s in a["sectors"] if s["kept"] > 0}
        smax = max((v[0] for v in kb.values() if len(v)), default=1.0)
        for q in set(ka) | set(kb):
            va, vb = ka.get(q, np.zeros(0)), kb.get(q, np.zeros(0))
            if len(va) != len(vb):
                mism += 1
                m = min(len(va), len(vb))
                va, vb = va[:m], vb[:m]
            if len(va):
                worst = max(worst, float(np.abs(va - vb).max()) / smax)
        n += 1
    return {"decompositions": n, "max_abs_diff_over_sigma_max": worst, "sectors_with_other_kept_count": mism}


def cpu_gemm_baseline(shapes, seconds):
    """The reference's CPU hot loop of operator* (TensorBase.cpp:262-279) restated by the oracle:
    one FP64 BLAS GEMM per matched sector block, BLAS threads on all host cores, over the block
    shapes of the fast step's contraction GEMMs (random FP64 data of those shapes)."""
    import numpy as np

    from oracle import tensor as oten

    rng = np.random.default_rng(0)
    cache = {}
    work = []
    flops = 0.0
    for m, n, k in shapes:
        key = (int(m), int(n), int(k))
        if key not in cache:
            cache[key] = (np.asfortranarray(rng.uniform(-1, 1, (key[0], key[2]))), np.asfortranarray(rng.uniform(-1, 1, (key[2], key[1]))),
                          np.empty((key[0], key[1]), order="F"))
        work.append(cache[key])
        flops += 2.0 * key[0] * key[1] * key[2]
    done, t0, n = 0.0, time.perf_counter(), 0
    while time.perf_counter() - t0 < seconds and n < len(work) * 50:
        a, b, c = work[n % len(work)]
        oten.block_gemm(a, b, c)
        done += 2.0 * a.shape[0] * a.shape[1] * b.shape[1]
        n += 1
    el = time.perf_counter() - t0
    threads = os.environ.get("OMP_NUM_THREADS") or os.environ.get("OPENBLAS_NUM_THREADS") or str(os.cpu_count())
    return {"value": round(done / el / 1e12, 6), "unit": "TFLOP/s", "cores": int(threads), "kind": "port",
            "sample": f"{n} sector-block GEMMs, {done / 1e12:.3f} TFLOP in {el:.1f} s: the {len(work)} contraction blocks of one fast step "
                      f"({flops / 1e12:.3f} TFLOP) cycled in step order; oracle per-block FP64 matmul (numpy/OpenBLAS)"}


def run_c4(args):
    import xped_amd.tensor as xt
    from xped_amd import configs
    from xped_amd.ipeps import random_ipeps

    fast = run_mode("fast", args, xt, configs, random_ipeps)
    prev = None if (args.no_previous or args.ranks == "previous") else run_mode("fast", args, xt, configs, random_ipeps, ranks="previous")
    fresh = None if (args.no_previous or args.ranks == "previous") else run_mode("fast", args, xt, configs, random_ipeps, fresh=True)
    ref = None if args.no_fp64 else run_mode("fp64", args, xt, configs, random_ipeps)
    # value: contraction products only (operator* of the step's contractions); all grouped GEMM
    # launches of the step, the randomized SVDs' included, are reported beside it
    mf_ms, mf_fl, mf_n, _ = fast["gemm_split"][1]
    sn_ms, sn_fl, sn_n, sn_by = fast["gemm_split"][2]
    sk_ms, sk_fl, sk_n, _ = fast["gemm_split"][0]
    con_ms, con_fl, con_n = mf_ms + sn_ms, mf_fl + sn_fl, mf_n + sn_n
    value = con_fl / (con_ms * 1e-3) / 1e12 if con_ms > 0 else 0.0
    mf_tf = mf_fl / (mf_ms * 1e-3) / 1e12 if mf_ms > 0 else 0.0
    all_tf = fast["gemm_timed_flops"] / (fast["gemm_kernel_ms"] * 1e-3) / 1e12 if fast["gemm_kernel_ms"] > 0 else 0.0
    peak = PEAKS[args.precision]
    res = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(fast["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPES[args.precision],
        "data": f"synthetic (site tensors N(0,1)*{SCALE:g}/sqrt(plain size), std::mt19937(seed + site))",
        "config": {"workload": f"U(1) iPEPS CTMRG step, 2x2 cell, D=8, chi={args.chi}, Heisenberg", "chi": args.chi, "D": 8,
                   "svd": f"randomized, l = {args.rank_factor:g} k + {args.oversample}, k from the {'learned predictor' if args.ranks == 'learned' else 'previous step'}, "
                          f"{args.power_iters} power iterations, FP64 fallback",
                   "gemm_precision": args.precision, "seed": args.seed},
        "fast": {"ms_per_step": round(fast["ms_per_step"], 3), "svd_ms_per_step": round(fast["svd_ms_per_step"], 3),
                 "gemm_gflop_per_step": round(fast["gemm_gflop_per_step"], 3), "permute_gb_per_step": round(fast["permute_gb_per_step"], 3),
                 "rsvd_calls_fallbacks_ratio": list(fast["rsvd"]), "energies": fast["energies"],
                 "fused_contractions_per_step": fast["fused_contractions_per_step"],
                 "device_allocs_cached_missed_per_step": fast["device_allocs_cached_missed_per_step"],
                 "predictor_features_previous_sketch": fast["predictor_features_previous_sketch"],
                 **({"autotune_entries_runs_hits_chose64": fast["autotune"]} if "autotune" in fast else {}),
                 "host_plan_ms_per_step": round(fast["host_plan_ms_per_step"], 1), "plans_per_step": fast["plans_per_step"],
                 "host_breakdown_ms_per_step": fast["host_breakdown_ms_per_step"], "symbolic_plan_hits_per_step": fast["symbolic_plan_hits_per_step"]},
        # the dominant kernel: the matrix-core GEMM of the contraction launches with tiled items
        "roofline": {"bound": "mfma", "achieved": round(mf_tf * (OZ_PAIRS if args.precision == "ozaki" else 1), 4), "peak": peak,
                     "unit": "TOP/s (int8)" if args.precision == "ozaki" else "TFLOP/s",
                     "frac": round(mf_tf * (OZ_PAIRS if args.precision == "ozaki" else 1) / peak, 5),
                     "traffic": measured_traffic("k_gemm_f64_ctm_step_counters.md") if args.precision == "fp64" and args.chi == 256 else None,
                     "kernel": ("contraction GEMM launches with matrix-core items: k_gemm_f64 + k_gemm_f64_128 (+ k_splitk_reduce, and the skinny items of the same launches)"
                                if args.precision == "fp64" else "k_oz_split + k_gemm_oz" if args.precision == "ozaki" else f"k_gemm_{args.precision}"),
                     "launches_per_step": mf_n, "flops_per_step": mf_fl, "bytes_per_step": fast["gemm_split"][1][3], "kernel_ms_per_step": round(mf_ms, 3)},
        # contraction launches whose every item is skinny (memory-bound), the gather GEMMs among them
        "skinny_contractions": {"bound": "hbm", "kernel": "k_gemm_skinny + k_gemm_skinny_gather",
                                "achieved": round(sn_by / (sn_ms * 1e-3) / 1e9, 1) if sn_ms > 0 else 0.0, "peak": 8000.0, "unit": "GB/s",
                                "frac": round(sn_by / (sn_ms * 1e-3) / 1e9 / 8000.0, 4) if sn_ms > 0 else 0.0,
                                "bytes_per_step": sn_by, "flops_per_step": sn_fl, "launches_per_step": sn_n, "kernel_ms_per_step": round(sn_ms, 3)},
        "contractions_all": {"achieved": round(value, 4), "unit": "TFLOP/s", "launches_per_step": con_n, "flops_per_step": con_fl,
                             "kernel_ms_per_step": round(con_ms, 3)},
        "all_gemm_launches": {"achieved": round(all_tf, 4), "unit": "TFLOP/s", "frac": round(all_tf / peak, 5), "launches_per_step": fast["gemm_launches"],
                              "flops_per_step": fast["gemm_timed_flops"], "kernel_ms_per_step": round(fast["gemm_kernel_ms"], 3),
                              "randomized_svd_products": {"achieved": round(sk_fl / (sk_ms * 1e-3) / 1e12, 4) if sk_ms > 0 else 0.0, "launches": sk_n,
                                                          "flops": sk_fl, "kernel_ms": round(sk_ms, 3)}},
        "permute_roofline": {"bound": "hbm", "kernel": "k_permute_gather + k_permute_transpose",
                             "achieved": round(fast["permute_bytes"] / (fast["permute_ms"] * 1e-3) / 1e9, 1) if fast["permute_ms"] > 0 else 0.0,
                             "peak": 8000.0, "unit": "GB/s",
                             "frac": round(fast["permute_bytes"] / (fast["permute_ms"] * 1e-3) / 1e9 / 8000.0, 4) if fast["permute_ms"] > 0 else 0.0,
                             "traffic": measured_traffic("k_permute_ctm_step_counters.md") if args.chi == 256 else None,
                             "bytes_per_step": fast["permute_bytes"], "launches_per_step": fast["permute_launches"],
                             "kernel_ms_per_step": round(fast["permute_ms"], 3)},
    }
    if fresh is not None:
        res["learned_fresh_sketch"] = {"ms_per_step": round(fresh["ms_per_step"], 3), "svd_ms_per_step": round(fresh["svd_ms_per_step"], 3),
                                       "rsvd_calls_fallbacks_ratio": list(fresh["rsvd"]), "energies": fresh["energies"]}
    if prev is not None:
        res["ranks_previous"] = {"ms_per_step": round(prev["ms_per_step"], 3), "svd_ms_per_step": round(prev["svd_ms_per_step"], 3),
                                 "rsvd_calls_fallbacks_ratio": list(prev["rsvd"]), "energies": prev["energies"]}
    if ref is not None:
        # step by step: the fast mode's warm-up and timed steps against as many FP64 steps
        nb = min(len(ref["energies"]), len(fast["energies"]))
        errs = [abs(fast["energies"][i] - ref["energies"][i]) / abs(ref["energies"][i]) for i in range(nb)]
        res["energy_rel_err_vs_fp64_per_step"] = errs
        if fresh is not None:
            nbf = min(len(ref["energies"]), len(fresh["energies"]))
            res["learned_fresh_sketch"]["energy_rel_err_vs_fp64_per_step"] = [abs(fresh["energies"][i] - ref["energies"][i]) / abs(ref["energies"][i])
                                                                             for i in range(nbf)]
        if prev is not None:
            nbp = min(len(ref["energies"]), len(prev["energies"]))
            res["ranks_previous"]["energy_rel_err_vs_fp64_per_step"] = [abs(prev["energies"][i] - ref["energies"][i]) / abs(ref["energies"][i]) for i in range(nbp)]
        res["fp64"] = {"ms_per_step": round(ref["ms_per_step"], 3), "svd_ms_per_step": round(ref["svd_ms_per_step"], 3), "energies": ref["energies"],
                       "host_plan_ms_per_step": round(ref["host_plan_ms_per_step"], 1), "plans_per_step": ref["plans_per_step"]}
        res["speedup_vs_fp64"] = round(ref["ms_per_step"] / fast["ms_per_step"], 3)
        res["energy_rel_err_vs_fp64"] = max(errs) if errs else None
        res["spectra_vs_fp64"] = compare_spectra(fast["spectra"], ref["spectra"])
    if not args.no_su:
        res["simple_update"] = run_simple_update(args, xt, configs, random_ipeps)
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_gemm_baseline(fast["gemm_shapes"], args.cpu_seconds)
        res["cpu_baseline"]["oracle_ctm_step_s"] = ORACLE_STEP
    return res


def run_simple_update(args, xt, configs, random_ipeps, steps=3):
    """Simple-update steps (TimePropagator::t_step, 8 bond updates) of the Heisenberg model on the
    C4 sites (D = 8, 2x2 cell), dt = 0.05, after one untimed step that builds the plans."""
    rt = xt.Runtime(0)
    A = random_ipeps(rt, configs.c4_virtual(), configs.c4_physical(), args.seed)
    H = xt.heisenberg_bond(rt, A[0].leg(4))
    tp = xt.TimePropagator(rt, [H] * 4, [H] * 4, 0.05)
    psi = xt.WeightedPEPS(rt, A, 8)
    tp.t_step(psi)
    rt.synchronize()
    c0 = rt.counters()
    t0 = time.perf_counter()
    for _ in range(steps):
        tp.t_step(psi)
    rt.synchronize()
    wall = time.perf_counter() - t0
    c1 = rt.counters()
    log(f"[bench] simple update: {1e3 * wall / steps:.1f} ms per step")
    return {"workload": "simple update t_step (8 bond updates), Heisenberg, 2x2 cell, D=8, dt=0.05", "steps": steps,
            "ms_per_step": round(1e3 * wall / steps, 3), "ms_per_bond_update": round(1e3 * wall / steps / 8, 3),
            "gemm_launches_per_step": (c1["n_gemm"] - c0["n_gemm"]) / steps, "permute_launches_per_step": (c1["n_permute"] - c0["n_permute"]) / steps}


# ------------------------------------------------------------------------------------------ C3
def build_c3(rt, leg_dim, xt, configs):
    def basis(entries):
        b = xt.Qbasis(xt.FU1xU1)
        for q, d in entries:
            b.push_back(q, d)
        return b

    legs = [basis(configs.c3_leg(s, leg_dim)) for s in range(6)]
    A = xt.Tensor(rt, legs[0:2], legs[2:4]).setRandom(1)
    B = xt.Tensor(rt, legs[2:4], legs[4:6]).setRandom(2)
    return A, B


def cpu_baseline_c3(leg_dim, seconds, configs):
    """operator* (TensorBase.cpp:262-279) restated by the oracle: one FP64 BLAS GEMM per matched
    sector, serial over sectors; block structures are built once outside the timed loop."""
    import numpy as np

    from oracle import tensor as oten
    from oracle.structure import Qbasis as OQ
    from oracle.symmetry import FU1, U1, Sym

    s = Sym(FU1, U1)

    def basis(entries):
        b = OQ(s, 1)
        for q, d in entries:
            b.push_back(q, d)
        return b

    legs = [basis(configs.c3_leg(i, leg_dim)) for i in range(6)]
    A = oten.Tensor(s, legs[0:2], legs[2:4]).set_random(1)
    B = oten.Tensor(s, legs[2:4], legs[4:6]).set_random(2)
    pairs = []
    flops = 0.0
    for i, q in enumerate(A.sector):
        j = B.find(q)
        if j >= 0:
            a = np.asfortranarray(A.block(i))
            b = np.asfortranarray(B.block(j))
            pairs.append((a, b, np.empty((a.shape[0], b.shape[1]), order="F")))
            flops += 2.0 * a.shape[0] * a.shape[1] * b.shape[1]

    def step():
        for a, b, c in pairs:
            oten.block_gemm(a, b, c)

    step()
    n, t0 = 0, time.perf_counter()
    while True:
        step()
        n += 1
        el = time.perf_counter() - t0
        if el >= seconds or n >= 100000:
            break
    threads = os.environ.get("OMP_NUM_THREADS") or os.environ.get("OPENBLAS_NUM_THREADS") or str(os.cpu_count())
    return {"value": round(flops * n / el / 1e12, 6), "unit": "TFLOP/s", "cores": int(threads), "kind": "port",
            "sample": f"{n} passes of the {len(pairs)} matched-sector FP64 GEMMs of the workload in {el:.1f} s "
                      f"(oracle: numpy/OpenBLAS matmul per sector, structures prebuilt)"}


def run_c3(args):
    import xped_amd.tensor as xt
    from xped_amd import _lib, configs

    L = _lib.lib()
    rt = xt.Runtime(0)
    if args.autotune:
        L.xh_set_gemm_autotune(rt.ctx(), 1)
    A, B = build_c3(rt, args.leg_dim, xt, configs)
    prec = args.precision
    a1, a2 = [-1, -2, 1, 2], [1, 2, -3, -4]
    steps = max(args.steps, 10)
    for _ in range(max(args.warmup, 2)):
        Cm = A.contract(B, a1, a2, 2, precision=prec)
    rt.synchronize()
    c0 = rt.counters()
    ctx = rt.ctx()
    L.xh_event_record(ctx, 0)
    t0 = time.perf_counter()
    for _ in range(steps):
        Cm = A.contract(B, a1, a2, 2, precision=prec)
    L.xh_event_record(ctx, 1)
    rt.synchronize()
    wall = time.perf_counter() - t0
    ms = C.c_float()
    L.xh_event_elapsed_ms(ctx, 0, 1, C.byref(ms))
    c1 = rt.counters()
    flops_step = (c1["gemm_flops"] - c0["gemm_flops"]) / steps
    step_ms = ms.value / steps
    value = flops_step / (step_ms * 1e-3) / 1e12
    # the product's error against the FP64 matrix cores on the same blocks (mixed modes)
    vs_fp64 = None
    if prec != "fp64":
        ref = A.contract(B, a1, a2, 2, precision="fp64")
        vs_fp64 = {"max_abs_diff_over_max_abs": (Cm - ref).maxNorm() / ref.maxNorm()}
        del ref
    graph = None
    if args.graph and prec == "ozaki":
        log("[bench] c3: no graph replay in the Ozaki mode (it allocates its slices per call)")
    elif args.graph:
        # the same product replayed from a HIP graph captured once (mul_into on the output of the
        # last eager step: no allocation under capture); timed like the eager steps
        out = Cm
        rt.graph_begin()
        A.mul_into(B, out, precision=prec)
        g = rt.graph_end()
        for _ in range(2):
            g.launch()
        rt.synchronize()
        L.xh_event_record(ctx, 0)
        for _ in range(steps):
            g.launch()
        L.xh_event_record(ctx, 1)
        rt.synchronize()
        gms = C.c_float()
        L.xh_event_elapsed_ms(ctx, 0, 1, C.byref(gms))
        graph = {"ms_per_step": round(gms.value / steps, 5), "tflops": round(flops_step / (gms.value / steps * 1e-3) / 1e12, 4)}
        g.close()
    # dominant kernel: the grouped GEMM launch alone, timed with HIP events on the context stream
    rt.gemm_timing(True, reset=True)
    for _ in range(20):
        Cm = A.mul(B, precision=prec)
    gemm_ms, gemm_fl, nblocks = rt.gemm_timing(False)
    achieved = gemm_fl / (gemm_ms * 1e-3) / 1e12
    peak = PEAKS[prec]
    if prec == "ozaki":
        fp64_equiv = achieved
        achieved = OZ_PAIRS * achieved  # int8 TOP/s of the slice products
    res = {
        "metric": METRIC, "value": round(value, 4), "unit": "TFLOP/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
        "ms_per_step": round(step_ms, 5), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": DTYPES[prec], "data": "synthetic (seeded mt19937 uniform[-1,1) blocks)",
        "config": {"workload": f"U(1)xU(1) fermionic grouped contraction A(l1,l2|l3,l4).B(l3,l4|l5,l6), leg dim {args.leg_dim}",
                   "matched_blocks": nblocks // 20, "gflop_per_step": round(flops_step / 1e9, 4), "precision": prec},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 4), "peak": peak, "unit": "TOP/s (int8)" if prec == "ozaki" else "TFLOP/s",
                     "frac": round(achieved / peak, 5), "traffic": None,
                     "kernel": {"fp64": "k_gemm_f64 + k_gemm_f64_128", "ozaki": "k_oz_split + k_gemm_oz"}.get(prec, f"k_gemm_{prec}"),
                     "kernel_ms": round(gemm_ms / 20, 5)},
        "wall_s_timed": round(wall, 4),
    }
    if prec == "ozaki":
        res["roofline"]["fp64_equivalent_tflops"] = round(fp64_equiv, 4)
    if vs_fp64 is not None:
        res["error_vs_fp64"] = vs_fp64
    if graph is not None:
        res["graph_replay"] = graph
    if args.autotune:
        res["autotune_entries_runs_hits_chose64"] = autotune_stats(rt)
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline_c3(args.leg_dim, args.cpu_seconds, configs)
    return res


def run_copy(args):
    """HBM copy calibration: y = 2 x over 1 GiB with k_axpby (b = 0 does not read y)."""
    import ctypes as C

    import xped_amd.tensor as xt
    from xped_amd._lib import lib

    rt = xt.Runtime(0)
    L, ctx = lib(), rt.ctx()
    n = 1 << 27
    x, y = C.c_void_p(), C.c_void_p()
    for d in (x, y):
        if L.xh_alloc(ctx, 8 * n, C.byref(d)) != 0:
            raise RuntimeError("allocation failed")
    L.xh_fill(ctx, x, n, 1.0)
    L.xh_axpby(ctx, y, x, n, 2.0, 0.0)  # warm-up
    rt.synchronize()
    steps = 3
    t0 = time.perf_counter()
    for _ in range(steps):
        L.xh_axpby(ctx, y, x, n, 2.0, 0.0)
    rt.synchronize()
    el = time.perf_counter() - t0
    gbs = 2 * 8 * n * steps / el / 1e9
    L.xh_free(ctx, x)
    L.xh_free(ctx, y)
    return {"metric": "HBM copy bandwidth (k_axpby, b = 0)", "value": round(gbs, 1), "unit": "GB/s", "n_gpus": 1, "steps": steps, "warmup": 1,
            "ms_per_step": round(1e3 * el / steps, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic (constant fill)", "config": {"workload": "copy of 2^27 doubles (1 GiB read + 1 GiB written per launch)"},
            "bytes_per_launch": 2 * 8 * n}


def autotune_stats(rt):
    """(cached signatures, tuning runs, cache hits, signatures that chose 64 x 64 tiles only)."""
    import xped_amd.tensor as xt

    v = [C.c_int64() for _ in range(4)]
    xt.lib().xh_gemm_autotune_stats(rt.ctx(), *[C.byref(x) for x in v])
    return [x.value for x in v]


def measured_traffic(name):
    """HBM bytes per launch from a committed rocprofv3 counter summary (a line
    `traffic_bytes_per_launch: <value>` in profiles/<name>), or None."""
    path = ROOT / "profiles" / name
    try:
        for line in path.read_text().splitlines():
            if line.strip().startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1].strip())
    except OSError:
        return None
    return None


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU: --gpus must be 1 (the multi-GPU sharding is a design note only)"}))
        sys.exit(2)
    res = {"c4": run_c4, "c3": run_c3, "copy": run_copy}[args.workload](args)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
