"""Build recipe for the in-tree shared library ``xped_amd/lib/libxped_hip.so``.

HIP sources are compiled by hipcc for gfx950 only; host-only C++ sources (the structure engine and
the C ABI) by g++.  Objects go to ``xped_amd/_build`` and the library into ``xped_amd/lib``.
Run ``python -m xped_amd.build`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIBDIR = PKG / "lib"
LIB = LIBDIR / "libxped_hip.so"

ARCH = "gfx950"
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
CXX = shutil.which("g++") or "g++"

HIP_FLAGS = [
    "-O3", "-std=c++20", "-fPIC", f"--offload-arch={ARCH}",
    # keep the compiler from emitting packed FP32 arithmetic (v_pk_*_f32)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]
CXX_FLAGS = ["-O2", "-std=c++20", "-fPIC", "-pthread", "-Wall", "-Wno-unused-result", "-Wno-sign-compare"]


def sources():
    hip = sorted((CSRC / "hip").glob("*.hip"))
    cpp = sorted((CSRC / "core").glob("*.cpp")) + sorted(CSRC.glob("*.cpp"))
    return hip, cpp


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def build(verbose: bool = False, jobs: int | None = None) -> Path:
    hip, cpp = sources()
    BUILD.mkdir(exist_ok=True)
    LIBDIR.mkdir(exist_ok=True)
    inc = ["-I", str(ROOT / "include")]
    steps = []
    for s in hip:
        o = BUILD / (s.stem + ".hip.o")
        steps.append(([HIPCC, *HIP_FLAGS, *inc, "-c", str(s), "-o", str(o)], o))
    for s in cpp:
        o = BUILD / (s.parent.name + "_" + s.stem + ".o")
        steps.append(([CXX, *CXX_FLAGS, *inc, "-c", str(s), "-o", str(o)], o))
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(_run, cmd) for cmd, _ in steps]
        for f, (cmd, _) in zip(futs, steps):
            out = f.result()
            if verbose:
                print(" ".join(cmd))
                if out.strip():
                    print(out)
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *[str(o) for _, o in steps], "-o", str(tmp),
          "-lpthread"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    p = build(verbose="-v" in sys.argv)
    print(p)
