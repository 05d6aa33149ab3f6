"""Builds tests/cpp/adapter_test.cpp against include/ and the in-tree library (used by the tests)."""
import json
import subprocess
import tempfile
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
LIBDIR = ROOT / "xped_amd" / "lib"


def build_and_run(mode: str) -> dict:
    out = Path(tempfile.mkdtemp(prefix="xped_adapter_")) / "adapter_test"
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", str(ROOT / "include"), str(ROOT / "tests" / "cpp" / "adapter_test.cpp"), "-o",
                    str(out), f"-L{LIBDIR}", "-lxped_hip", f"-Wl,-rpath,{LIBDIR}"], check=True, capture_output=True, text=True)
    r = subprocess.run([str(out), mode], check=True, capture_output=True, text=True)
    return json.loads(r.stdout)
