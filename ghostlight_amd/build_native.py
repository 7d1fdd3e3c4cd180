"""Build the gfx950 HIP library in-tree: ghostlight_amd/_lib/libghostlight_amd.so.

Every HIP translation unit is compiled with -ffp-contract=off and IEEE fp32
division / square root (the tracer's bit-exact contract with its CPU oracle)
and with packed-FP32 instruction selection disabled.  Host-only .cpp units
are compiled by the system C++ compiler.  Usage:
    python -m ghostlight_amd.build_native [--jobs N]
"""
from __future__ import annotations

import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_lib", "obj")
OUT = os.path.join(HERE, "_lib", "libghostlight_amd.so")

ARCH = "gfx950"
HIP_FLAGS = [
    "-x", "hip", f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-command-line-argument",
]


# Host-only translation units, built by the system C++ compiler (see the
# header of each file for why it is not a HIP unit).
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall"]


def sources():
    return sorted(f for f in os.listdir(CSRC) if f.endswith(".hip") or f.endswith(".cpp"))


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.sep not in c or os.path.exists(c)):
            return c
    return "hipcc"


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    if src.endswith(".cpp"):
        cmd = [os.environ.get("CXX", "g++")] + CXX_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
    else:
        cmd = [hipcc()] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-4000:]}")
    return obj


def build(jobs: int = 8) -> str:
    os.makedirs(OBJ, exist_ok=True)
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(_compile, sources()))
    tmp = OUT + ".tmp"
    # Host runtime helpers (complex float division and the like) come from
    # libgcc, as for code built by the system compiler.
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "--rtlib=libgcc", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    a = ap.parse_args()
    print(build(a.jobs))
    sys.exit(0)
