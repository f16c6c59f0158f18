"""Build libcantera_amd.so (HIP, gfx950) in-tree with hipcc.

Usage: python -m cantera_amd.build [--jobs N]
"""
from __future__ import annotations

import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libcantera_amd.so")
OBJDIR = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["mech.cpp", "api.cpp", "rates.hip", "jacobian.hip", "order.hip", "bdf.hip", "bdf_r16.hip", "bdf_r32.hip", "bdf_lr56.hip", "bdf_lr64.hip",
           "bdf_gr56.hip", "bdf_gr64.hip",
           "bdf_l2.hip", "bdf_mole.hip", "sparse_lu.cpp", "surrogate.hip", "surrogate_reactors.hip", "equil.hip"]

COMMON = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    # packed-FP32 instructions (v_pk_*_f32) off: some of their forms have given
    # wrong results without an error on MI355X, so none are generated
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-variable", "-Wno-unused-but-set-variable",
]


def _compile(src: str) -> str:
    obj = os.path.join(OBJDIR, os.path.basename(src) + ".o")
    cmd = [HIPCC, *COMMON, "-c", os.path.join(CSRC, src), "-o", obj]
    if src.endswith(".hip"):
        cmd[1:1] = ["-x", "hip"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    if r.stderr.strip():
        lines = [l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and "warning: feature flag" not in l]
        if lines:
            print("\n".join(lines), file=sys.stderr)
    return obj


def build(jobs: int = 8) -> str:
    os.makedirs(OBJDIR, exist_ok=True)
    with ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    return OUT


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=4)
    a = ap.parse_args(argv)
    print(build(a.jobs))


if __name__ == "__main__":
    main()
